// Device-wide scan and reduce for gfx950.
//
// Reference: vk::scan (inc/mgpu/vk/scan.hxx:10-115) is an exclusive,
// in-place, three-phase scan (upsweep -> recursive scan of tile partials ->
// downsweep) that moves 12 bytes per 4-byte element.  Here the scan is a
// single pass with decoupled look-back (Merrill & Garland, NVR-2016-002): each
// tile publishes its aggregate, then walks back over its predecessors' words
// 64 at a time (one wave) until it meets an inclusive prefix, so every element
// is read once and written once (8 bytes per element).
//
// Summation order inside a tile is hierarchical as in the reference
// (thread-serial over 4 consecutive elements, cta_scan.hxx:150; a wave scan,
// :100-104; the wave totals in order, :107-130; then the carry-in, :157-171).
// A tile is NT x VT elements in rows of 4 * NT; each wave scans an aligned
// block of 256 consecutive elements of a row in the element type, and every
// sum above such a block (earlier blocks of the tile, the carry between
// tiles) is taken in the carry type: float64 for float32 data, so a float32
// result has at most 10 float32 additions on its path (3 thread-serial,
// 6 wave-scan levels, the wave prefix) and does not depend on the tile shape.
// For integers all arithmetic is modulo 2^32.  Tiles are large (20480
// elements by default) because each tile's look-back costs a few memory
// round trips: fewer tiles per byte is what moves the scan toward the copy
// bandwidth.
#pragma once
#include <functional>

#include "meta.hxx"

namespace mgpu {
namespace hip {

template <typename T> struct scan_traits;
template <> struct scan_traits<int32_t> { typedef uint32_t acc_t; typedef uint32_t carry_t; };
template <> struct scan_traits<uint32_t> { typedef uint32_t acc_t; typedef uint32_t carry_t; };
template <> struct scan_traits<float> { typedef float acc_t; typedef double carry_t; };

// Look-back status word: the payload and a 2-bit flag (0 = not ready,
// 1 = tile aggregate, 2 = inclusive prefix) in one 64-bit word.
enum : uint32_t { status_invalid = 0, status_aggregate = 1, status_inclusive = 2 };

__device__ inline uint64_t pack_status(uint32_t v, uint32_t flag) {
  return ((uint64_t)v << 32) | flag;
}
__device__ inline uint64_t pack_status(double v, uint32_t flag) {
  return ((uint64_t)__double_as_longlong(v) & ~3ull) | flag;
}
template <typename carry_t> __device__ inline carry_t unpack_status(uint64_t w);
template <> __device__ inline uint32_t unpack_status<uint32_t>(uint64_t w) {
  return (uint32_t)(w >> 32);
}
template <> __device__ inline double unpack_status<double>(uint64_t w) {
  return __longlong_as_double((long long)(w & ~3ull));
}

struct identity_t {
  template <typename T>
  __host__ __device__ T operator()(T x) const { return x; }
};

// Sum of xform(input[first, first + n)) computed by one wave, in the carry
// type (used when a predecessor tile has not published its aggregate).
template <typename carry_t, typename acc_t, typename in_t, typename T, typename xform_t>
__device__ inline carry_t wave_tile_sum(const in_t* input, int64_t first, int n, xform_t xform) {
  carry_t sum = 0;
  for (int e = lane_id(); e < n; e += warp_size) sum += (carry_t)(acc_t)(T)xform(input[first + e]);
  return wave_reduce_all(sum);
}

// Exclusive prefix of tile `tile` from the words of its predecessors.  Called
// by one full wave; every lane returns the same value.  Words are read 64 at a
// time; a window with a not-yet-published word is re-polled up to
// lb.spin_limit times, then the first missing tile's aggregate is computed
// from its input.  For an in-place scan the owner of that tile may be writing
// its output concurrently, but only after its aggregate word is visible, so
// the word is read again after the data: if it is visible now, the computed
// sum is discarded.
template <typename carry_t, typename acc_t, int NV, typename in_t, typename T, typename xform_t>
__device__ inline carry_t scan_lookback(const uint64_t* status, int64_t tile, carry_t init,
                                        const in_t* input, xform_t xform,
                                        lookback_config_t lb) {
  const int lane = lane_id();
  carry_t excl = 0;
  int64_t j = tile - 1;
  uint32_t spins = 0;
  while (true) {
    const int64_t idx = j - lane;
    uint64_t w;
    if (idx >= 0) w = atomic_load_agent(status + idx);
    else w = pack_status(idx == -1 ? init : carry_t(0), status_inclusive);
    const uint32_t flag = (uint32_t)(w & 3);
    const uint64_t inc_mask = __ballot(flag == status_inclusive);
    const uint64_t bad_mask = __ballot(flag == status_invalid);
    const uint64_t lowest = inc_mask & (~inc_mask + 1);
    const uint64_t range = inc_mask ? (lowest | (lowest - 1)) : ~0ull;
    if (bad_mask & range) {
      if (++spins <= lb.spin_limit) {
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
      const int kb = __builtin_ctzll(bad_mask & range);
      const int64_t jj = j - kb;
      carry_t agg = wave_tile_sum<carry_t, acc_t, in_t, T>(input, jj * NV, NV, xform);
      acquire_agent();
      const uint64_t again = atomic_load_agent(status + jj);
      carry_t v = lane < kb ? unpack_status<carry_t>(w) : carry_t(0);
      excl += wave_reduce_all(v);
      spins = 0;
      if ((again & 3) != status_invalid) {
        j = jj;  // published meanwhile: continue from its word
        continue;
      }
      excl += agg;
      if (lane == 0 && lb.fallbacks) atomic_add_agent(lb.fallbacks, 1u);
      j = jj - 1;
      continue;
    }
    carry_t v = ((range >> lane) & 1) ? unpack_status<carry_t>(w) : carry_t(0);
    excl += wave_reduce_all(v);
    if (inc_mask) break;
    j -= warp_size;
  }
  return excl;
}

__device__ inline uint32_t carry_combine(uint32_t carry, uint32_t local) { return carry + local; }
__device__ inline float carry_combine(double carry, float local) {
  return (float)(carry + (double)local);
}

// Load tile `tile` (NT*VT elements) as acc_t: 128-bit row-strided loads for a
// full, aligned tile, guarded scalar loads otherwise.
template <int NT, int VT, typename acc_t, typename in_t, typename T, typename xform_t>
__device__ inline void scan_load_tile(const in_t* input, int64_t count, int64_t tile, bool vec_ok,
                                      xform_t xform, acc_t (&x)[VT]) {
  constexpr int NV = NT * VT;
  constexpr int NR = VT / 4;
  const int tid = threadIdx.x;
  const int64_t base = tile * NV;
  if (vec_ok && count - base >= NV) {
    typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      u32x4 v = __builtin_nontemporal_load((const u32x4*)(input + base) + (r * NT + tid));
      in_t e[4];
      __builtin_memcpy(e, &v, 16);
#pragma unroll
      for (int q = 0; q < 4; ++q) x[4 * r + q] = (acc_t)(T)xform(e[q]);
    }
  } else {
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int64_t idx = base + (int64_t)(r * NT + tid) * 4 + q;
        x[4 * r + q] = idx < count ? (acc_t)(T)xform(input[idx]) : acc_t(0);
      }
  }
}

// Single-pass scan: block b scans tile b.  Wave 0 publishes the tile's
// aggregate and looks back over its predecessors.
template <int NT, int VT, bool inclusive, typename in_t, typename T, typename xform_t>
__global__ __launch_bounds__(NT) void kernel_scan_lookback(
    const in_t* input, T* output, int64_t count, int64_t num_tiles, T init, uint64_t* status,
    lookback_config_t lb, int vec_ok, xform_t xform) {
  typedef scan_traits<T> traits;
  typedef typename traits::acc_t acc_t;
  typedef typename traits::carry_t carry_t;
  constexpr int NW = NT / warp_size;
  constexpr int NV = NT * VT;
  constexpr int NR = VT / 4;
  static_assert(VT % 4 == 0, "VT must be a multiple of 4");
  static_assert(sizeof(T) == 4, "scan works on 32-bit element types");

  __shared__ acc_t s_wave[NR][NW];
  __shared__ carry_t s_carry;

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int warp = tid / warp_size;
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * NV;

  acc_t x[VT];
  scan_load_tile<NT, VT, acc_t, in_t, T>(input, count, tile, vec_ok, xform, x);

  // Thread-serial inclusive scan of each group of 4, then a wave scan of the
  // group totals for every row.
  acc_t w_excl[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    x[4 * r + 1] += x[4 * r + 0];
    x[4 * r + 2] += x[4 * r + 1];
    x[4 * r + 3] += x[4 * r + 2];
    acc_t incl = wave_scan_incl(x[4 * r + 3]);
    if (lane == warp_size - 1) s_wave[r][warp] = incl;
    w_excl[r] = wave_shift_up1(incl, acc_t(0));
  }
  __syncthreads();

  // Tile aggregate in the carry type (float64 for float32 data), summed by
  // wave 0 over the NR x NW wave totals.
  carry_t aggregate = 0;
  if (warp == 0) {
    const acc_t* flat = &s_wave[0][0];
    for (int i = lane; i < NR * NW; i += warp_size) aggregate += (carry_t)flat[i];
    aggregate = wave_reduce_all(aggregate);
  }

  if (warp == 0) {
    carry_t carry;
    if (tile == 0) {
      carry = (carry_t)(acc_t)init;
    } else {
      if (lane == 0) {
        atomic_store_agent(status + tile, pack_status(aggregate, status_aggregate));
        // The aggregate must be visible before this tile's output (possibly
        // over its own input) can be.
        status_store_complete();
      }
      carry = scan_lookback<carry_t, acc_t, NV, in_t, T>(status, tile, (carry_t)(acc_t)init,
                                                         input, xform, lb);
    }
    if (lane == 0) {
      atomic_store_agent(status + tile,
                         pack_status((carry_t)(carry + aggregate), status_inclusive));
      s_carry = carry;
    }
  }
  __syncthreads();
  const carry_t carry = s_carry;

  // Everything above one wave row (the carry, earlier rows, earlier waves of
  // this row) is summed in the carry type; inside a row the sums are acc_t:
  // at most 3 thread-serial additions, 6 wave-scan levels and the wave's
  // exclusive prefix.
  T o[VT];
  carry_t run = carry;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    carry_t pre = run, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const carry_t v = (carry_t)s_wave[r][w];
      if (w < warp) pre += v;
      tot += v;
    }
    run += tot;
    const acc_t p = w_excl[r];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      acc_t local = inclusive ? (acc_t)(p + x[4 * r + q])
                              : (q == 0 ? p : (acc_t)(p + x[4 * r + q - 1]));
      o[4 * r + q] = (T)carry_combine(pre, local);
    }
  }
  if (vec_ok && count - base >= NV) {
    typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      u32x4 v;
      __builtin_memcpy(&v, &o[4 * r], 16);
      __builtin_nontemporal_store(v, (u32x4*)(output + base) + (r * NT + tid));
    }
  } else {
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int64_t idx = base + (int64_t)(r * NT + tid) * 4 + q;
        if (idx < count) output[idx] = o[4 * r + q];
      }
  }
}

// The scan kernel of transform_scan, for a caller that has already zeroed the
// status words (one 64-bit word per NT * VT tile; transform_scan keeps them at
// the start of its aux).  Arguments are not checked again.
template <int NT, int VT, typename in_t, typename T, typename xform_t>
hipError_t scan_launch_cleared(hipStream_t stream, const in_t* input, T* output, int64_t count,
                               mgpu_scan_type_t type, T init, xform_t xform, uint64_t* status,
                               lookback_config_t lb) {
  constexpr int NV = NT * VT;
  const int64_t num_tiles = div_up(count, (int64_t)NV);
  int vec_ok = sizeof(in_t) == 4 && 0 == ((uintptr_t)input & 15) && 0 == ((uintptr_t)output & 15);
  if (type == MGPU_SCAN_INCLUSIVE)
    return launch_kernel("mgpu_scan_lookback", kernel_scan_lookback<NT, VT, true, in_t, T, xform_t>,
                         dim3((unsigned)num_tiles), dim3(NT), 0, stream, input, output, count,
                         num_tiles, init, status, lb, vec_ok, xform);
  return launch_kernel("mgpu_scan_lookback", kernel_scan_lookback<NT, VT, false, in_t, T, xform_t>,
                       dim3((unsigned)num_tiles), dim3(NT), 0, stream, input, output, count,
                       num_tiles, init, status, lb, vec_ok, xform);
}
template <int NT, int VT, typename in_t, typename T, typename xform_t>
hipError_t scan_launch_cleared(hipStream_t stream, const in_t* input, T* output, int64_t count,
                               mgpu_scan_type_t type, T init, xform_t xform, uint64_t* status) {
  return scan_launch_cleared<NT, VT>(stream, input, output, count, type, init, xform, status,
                                     lookback_config(stream));
}

// transform_scan: scan of xform(input[i]).  input/output may alias (in place).
template <int NT, int VT, typename in_t, typename T, typename xform_t>
mgpu_status_t transform_scan_impl(void* aux_data, size_t& aux_size, hipStream_t stream,
                                  const in_t* input, T* output, int64_t count,
                                  mgpu_scan_type_t type, T init, xform_t xform) {
  constexpr int NV = NT * VT;
  if (count < 0) return MGPU_INVALID_ARG;
  if (type != MGPU_SCAN_EXCLUSIVE && type != MGPU_SCAN_INCLUSIVE) return MGPU_INVALID_ARG;
  if (count == 0) return MGPU_OK;
  int64_t num_tiles = div_up(count, (int64_t)NV);
  if (num_tiles >= ((int64_t)1 << 31)) return MGPU_UNSUPPORTED;
  size_t need = aux_bytes<uint64_t>(num_tiles);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if (!input || !output) return MGPU_INVALID_ARG;
  if (mgpu_status_t st = check_aux(aux_data, aux_size, need)) return st;
  void* p = aux_data;
  uint64_t* status = advance_pointer<uint64_t>(p, num_tiles);
  hipError_t e = fill_u32((uint32_t*)status, 2 * num_tiles, 0u, stream);
  if (e != hipSuccess) return hip_status(e);
  return hip_status(scan_launch_cleared<NT, VT>(stream, input, output, count, type, init, xform,
                                                status));
}

// nt and vt only tune: a shape this kernel cannot run (VT not a multiple of 4,
// NT not whole waves, such as the reference's default 128 x 7) runs as the
// default 256 x 80; the result does not depend on the shape.
template <int NT = 256, int VT = 80, typename in_t, typename T, typename xform_t>
mgpu_status_t transform_scan(void* aux_data, size_t& aux_size, hipStream_t stream,
                             const in_t* input, T* output, int64_t count,
                             mgpu_scan_type_t type, T init, xform_t xform) {
  if constexpr (VT >= 4 && VT % 4 == 0 && NT >= warp_size && NT % warp_size == 0 && NT <= 1024)
    return transform_scan_impl<NT, VT>(aux_data, aux_size, stream, input, output, count, type,
                                       init, xform);
  else
    return transform_scan_impl<256, 80>(aux_data, aux_size, stream, input, output, count, type,
                                        init, xform);
}

template <int NT = 256, int VT = 80, typename T>
mgpu_status_t scan(void* aux_data, size_t& aux_size, hipStream_t stream, const T* input,
                   T* output, int64_t count, mgpu_scan_type_t type = MGPU_SCAN_EXCLUSIVE,
                   T init = T()) {
  return transform_scan<NT, VT>(aux_data, aux_size, stream, input, output, count, type, init,
                                identity_t());
}

// In-place exclusive scan with the reference's argument order
// (inc/mgpu/vk/scan.hxx:10-13).
template <int NT = 256, int VT = 80, typename T>
mgpu_status_t scan(void* aux_data, size_t& aux_size, hipStream_t stream, T* data,
                   int64_t count, T init = T()) {
  return scan<NT, VT>(aux_data, aux_size, stream, (const T*)data, data, count,
                      MGPU_SCAN_EXCLUSIVE, init);
}

// The reference's full form, scan<nt, vt>(aux, aux_size, cmd, data, count,
// init, op) (inc/mgpu/vk/scan.hxx:10-13).  Its thread-local steps add with
// + whatever op is (common/cta_scan.hxx:150,164), so plus is the only op with
// a defined result there, and the only one accepted here.
template <int NT = 256, int VT = 80, typename T, typename op_t>
mgpu_status_t scan(void* aux_data, size_t& aux_size, hipStream_t stream, T* data,
                   int64_t count, T init, op_t) {
  static_assert(std::is_same_v<op_t, std::plus<T>> || std::is_same_v<op_t, std::plus<>>,
                "the scan operator is addition, as the reference computes it");
  return scan<NT, VT>(aux_data, aux_size, stream, data, count, init);
}

////////////////////////////////////////////////////////////////////////////////
// Reduce: each block sums one 32 KB tile (8 non-temporal 128-bit loads per
// thread, all in flight together) into a partial, then one block adds the
// partials in index order, so the result is deterministic for a given count.
// float32 sums in float64.  One tile per block, with the dispatcher refilling
// CUs as blocks finish, streams faster than a persistent grid-stride loop:
// 2^29 int32 0.33 ms vs 0.36 ms on MI355X (6.5 vs 5.9 TB/s).
constexpr int reduce_nt = 256, reduce_vpt = 8;

template <typename T> struct reduce_traits;
template <> struct reduce_traits<int32_t> { typedef uint32_t acc_t; };
template <> struct reduce_traits<uint32_t> { typedef uint32_t acc_t; };
template <> struct reduce_traits<float> { typedef double acc_t; };
template <> struct reduce_traits<uint64_t> { typedef uint64_t acc_t; };

template <int NT, typename acc_t>
__device__ inline acc_t block_reduce(acc_t v, acc_t* s_warp) {
  constexpr int NW = NT / warp_size;
  v = wave_reduce_all(v);
  int warp = threadIdx.x / warp_size;
  if (lane_id() == 0) s_warp[warp] = v;
  __syncthreads();
  acc_t total = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) total += s_warp[w];
  return total;
}

template <int NT, int V, typename T>
__global__ __launch_bounds__(NT) void kernel_reduce_partial(
    const T* input, int64_t count, int64_t head,
    typename reduce_traits<T>::acc_t* partials) {
  typedef typename reduce_traits<T>::acc_t acc_t;
  typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
  constexpr int EPV = 16 / (int)sizeof(T);  // elements per 128-bit vector
  __shared__ acc_t s_warp[NT / warp_size];
  const int tid = threadIdx.x;
  acc_t sum = 0;
  // 128-bit vectors start at input + head (16-byte aligned); this block's
  // tile is vectors [v0, v0 + NT * V).
  const int64_t nvec = (count - head) / EPV;
  const int64_t v0 = (int64_t)blockIdx.x * NT * V;
  const u32x4* vp = (const u32x4*)(input + head) + v0 + tid;
  auto addv = [&](u32x4 v) {
    T e[EPV];
    __builtin_memcpy(e, &v, 16);
#pragma unroll
    for (int q = 0; q < EPV; ++q) sum += (acc_t)e[q];
  };
  if (v0 + NT * V <= nvec) {
    u32x4 v[V];
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = __builtin_nontemporal_load(vp + i * NT);
#pragma unroll
    for (int i = 0; i < V; ++i) addv(v[i]);
  } else {
    for (int i = 0; i < V; ++i)
      if (v0 + i * NT + tid < nvec) addv(__builtin_nontemporal_load(vp + i * NT));
  }
  // The unaligned head and the elements after the last whole vector.
  if (blockIdx.x == 0) {
    if (tid < head) sum += (acc_t)input[tid];
    int64_t t = head + nvec * EPV + tid;
    if (tid < EPV && t < count) sum += (acc_t)input[t];
  }
  acc_t total = block_reduce<NT>(sum, s_warp);
  if (tid == 0) partials[blockIdx.x] = total;
}

template <int NT, typename T>
__global__ __launch_bounds__(NT) void kernel_reduce_final(
    const typename reduce_traits<T>::acc_t* partials, int num, T* out, T init) {
  typedef typename reduce_traits<T>::acc_t acc_t;
  __shared__ acc_t s_warp[NT / warp_size];
  // Eight loads in flight per thread; the order of the additions is fixed by
  // num alone, so the result stays deterministic.
  acc_t sum = 0;
  int i = threadIdx.x;
  for (; i + 7 * NT < num; i += 8 * NT) {
    acc_t x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = partials[i + j * NT];
#pragma unroll
    for (int j = 0; j < 8; ++j) sum += x[j];
  }
  for (; i < num; i += NT) sum += partials[i];
  acc_t total = block_reduce<NT>(sum, s_warp);
  if (threadIdx.x == 0) {
    if constexpr (std::is_same_v<T, float>)
      out[0] = (float)((double)init + total);
    else
      out[0] = (T)((acc_t)init + total);
  }
}

template <int NT = reduce_nt, typename T>
mgpu_status_t reduce(void* aux_data, size_t& aux_size, hipStream_t stream, const T* input,
                     int64_t count, T* out, T init = T()) {
  typedef typename reduce_traits<T>::acc_t acc_t;
  constexpr int V = reduce_vpt;
  constexpr int64_t tile = (int64_t)NT * V * (16 / (int)sizeof(T));  // elements per block
  if (count < 0) return MGPU_INVALID_ARG;
  // One partial per tile (the misaligned head shifts the tiles by < 16 B,
  // so div_up(count, tile) tiles always cover the vectors).
  const int64_t blocks = max_<int64_t>(div_up(count, tile), 1);
  if (blocks >= ((int64_t)1 << 31)) return MGPU_UNSUPPORTED;
  size_t need = aux_bytes<acc_t>(blocks);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if (!out || (count > 0 && !input)) return MGPU_INVALID_ARG;
  if (mgpu_status_t st = check_aux(aux_data, aux_size, need)) return st;
  void* p = aux_data;
  acc_t* partials = advance_pointer<acc_t>(p, blocks);
  int64_t mis = ((uintptr_t)input & 15) / sizeof(T);
  int64_t head = mis ? min_<int64_t>((16 / (int64_t)sizeof(T)) - mis, count) : 0;
  hipError_t e = launch_kernel("mgpu_reduce_partial", kernel_reduce_partial<NT, V, T>,
                               dim3((unsigned)blocks), dim3(NT), 0, stream, input, count, head,
                               partials);
  if (e != hipSuccess) return hip_status(e);
  e = launch_kernel("mgpu_reduce_final", kernel_reduce_final<1024, T>, dim3(1), dim3(1024), 0,
                    stream, (const acc_t*)partials, (int)blocks, out, init);
  return hip_status(e);
}

////////////////////////////////////////////////////////////////////////////////
// 64-bit integer scan (modulo 2^64), single pass with decoupled look-back like
// the 32-bit scan.  A 64-bit value and a flag do not fit one atomic word, so a
// tile's status is two 64-bit words, each carrying the flag in its high half
// and one 32-bit half of the value in its low half.  Each word is written and
// read atomically; a reader takes the status only when both words carry the
// same flag (a torn pair, one word updated and the other not yet, is re-read),
// so no ordering between the two stores is needed.  8 bytes read and 8
// written per element.

__device__ inline void scan64_publish(uint64_t* status, int64_t tile, uint64_t v, uint32_t flag) {
  const uint64_t hi_flag = (uint64_t)flag << 32;
  atomic_store_agent(status + 2 * tile, hi_flag | (v & (uint64_t)0xFFFFFFFFu));
  atomic_store_agent(status + 2 * tile + 1, hi_flag | (v >> 32));
}

// Flag of a word pair: the common flag, or 3 for a torn pair.
constexpr uint32_t status64_torn = 3;
__device__ inline uint32_t scan64_read(const uint64_t* status, int64_t tile, uint64_t& v) {
  const uint64_t w0 = atomic_load_agent(status + 2 * tile);
  const uint64_t w1 = atomic_load_agent(status + 2 * tile + 1);
  v = (w1 << 32) | (w0 & (uint64_t)0xFFFFFFFFu);
  const uint32_t f0 = (uint32_t)(w0 >> 32), f1 = (uint32_t)(w1 >> 32);
  return f0 == f1 ? f0 : status64_torn;
}

// Exclusive prefix of tile `tile` (one full wave; every lane returns it), as
// scan_lookback: 64 predecessors per round, bounded spin, then the first
// missing tile's sum from its input with the in-place re-check.  A torn pair
// means its owner is between its two stores, so it is only ever re-polled.
template <int NV, typename T>
__device__ inline uint64_t scan64_lookback(const uint64_t* status, int64_t tile, uint64_t init,
                                           const T* input, lookback_config_t lb) {
  const int lane = lane_id();
  uint64_t excl = 0;
  int64_t j = tile - 1;
  uint32_t spins = 0;
  while (true) {
    const int64_t idx = j - lane;
    uint64_t v = 0;
    uint32_t flag;
    if (idx >= 0) {
      flag = scan64_read(status, idx, v);
    } else {
      flag = status_inclusive;
      v = idx == -1 ? init : 0;
    }
    const uint64_t inc_mask = __ballot(flag == status_inclusive);
    const uint64_t bad_mask = __ballot(flag == status_invalid || flag == status64_torn);
    const uint64_t lowest = inc_mask & (~inc_mask + 1);
    const uint64_t range = inc_mask ? (lowest | (lowest - 1)) : ~0ull;
    if (bad_mask & range) {
      const int kb = __builtin_ctzll(bad_mask & range);
      const bool torn = __shfl(flag, kb, warp_size) == status64_torn;
      if (++spins <= lb.spin_limit || torn) {
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
      const int64_t jj = j - kb;
      uint64_t agg = 0;
      for (int e = lane; e < NV; e += warp_size) agg += (uint64_t)input[jj * NV + e];
      agg = wave_reduce_all(agg);
      uint64_t again;
      acquire_agent();
      const uint32_t f_again = scan64_read(status, jj, again);
      excl += wave_reduce_all(lane < kb ? v : (uint64_t)0);
      spins = 0;
      if (f_again != status_invalid) {
        j = jj;  // published meanwhile: continue from its words
        continue;
      }
      excl += agg;
      if (lane == 0 && lb.fallbacks) atomic_add_agent(lb.fallbacks, 1u);
      j = jj - 1;
      continue;
    }
    excl += wave_reduce_all(((range >> lane) & 1) ? v : (uint64_t)0);
    if (inc_mask) break;
    j -= warp_size;
  }
  return excl;
}

// One tile of NT * VT elements in VT / 2 rows of 2 * NT: each thread loads two
// consecutive elements of every row (one 16-byte load), sums them, and a wave
// scan of the pair sums plus the row's wave totals (in LDS) give every prefix
// inside the tile; the tile's carry comes from the look-back.  Large tiles
// keep the look-back off the critical path (the tiles' status words are
// exchanged across XCDs at a few memory round trips per tile).
template <int NT, int VT, bool inclusive, typename T>
__global__ __launch_bounds__(NT) void kernel_scan64_lookback(const T* input, T* output,
                                                             int64_t count, T init,
                                                             uint64_t* status,
                                                             lookback_config_t lb) {
  constexpr int NV = NT * VT;
  constexpr int NW = NT / warp_size;
  constexpr int NR = VT / 2;
  static_assert(VT % 2 == 0, "VT must be even (pairs per row)");
  __shared__ uint64_t s_wave[NR][NW];
  __shared__ uint64_t s_carry;
  const int tid = threadIdx.x, lane = lane_id(), warp = tid / warp_size;
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * NV;
  const bool full = count - base >= NV && 0 == ((uintptr_t)input & 15) &&
                    0 == ((uintptr_t)output & 15);
  typedef __attribute__((ext_vector_type(2))) uint64_t u64x2;
  uint64_t x[VT];
  if (full) {
    const u64x2* in2 = (const u64x2*)(input + base);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const u64x2 v = __builtin_nontemporal_load(in2 + r * NT + tid);
      x[2 * r] = v.x;
      x[2 * r + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int64_t idx = base + (int64_t)(r * NT + tid) * 2 + q;
        x[2 * r + q] = idx < count ? (uint64_t)input[idx] : 0ull;
      }
  }
  uint64_t w_excl[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    x[2 * r + 1] += x[2 * r];
    uint64_t incl = x[2 * r + 1];
#pragma unroll
    for (int off = 1; off < warp_size; off <<= 1) {
      const uint64_t y = __shfl_up(incl, off, warp_size);
      if (lane >= off) incl += y;
    }
    if (lane == warp_size - 1) s_wave[r][warp] = incl;
    w_excl[r] = wave_shift_up1(incl, (uint64_t)0);
  }
  __syncthreads();
  if (warp == 0) {
    const uint64_t* flat = &s_wave[0][0];
    uint64_t aggregate = 0;
    for (int i = lane; i < NR * NW; i += warp_size) aggregate += flat[i];
    aggregate = wave_reduce_all(aggregate);
    uint64_t carry;
    if (tile == 0) {
      carry = (uint64_t)init;
    } else {
      if (lane == 0) {
        scan64_publish(status, tile, aggregate, status_aggregate);
        // The aggregate must be visible before this tile's output (possibly
        // over its own input) can be.
        status_store_complete();
      }
      carry = scan64_lookback<NV>(status, tile, (uint64_t)init, input, lb);
    }
    if (lane == 0) {
      scan64_publish(status, tile, carry + aggregate, status_inclusive);
      s_carry = carry;
    }
  }
  __syncthreads();
  // Outputs replace the inputs in x: x[2r] is the first element of the
  // thread's pair, x[2r + 1] the pair's sum.
  uint64_t run = s_carry;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    uint64_t pre = run, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint64_t v = s_wave[r][w];
      if (w < warp) pre += v;
      tot += v;
    }
    run += tot;
    pre += w_excl[r];
    const uint64_t first = x[2 * r], pair = x[2 * r + 1];
    x[2 * r] = inclusive ? pre + first : pre;
    x[2 * r + 1] = inclusive ? pre + pair : pre + first;
  }
  if (full) {
    u64x2* out2 = (u64x2*)(output + base);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      u64x2 v;
      v.x = x[2 * r];
      v.y = x[2 * r + 1];
      __builtin_nontemporal_store(v, out2 + r * NT + tid);
    }
  } else {
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int64_t idx = base + (int64_t)(r * NT + tid) * 2 + q;
        if (idx < count) output[idx] = (T)x[2 * r + q];
      }
  }
}

constexpr int scan64_nt = 256, scan64_vt = 48;
// Scans of up to 2^20 elements take 256 x 8 tiles instead: a few large
// tiles leave the look-back chain and each tile's loads exposed, more small
// ones overlap them.
constexpr int scan64_small_vt = 8;
constexpr int64_t scan64_small_count = (int64_t)1 << 20;
inline int64_t scan64_tile_size(int64_t count) {
  return (int64_t)scan64_nt * (count <= scan64_small_count ? scan64_small_vt : scan64_vt);
}

// The kernel of scan64, for a caller that has already zeroed the status words
// (two 64-bit words per tile of scan64_tile_size(count) elements).  Arguments are not
// checked again.
template <typename T>
hipError_t scan64_launch_cleared(hipStream_t stream, const T* input, T* output, int64_t count,
                                 mgpu_scan_type_t type, T init, uint64_t* status) {
  constexpr int NT = scan64_nt;
  const int64_t num_tiles = div_up(count, scan64_tile_size(count));
  const lookback_config_t lb = lookback_config(stream);
  auto go = [&](auto vt, auto inclusive) {
    constexpr int VT = decltype(vt)::value;
    constexpr bool INC = decltype(inclusive)::value;
    return launch_kernel("mgpu_scan64_lookback", kernel_scan64_lookback<NT, VT, INC, T>,
                         dim3((unsigned)num_tiles), dim3(NT), 0, stream, input, output, count,
                         init, status, lb);
  };
  const bool inc = type == MGPU_SCAN_INCLUSIVE;
  if (count <= scan64_small_count)
    return inc ? go(std::integral_constant<int, scan64_small_vt>(), std::true_type())
               : go(std::integral_constant<int, scan64_small_vt>(), std::false_type());
  return inc ? go(std::integral_constant<int, scan64_vt>(), std::true_type())
             : go(std::integral_constant<int, scan64_vt>(), std::false_type());
}

// Scan of 64-bit integers (uint64_t / int64_t, wrapping modulo 2^64); input
// and output may alias.
template <typename T>
mgpu_status_t scan64(void* aux_data, size_t& aux_size, hipStream_t stream, const T* input,
                     T* output, int64_t count, mgpu_scan_type_t type, T init) {
  static_assert(sizeof(T) == 8, "64-bit element types");
  if (count < 0) return MGPU_INVALID_ARG;
  if (type != MGPU_SCAN_EXCLUSIVE && type != MGPU_SCAN_INCLUSIVE) return MGPU_INVALID_ARG;
  if (count == 0) return MGPU_OK;
  const int64_t num_tiles = div_up(count, scan64_tile_size(count));
  if (num_tiles >= ((int64_t)1 << 31)) return MGPU_UNSUPPORTED;
  const size_t need = aux_bytes<uint64_t>(2 * num_tiles);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if (!input || !output) return MGPU_INVALID_ARG;
  if (mgpu_status_t st = check_aux(aux_data, aux_size, need)) return st;
  uint64_t* status = (uint64_t*)aux_data;
  hipError_t e = fill_u32((uint32_t*)status, 4 * num_tiles, 0u, stream);
  if (e != hipSuccess) return hip_status(e);
  return hip_status(scan64_launch_cleared<T>(stream, input, output, count, type, init, status));
}

}  // namespace hip
}  // namespace mgpu
