// Batched segmented scan for gfx950: one prefix sum per segment, all
// segments in one single-pass launch.
//
// No counterpart in the reference (its scan, inc/mgpu/vk/scan.hxx:10-115, has
// no segmented form); the call shape is the scan's with the segmented sort's
// offsets array: `segments[i]` = first element of segment i, sorted
// ascending, index 0 always starts a segment.  Each segment is scanned from
// `init`, exclusive or inclusive:
//   exclusive: out[i] = init + x[h] + ... + x[i-1]      (h = start of i's segment)
//   inclusive: out[i] = init + x[h] + ... + x[i]
//
// Scan of (head flag, value) pairs under (fa, va) . (fb, vb) = (fa | fb,
// fb ? vb : va + vb), in the hierarchical order of the plain scan (thread-
// serial over 4 elements, a DPP wave scan, the wave totals, the carry); as
// there, every combination above an aligned 256-element wave block is taken
// in the carry type (float64 for float32 data), so float32 results have at
// most 10 float32 additions on their path and do not depend on the tile
// shape.  Tiles are 256 x 32 by default (8192 elements): fewer tiles, fewer
// look-backs per byte.
// Tile heads come from a small search kernel, tile_first[t] = number of
// segments starting before t * NV, so a tile's heads are
// segments[tile_first[t] .. tile_first[t + 1]) and are marked in an LDS
// bitmap.  A tile that contains a head knows its carry-out without looking
// back (it publishes the inclusive word at once); look-back stops at the
// first such tile, so short segments need no waiting at all.
#pragma once
#include "scan.hxx"

namespace mgpu {
namespace hip {

// tile_first[t] = lower_bound(segments, t * NV) for t in [0, num_tiles].
template <int NT>
__global__ __launch_bounds__(NT) void kernel_segscan_tiles(const int32_t* segments,
                                                           int64_t num_segments, int64_t nv,
                                                           int64_t num_tiles,
                                                           int32_t* tile_first) {
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t > num_tiles) return;
  const int64_t key = t * nv;
  int64_t lo = 0, hi = num_segments;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)segments[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  tile_first[t] = (int32_t)lo;
}

// v of the DPP source lane (0 where the row mask or the shift leaves none).
template <typename V, int ctrl, int rmask> struct dpp_mover;
template <int ctrl, int rmask> struct dpp_mover<uint32_t, ctrl, rmask> {
  static __device__ uint32_t move(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, rmask, 0xf, false);
  }
};
template <int ctrl, int rmask> struct dpp_mover<float, ctrl, rmask> {
  static __device__ float move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, rmask, 0xf,
                                                      false));
  }
};
template <typename V, int ctrl, int rmask>
__device__ inline V dpp_move(V v) {
  return dpp_mover<V, ctrl, rmask>::move(v);
}

// Inclusive segmented wave scan of (f, v) over 64 lanes (lane order).
template <typename V>
__device__ inline void wave_segscan_incl(uint32_t& f, V& v) {
#define MGPU_SEG_STEP(ctrl, rmask)                                                       \
  {                                                                                      \
    uint32_t fp = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)f, ctrl, rmask, 0xf,     \
                                                        false);                          \
    V vp = dpp_move<V, ctrl, rmask>(v);                                                  \
    v = f ? v : (V)(vp + v);                                                             \
    f |= fp;                                                                             \
  }
  MGPU_SEG_STEP(0x111, 0xf)
  MGPU_SEG_STEP(0x112, 0xf)
  MGPU_SEG_STEP(0x114, 0xf)
  MGPU_SEG_STEP(0x118, 0xf)
  MGPU_SEG_STEP(0x142, 0xa)
  MGPU_SEG_STEP(0x143, 0xc)
#undef MGPU_SEG_STEP
}

// Carry-out contribution of tile jj computed from its input by one wave:
// (has_head, sum from its last head, or from its first element, to its end).
template <typename carry_t, typename acc_t, int NV, typename in_t, typename T, typename xform_t>
__device__ inline carry_t segscan_tile_tail(const in_t* input, const int32_t* segments,
                                            const int32_t* tile_first, int64_t jj,
                                            xform_t xform, bool& has_head) {
  const int64_t base = jj * NV;
  int64_t first = base;
  const int32_t k0 = tile_first[jj], k1 = tile_first[jj + 1];
  has_head = false;
  if (k1 > k0) {
    const int64_t h = segments[k1 - 1];
    if (h >= base && h < base + NV) {
      first = h;
      has_head = true;
    }
  }
  carry_t sum = 0;
  for (int64_t e = first + lane_id(); e < base + NV; e += warp_size)
    sum += (carry_t)(acc_t)(T)xform(input[e]);
  return wave_reduce_all(sum);
}

// Exclusive carry of tile `tile` (the scan_lookback protocol: bounded spin,
// then the missing predecessor's contribution from its input, re-checked for
// in-place scans).  A predecessor whose carry-out is complete on its own
// (it holds a head) publishes an inclusive word, so the walk stops there.
template <typename carry_t, typename acc_t, int NV, typename in_t, typename T, typename xform_t>
__device__ inline carry_t segscan_lookback(const uint64_t* status, int64_t tile, carry_t init,
                                           const in_t* input, const int32_t* segments,
                                           const int32_t* tile_first, xform_t xform,
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
      bool head;
      carry_t tail = segscan_tile_tail<carry_t, acc_t, NV, in_t, T>(input, segments, tile_first,
                                                                    jj, xform, head);
      acquire_agent();
      const uint64_t again = atomic_load_agent(status + jj);
      carry_t v = lane < kb ? unpack_status<carry_t>(w) : carry_t(0);
      excl += wave_reduce_all(v);
      spins = 0;
      if ((again & 3) != status_invalid) {
        j = jj;  // published meanwhile: continue from its word
        continue;
      }
      if (lane == 0 && lb.fallbacks) atomic_add_agent(lb.fallbacks, 1u);
      if (head) {
        excl += init + tail;
        break;
      }
      excl += tail;
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

template <typename acc_t>
struct seg_pair_t {
  uint32_t f;
  acc_t v;
};
template <typename acc_t>
__device__ inline seg_pair_t<acc_t> seg_combine(seg_pair_t<acc_t> a, seg_pair_t<acc_t> b) {
  return {a.f | b.f, b.f ? b.v : (acc_t)(a.v + b.v)};
}

// (flag, value) pair in the carry type, for sums above a wave row.
template <typename carry_t>
struct seg_cpair_t {
  uint32_t f;
  carry_t v;
};
template <typename carry_t>
__device__ inline seg_cpair_t<carry_t> seg_combine_c(seg_cpair_t<carry_t> a, uint32_t bf,
                                                     carry_t bv) {
  return {a.f | bf, bf ? bv : (carry_t)(a.v + bv)};
}

// Result of one element: `pre` is the pair of everything in the tile before
// its wave row segment (carry type), `e` its prefix inside the segment.
__device__ inline uint32_t segscan_finish(uint32_t carry, uint32_t init,
                                          seg_cpair_t<uint32_t> pre, seg_pair_t<uint32_t> e) {
  if (e.f) return init + e.v;
  return (pre.f ? init : carry) + pre.v + e.v;
}
__device__ inline float segscan_finish(double carry, float init, seg_cpair_t<double> pre,
                                       seg_pair_t<float> e) {
  if (e.f) return (float)((double)init + (double)e.v);
  return (float)((pre.f ? (double)init : carry) + pre.v + (double)e.v);
}

template <int NT, int VT, bool inclusive, typename in_t, typename T, typename xform_t>
__global__ __launch_bounds__(NT) void kernel_segscan(
    const in_t* input, T* output, int64_t count, const int32_t* segments,
    const int32_t* tile_first, T init, uint64_t* status, lookback_config_t lb, int vec_ok,
    xform_t xform) {
  typedef scan_traits<T> traits;
  typedef typename traits::acc_t acc_t;
  typedef typename traits::carry_t carry_t;
  typedef seg_pair_t<acc_t> pair_t;
  typedef seg_cpair_t<carry_t> cpair_t;
  constexpr int NW = NT / warp_size;
  constexpr int NV = NT * VT;
  constexpr int NR = VT / 4;
  static_assert(VT % 4 == 0, "VT must be a multiple of 4");
  static_assert(sizeof(T) == 4, "segmented scan works on 32-bit element types");

  __shared__ uint32_t s_heads[NV / 32];
  __shared__ uint32_t s_wave_f[NR][NW];
  __shared__ acc_t s_wave_v[NR][NW];
  __shared__ carry_t s_carry;

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int warp = tid / warp_size;
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * NV;

  acc_t x[VT];
  scan_load_tile<NT, VT, acc_t, in_t, T>(input, count, tile, vec_ok, xform, x);

  // Head bitmap of this tile: bit e of the tile <-> element base + e.
  for (int i = tid; i < NV / 32; i += NT) s_heads[i] = 0;
  __syncthreads();
  {
    const int32_t k0 = tile_first[tile], k1 = tile_first[tile + 1];
    for (int32_t k = k0 + tid; k < k1; k += NT) {
      const int64_t h = (int64_t)segments[k] - base;
      if (h >= 0 && h < NV) atomicOr(&s_heads[h >> 5], 1u << (h & 31));
    }
  }
  __syncthreads();

  // Thread-serial pairs over each group of 4, then a wave scan of the group
  // totals for every row.
  uint32_t fl[VT];
  pair_t p[VT];
  pair_t w_excl[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int g = r * NT + tid;  // group of 4 elements: bits 4g .. 4g+3
    const uint32_t bits = (s_heads[g >> 3] >> ((g & 7) * 4)) & 15u;
#pragma unroll
    for (int q = 0; q < 4; ++q) fl[4 * r + q] = (bits >> q) & 1u;
    p[4 * r] = {fl[4 * r], x[4 * r]};
#pragma unroll
    for (int q = 1; q < 4; ++q) p[4 * r + q] = seg_combine(p[4 * r + q - 1], {fl[4 * r + q], x[4 * r + q]});
    uint32_t f = p[4 * r + 3].f;
    acc_t v = p[4 * r + 3].v;
    wave_segscan_incl(f, v);
    if (lane == warp_size - 1) {
      s_wave_f[r][warp] = f;
      s_wave_v[r][warp] = v;
    }
    w_excl[r] = {wave_shift_up1(f, 0u), wave_shift_up1(v, acc_t(0))};
  }
  __syncthreads();

  // Tile aggregate as a (flag, carry-type) pair, combined by wave 0 over the
  // NR x NW wave totals in order.
  cpair_t aggregate = {0u, carry_t(0)};
  if (warp == 0) {
#pragma unroll 4
    for (int i = 0; i < NR * NW; ++i) {
      const int r = i / NW, w = i % NW;
      aggregate = seg_combine_c(aggregate, s_wave_f[r][w], (carry_t)s_wave_v[r][w]);
    }
  }

  // A tile holding a head knows its carry-out, init + (sum after its last
  // head), without looking back and publishes it as inclusive at once; it
  // still needs the carry-in for its elements before the first head.
  if (warp == 0) {
    const bool head_tile = aggregate.f != 0;
    carry_t carry = (carry_t)(acc_t)init;
    if (tile > 0) {
      if (lane == 0) {
        atomic_store_agent(status + tile,
                           head_tile ? pack_status((carry_t)((carry_t)(acc_t)init + aggregate.v),
                                                   status_inclusive)
                                     : pack_status(aggregate.v, status_aggregate));
        // The word must be visible before this tile's output (possibly over
        // its own input) can be.
        status_store_complete();
      }
      if (!head_tile || !(s_heads[0] & 1u))
        carry = segscan_lookback<carry_t, acc_t, NV, in_t, T>(
            status, tile, (carry_t)(acc_t)init, input, segments, tile_first, xform, lb);
    }
    if (lane == 0) {
      if (tile == 0 || !head_tile)
        atomic_store_agent(status + tile,
                           pack_status((carry_t)((head_tile ? (carry_t)(acc_t)init : carry) +
                                                 aggregate.v),
                                       status_inclusive));
      s_carry = carry;
    }
  }
  __syncthreads();
  const carry_t carry = s_carry;

  // Everything above one wave row (the carry, earlier rows, earlier waves of
  // this row) is combined in the carry type; inside a row the sums are acc_t.
  T o[VT];
  cpair_t run = {0u, carry_t(0)};
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    cpair_t pre = run;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t wf = s_wave_f[r][w];
      const carry_t wv = (carry_t)s_wave_v[r][w];
      if (w < warp) pre = seg_combine_c(pre, wf, wv);
      run = seg_combine_c(run, wf, wv);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      pair_t e;  // in-row prefix pair of this element (acc_t)
      if (inclusive) {
        e = seg_combine(w_excl[r], p[4 * r + q]);
      } else if (fl[4 * r + q]) {
        e = {1u, acc_t(0)};
      } else {
        e = q == 0 ? w_excl[r] : seg_combine(w_excl[r], p[4 * r + q - 1]);
      }
      o[4 * r + q] = (T)segscan_finish(carry, (acc_t)init, pre, e);
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

template <int NT = 256, int VT = 32>
size_t segmented_scan_aux_bytes(int64_t count) {
  const int64_t num_tiles = div_up(count, (int64_t)(NT * VT));
  return aux_bytes<uint64_t>(num_tiles) + aux_bytes<int32_t>(num_tiles + 1);
}

// Segmented transform_scan: per segment, a scan of xform(input[i]) from
// `init`.  segments: ascending starts, num_segments >= 0 (index 0 always
// starts a segment; starts at or beyond count are empty trailing segments).
// input/output may alias.
template <int NT = 256, int VT = 32, typename in_t, typename T, typename xform_t>
mgpu_status_t segmented_transform_scan(void* aux_data, size_t& aux_size, hipStream_t stream,
                                       const in_t* input, T* output, int64_t count,
                                       const int32_t* segments, int64_t num_segments,
                                       mgpu_scan_type_t type, T init, xform_t xform) {
  constexpr int NV = NT * VT;
  if (count < 0 || num_segments < 0) return MGPU_INVALID_ARG;
  if (type != MGPU_SCAN_EXCLUSIVE && type != MGPU_SCAN_INCLUSIVE) return MGPU_INVALID_ARG;
  if (count == 0) return MGPU_OK;
  if (count >= ((int64_t)1 << 31) || num_segments >= ((int64_t)1 << 31)) return MGPU_UNSUPPORTED;
  const int64_t num_tiles = div_up(count, (int64_t)NV);
  const size_t need = segmented_scan_aux_bytes<NT, VT>(count);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if (!input || !output || (num_segments > 0 && !segments)) return MGPU_INVALID_ARG;
  if (mgpu_status_t st = check_aux(aux_data, aux_size, need)) return st;
  void* p = aux_data;
  uint64_t* status = advance_pointer<uint64_t>(p, num_tiles);
  int32_t* tile_first = advance_pointer<int32_t>(p, num_tiles + 1);
  const lookback_config_t lb = lookback_config(stream);
  hipError_t e = fill_u32((uint32_t*)status, 2 * num_tiles, 0u, stream);
  if (e != hipSuccess) return hip_status(e);
  if (num_segments > 0) {
    e = launch_kernel("mgpu_segscan_tiles", kernel_segscan_tiles<256>,
                      dim3((unsigned)div_up(num_tiles + 1, (int64_t)256)), dim3(256), 0, stream,
                      segments, num_segments, (int64_t)NV, num_tiles, tile_first);
  } else {
    e = fill_u32((uint32_t*)tile_first, num_tiles + 1, 0u, stream);
  }
  if (e != hipSuccess) return hip_status(e);
  const int vec_ok =
      sizeof(in_t) == 4 && 0 == ((uintptr_t)input & 15) && 0 == ((uintptr_t)output & 15);
  if (type == MGPU_SCAN_INCLUSIVE)
    e = launch_kernel("mgpu_segscan", kernel_segscan<NT, VT, true, in_t, T, xform_t>,
                      dim3((unsigned)num_tiles), dim3(NT), 0, stream, input, output, count,
                      segments, (const int32_t*)tile_first, init, status, lb, vec_ok, xform);
  else
    e = launch_kernel("mgpu_segscan", kernel_segscan<NT, VT, false, in_t, T, xform_t>,
                      dim3((unsigned)num_tiles), dim3(NT), 0, stream, input, output, count,
                      segments, (const int32_t*)tile_first, init, status, lb, vec_ok, xform);
  return hip_status(e);
}

template <int NT = 256, int VT = 32, typename T>
mgpu_status_t segmented_scan(void* aux_data, size_t& aux_size, hipStream_t stream,
                             const T* input, T* output, int64_t count, const int32_t* segments,
                             int64_t num_segments, mgpu_scan_type_t type = MGPU_SCAN_EXCLUSIVE,
                             T init = T()) {
  return segmented_transform_scan<NT, VT>(aux_data, aux_size, stream, input, output, count,
                                          segments, num_segments, type, init, identity_t());
}

}  // namespace hip
}  // namespace mgpu
