// Merge-path mergesort for gfx950.
//
// Structure as in the reference (inc/mgpu/vk/mergesort.hxx:10-84):
//   1. blocksort: each nv-tile is sorted by a stable odd-even transposition
//      network over vt registers (common/sort_networks.hxx:11-39, with the
//      partial-tile mask of common/cta_mergesort.hxx:7-15) and log2(nt) merge
//      passes in LDS (common/cta_mergesort.hxx:87-133);
//   2. ceil(log2(tiles)) global passes; each pass partitions the coop groups of
//      2<<pass tiles with merge path (vk/mergesort.hxx:65-72) and merges one
//      output tile per block (common/kernel_mergesort.hxx:72-121).
// The ping-pong parity lands the result in the caller's buffers
// (vk/mergesort.hxx:47-48).  Every comparison is strict (swap only when
// comp(b, a)), and merges take A on ties, so the sort is stable and its output
// equals std::stable_sort.
#pragma once
#include "merge.hxx"

namespace mgpu {
namespace hip {

struct merge_range_t {
  int64_t a_begin, a_end, b_begin, b_end;
  __host__ __device__ int64_t a_count() const { return a_end - a_begin; }
  __host__ __device__ int64_t b_count() const { return b_end - b_begin; }
};

// common/cta_mergesort.hxx:17-31.
__host__ __device__ inline merge_range_t compute_mergesort_frame(int64_t partition, int64_t coop,
                                                                 int64_t spacing) {
  int64_t size = spacing * (coop / 2);
  int64_t start = ~(coop - 1) & partition;
  int64_t a_begin = spacing * start;
  int64_t b_begin = spacing * start + size;
  return merge_range_t{a_begin, a_begin + size, b_begin, b_begin + size};
}

// common/cta_mergesort.hxx:33-43.
__host__ __device__ inline merge_range_t compute_mergesort_range(int64_t count, int64_t partition,
                                                                 int64_t coop, int64_t spacing) {
  merge_range_t f = compute_mergesort_frame(partition, coop, spacing);
  return merge_range_t{f.a_begin, min_(count, f.a_end), min_(count, f.b_begin),
                       min_(count, f.b_end)};
}

// common/cta_mergesort.hxx:45-68: the range of partition `partition` given its
// two merge-path splits; the last partition of a coop group keeps the group's
// end because its mp1 belongs to the next group's coordinates.
__host__ __device__ inline merge_range_t compute_mergesort_range(int64_t count, int64_t partition,
                                                                 int64_t coop, int64_t spacing,
                                                                 int64_t mp0, int64_t mp1) {
  merge_range_t range = compute_mergesort_range(count, partition, coop, spacing);
  int64_t diag = spacing * partition - range.a_begin;
  if (coop - 1 != ((coop - 1) & partition)) {
    range.a_end = range.a_begin + mp1;
    range.b_end = min_(count, range.b_begin + diag + spacing - mp1);
  }
  range.a_begin = range.a_begin + mp0;
  range.b_begin = min_(count, range.b_begin + diag - mp0);
  return range;
}

template <int NT, int VT, typename key_t, typename val_t>
struct blocksort_storage_t {
  static constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  key_t keys[NT * VT + 1];
  typename std::conditional<has_values, val_t, char>::type vals[has_values ? NT * VT + 1 : 1];
};

// Stable odd-even transposition sort of VT registers; bit j of `oor` marks
// slot j as out of range (it never moves and nothing swaps with it).
// Each compare-exchange is written as selects (v_cndmask), not as a
// conditional swap: the compiler turns a conditional swap of several
// registers into an exec-masked branch with copies on both sides.
template <int VT, bool has_values, bool masked = true, typename key_t, typename val_t,
          typename comp_t>
__device__ inline void odd_even_sort(key_t (&k)[VT], val_t (&v)[VT], uint32_t oor, comp_t comp) {
#pragma unroll
  for (int I = 0; I < VT; ++I) {
#pragma unroll
    for (int i = 1 & I; i < VT - 1; i += 2) {
      const bool swap = (!masked || !((oor >> (i + 1)) & 1)) && comp(k[i + 1], k[i]);
      const key_t lo = swap ? k[i + 1] : k[i];
      const key_t hi = swap ? k[i] : k[i + 1];
      k[i] = lo;
      k[i + 1] = hi;
      if constexpr (has_values) {
        const val_t vlo = swap ? v[i + 1] : v[i];
        const val_t vhi = swap ? v[i] : v[i + 1];
        v[i] = vlo;
        v[i + 1] = vhi;
      }
    }
  }
}

// One in-LDS merge pass of the blocksort (common/cta_mergesort.hxx:87-113).
template <int NT, int VT, bool has_values, typename key_t, typename val_t, typename comp_t>
__device__ inline void block_merge_pass(key_t (&k)[VT], val_t (&v)[VT], int tid, int count,
                                        int pass, comp_t comp,
                                        blocksort_storage_t<NT, VT, key_t, val_t>& s) {
  const int coop = 2 << pass;
  merge_range_t range = compute_mergesort_range(count, tid, coop, VT);
  const int a_begin = (int)range.a_begin, a_end = (int)range.a_end;
  const int b_begin = (int)range.b_begin, b_end = (int)range.b_end;
  const int diag = VT * tid - a_begin;
#pragma unroll
  for (int i = 0; i < VT; ++i) s.keys[VT * tid + i] = k[i];
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[VT * tid + i] = v[i];
  }
  __syncthreads();
  int mp = merge_path_lds(s.keys + a_begin, a_end - a_begin, s.keys + b_begin, b_end - b_begin,
                          diag, comp);
  int idx[VT];
  serial_merge<VT>(s.keys, a_begin + mp, a_end, b_begin + diag - mp, b_end, k, idx, comp);
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) v[i] = s.vals[idx[i]];
  }
  __syncthreads();
}

// Sort tile_count (<= NT*VT) keys (and values) of one tile in LDS and write
// them out: the body of the blocksort kernel (common/kernel_mergesort.hxx:
// 12-67), shared with the segmented sort.  With sort_indices the values are
// index_base + position (common/kernel_mergesort.hxx:43-49).  keys_in may
// equal keys_out (the tile is fully read before it is written).
template <int NT, int VT, bool sort_indices, typename key_t, typename val_t, typename comp_t>
__device__ inline void cta_blocksort(const key_t* keys_in, const val_t* vals_in, key_t* keys_out,
                                     val_t* vals_out, int tile_count, int64_t index_base,
                                     comp_t comp, blocksort_storage_t<NT, VT, key_t, val_t>& s) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr int NV = NT * VT;
  const int tid = threadIdx.x;
  if (tile_count <= 0) return;
  // Unguarded loads of a clamped index (slots past the end get a copy of the
  // last key; the out-of-range mask keeps them out of the sort), so that all
  // loads are in flight together instead of one wait per guarded load.
  {
    key_t x[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i) x[i] = keys_in[min_(i * NT + tid, tile_count - 1)];
#pragma unroll
    for (int i = 0; i < VT; ++i) s.keys[i * NT + tid] = x[i];
  }
  if constexpr (has_values && !sort_indices) {
    val_t x[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i) x[i] = vals_in[min_(i * NT + tid, tile_count - 1)];
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[i * NT + tid] = x[i];
  }
  __syncthreads();
  key_t k[VT];
  val_t v[VT];
#pragma unroll
  for (int i = 0; i < VT; ++i) k[i] = s.keys[VT * tid + i];
  if constexpr (sort_indices) {
#pragma unroll
    for (int i = 0; i < VT; ++i) v[i] = (val_t)(index_base + VT * tid + i);
  } else if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) v[i] = s.vals[VT * tid + i];
  }
  __syncthreads();

  uint32_t oor = 0;
  if (tile_count < NV) {
    int valid = tile_count - VT * tid;
    for (int i = 0; i < VT; ++i)
      if (i >= valid) oor |= 1u << i;
  }
  // Full tiles (every thread's slots in range) skip the mask test.
  if (tile_count == NV) odd_even_sort<VT, has_values, false>(k, v, 0u, comp);
  else odd_even_sort<VT, has_values, true>(k, v, oor, comp);
#pragma unroll 1
  for (int pass = 0; (1 << pass) < NT; ++pass) {
    if ((VT << pass) >= tile_count) break;  // one sorted run covers the tile
    block_merge_pass<NT, VT, has_values>(k, v, tid, tile_count, pass, comp, s);
  }

#pragma unroll
  for (int i = 0; i < VT; ++i) s.keys[VT * tid + i] = k[i];
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[VT * tid + i] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const int j = i * NT + tid;
    const key_t x = s.keys[j];
    if (j < tile_count) keys_out[j] = x;
  }
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int j = i * NT + tid;
      const val_t x = s.vals[j];
      if (j < tile_count) vals_out[j] = x;
    }
  }
}

// common/kernel_mergesort.hxx:12-67.
template <int NT, int VT, bool sort_indices, typename key_t, typename val_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_blocksort(const key_t* keys_in, const val_t* vals_in,
                                                       key_t* keys_out, val_t* vals_out,
                                                       int64_t count, comp_t comp) {
  constexpr int NV = NT * VT;
  static_assert(is_pow2(NT), "NT must be a power of two");
  static_assert(VT < 32, "VT must fit the out-of-range mask");
  __shared__ blocksort_storage_t<NT, VT, key_t, val_t> s;
  const int64_t base = (int64_t)blockIdx.x * NV;
  const int tile_count = (int)min_<int64_t>(NV, count - base);
  cta_blocksort<NT, VT, sort_indices>(keys_in + base, vals_in ? vals_in + base : vals_in,
                                      keys_out + base, vals_out ? vals_out + base : vals_out,
                                      tile_count, base, comp, s);
}

// Partition of one mergesort pass (vk/mergesort.hxx:65-72); one per thread.
template <int NT, typename key_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_mergesort_partition(
    const key_t* keys, int64_t count, int64_t coop, int64_t spacing, int64_t num_partitions,
    int64_t* mp, comp_t comp) {
  int64_t index = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (index >= num_partitions) return;
  merge_range_t range = compute_mergesort_range(count, index, coop, spacing);
  int64_t diag = min_(spacing * index, count) - range.a_begin;
  mp[index] = merge_path_global(keys + range.a_begin, range.a_count(), keys + range.b_begin,
                                range.b_count(), diag, comp);
}

// common/kernel_mergesort.hxx:72-121.
template <int NT, int VT, typename key_t, typename val_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_mergesort_pass(
    const int64_t* mp, const key_t* keys_in, const val_t* vals_in, key_t* keys_out,
    val_t* vals_out, int64_t count, int64_t coop, comp_t comp) {
  constexpr int NV = NT * VT;
  __shared__ merge_storage_t<NT, VT, key_t, val_t> s;
  const int64_t cta = blockIdx.x;
  merge_range_t range = compute_mergesort_range(count, cta, coop, NV, mp[cta], mp[cta + 1]);
  cta_merge_tile<NT, VT>(keys_in, vals_in, range.a_begin, (int)range.a_count(), keys_in, vals_in,
                         range.b_begin, (int)range.b_count(), keys_out, vals_out, cta * NV,
                         (int)min_<int64_t>(NV, count - cta * NV), comp, s);
}

// Bytes of aux a mergesort call needs (vk/mergesort.hxx:35-41).
template <int NT, int VT, typename key_t, typename val_t>
inline size_t mergesort_aux_bytes(int64_t count) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr int NV = NT * VT;
  int64_t num_ctas = div_up(count, (int64_t)NV);
  if (find_log2(num_ctas, true) == 0) return 0;
  size_t need = aux_bytes<int64_t>(num_ctas + 1) + aux_bytes<key_t>(count);
  if (has_values) need += aux_bytes<val_t>(count);
  return need;
}

template <int NT, int VT, bool sort_indices, typename key_t, typename val_t, typename comp_t>
mgpu_status_t mergesort_kv_impl(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                                val_t* vals, int64_t count, comp_t comp) {
  static_assert(!sort_indices || std::is_same_v<val_t, int32_t>, "indices are int32");
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr int NV = NT * VT;
  if (count < 0) return MGPU_INVALID_ARG;
  if (count == 0) return MGPU_OK;
  if (count >= ((int64_t)1 << 31)) return MGPU_UNSUPPORTED;
  const int64_t num_ctas = div_up(count, (int64_t)NV);
  const int num_passes = find_log2(num_ctas, true);
  const size_t need = mergesort_aux_bytes<NT, VT, key_t, val_t>(count);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if (!keys || (has_values && !vals)) return MGPU_INVALID_ARG;
  if (mgpu_status_t st = check_aux(aux_data, aux_size, need)) return st;
  hipError_t e;
  if (num_passes == 0) {
    e = launch_kernel("mgpu_blocksort", kernel_blocksort<NT, VT, sort_indices, key_t, val_t, comp_t>,
                      dim3(1), dim3(NT), 0, stream, (const key_t*)keys, (const val_t*)vals, keys,
                      vals, count, comp);
    return hip_status(e);
  }
  void* p = aux_data;
  const int64_t num_partitions = num_ctas + 1;
  int64_t* mp = advance_pointer<int64_t>(p, num_partitions);
  key_t* keys2 = advance_pointer<key_t>(p, count);
  val_t* vals2 = has_values ? advance_pointer<val_t>(p, count) : nullptr;
  key_t* keys_bs = (1 & num_passes) ? keys2 : keys;
  val_t* vals_bs = (1 & num_passes) ? vals2 : vals;
  e = launch_kernel("mgpu_blocksort", kernel_blocksort<NT, VT, sort_indices, key_t, val_t, comp_t>,
                    dim3((unsigned)num_ctas), dim3(NT), 0, stream, (const key_t*)keys,
                    (const val_t*)vals, keys_bs, vals_bs, count, comp);
  if (e != hipSuccess) return hip_status(e);
  key_t* k_in = keys_bs;
  val_t* v_in = vals_bs;
  key_t* k_out = (1 & num_passes) ? keys : keys2;
  val_t* v_out = (1 & num_passes) ? vals : vals2;
  constexpr int PNT = 256;
  for (int pass = 0; pass < num_passes; ++pass) {
    int64_t coop = (int64_t)2 << pass;
    e = launch_kernel("mgpu_mergesort_partition", kernel_mergesort_partition<PNT, key_t, comp_t>,
                      dim3((unsigned)div_up(num_partitions, (int64_t)PNT)),
                      dim3(PNT), 0, stream, (const key_t*)k_in, count, coop, (int64_t)NV,
                      num_partitions, mp, comp);
    if (e != hipSuccess) return hip_status(e);
    e = launch_kernel("mgpu_mergesort_pass", kernel_mergesort_pass<NT, VT, key_t, val_t, comp_t>,
                      dim3((unsigned)num_ctas), dim3(NT), 0, stream, (const int64_t*)mp,
                      (const key_t*)k_in, (const val_t*)v_in, k_out, v_out, count, coop, comp);
    if (e != hipSuccess) return hip_status(e);
    std::swap(k_in, k_out);
    std::swap(v_in, v_out);
  }
  return MGPU_OK;
}

// nt and vt only tune: a shape the blocksort cannot run (NT not a power of two
// up to 1024, VT beyond the 31-slot out-of-range mask, LDS beyond 160 KB) runs
// as 256 x 15; the sorted result does not depend on the shape.  The default
// 512 x 19 is the tile measured fastest for 32-bit keys on MI355X.
template <int NT = 512, int VT = 19, bool sort_indices = false, typename key_t, typename val_t,
          typename comp_t = less_t<key_t>>
mgpu_status_t mergesort_kv(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                           val_t* vals, int64_t count, comp_t comp = comp_t()) {
  if constexpr (NT >= warp_size && NT <= 1024 && is_pow2(NT) && VT >= 1 && VT < 32 &&
                sizeof(blocksort_storage_t<NT, VT, key_t, val_t>) <= 160 * 1024 &&
                sizeof(merge_storage_t<NT, VT, key_t, val_t>) <= 160 * 1024)
    return mergesort_kv_impl<NT, VT, sort_indices>(aux_data, aux_size, stream, keys, vals, count,
                                                   comp);
  else
    return mergesort_kv_impl<256, 15, sort_indices>(aux_data, aux_size, stream, keys, vals, count,
                                                    comp);
}

template <int NT = 512, int VT = 19, typename key_t, typename comp_t = less_t<key_t>>
mgpu_status_t mergesort_keys(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                             int64_t count, comp_t comp = comp_t()) {
  return mergesort_kv<NT, VT, false>(aux_data, aux_size, stream, keys, (empty_t*)nullptr, count,
                                     comp);
}

template <int NT = 512, int VT = 19, typename key_t, typename comp_t = less_t<key_t>>
mgpu_status_t mergesort_indices(void* aux_data, size_t& aux_size, hipStream_t stream,
                                key_t* keys, int32_t* indices, int64_t count,
                                comp_t comp = comp_t()) {
  return mergesort_kv<NT, VT, true>(aux_data, aux_size, stream, keys, indices, count, comp);
}

}  // namespace hip
}  // namespace mgpu
