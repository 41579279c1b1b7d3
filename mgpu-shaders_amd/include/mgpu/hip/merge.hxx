// Merge-path primitives and the two-way stable merge for gfx950.
//
// merge_path follows inc/mgpu/common/cta_merge.hxx:10-31 with bounds_lower
// (ties go to A): it returns the number of A elements among the first `diag`
// outputs of the stable merge.  The partition search is one binary search
// per thread in global memory, as in vk/partition.hxx:9-20 (a wave-wide
// 64-probe search was measured: fewer dependent rounds, but four times the
// memory traffic, and slower).
//
// A tile merge (kernel_merge.hxx:9-58 and cta_merge.hxx:136-162) stages the A
// and B slices of one output tile in LDS with coalesced loads, runs one merge
// path per thread and a vt-step serial merge (cta_merge.hxx:110-133), and
// stores keys and values through LDS; values are gathered from LDS rather than
// from global memory (the reference gathers from global, cta_merge.hxx:84-105).
#pragma once
#include "meta.hxx"

namespace mgpu {
namespace hip {

// Which input wins ties on a merge-path diagonal (common/cta_merge.hxx:18-31):
// bounds_lower puts A's equal keys first (the stable merge of A then B);
// bounds_upper puts B's first.
enum bounds_t { bounds_lower, bounds_upper };

// Number of A elements among the first `diag` outputs of the merge of sorted
// a and b whose ties follow `bounds`: a binary search over the diagonal.
template <bounds_t bounds = bounds_lower, typename index_t, typename key_t, typename comp_t>
__host__ __device__ inline index_t merge_path_search(const key_t* a, index_t a_count,
                                                     const key_t* b, index_t b_count,
                                                     index_t diag, comp_t comp) {
  index_t begin = diag > b_count ? diag - b_count : 0;
  index_t end = diag < a_count ? diag : a_count;
  while (begin < end) {
    const index_t mid = (begin + end) >> 1;
    const bool pred = bounds == bounds_upper ? comp(a[mid], b[diag - 1 - mid])
                                             : !comp(b[diag - 1 - mid], a[mid]);
    if (pred) begin = mid + 1;
    else end = mid;
  }
  return begin;
}

// Per-thread merge path (binary search), used in LDS.
template <typename key_t, typename comp_t>
__device__ inline int merge_path_lds(const key_t* a, int a_count, const key_t* b,
                                     int b_count, int diag, comp_t comp) {
  return merge_path_search<bounds_lower, int>(a, a_count, b, b_count, diag, comp);
}

// Merge path in global memory, one binary search per thread (the upper levels
// of every search touch the same few cache lines, which stay in L2).
template <bounds_t bounds = bounds_lower, typename key_t, typename comp_t>
__device__ inline int64_t merge_path_global(const key_t* a, int64_t a_count, const key_t* b,
                                            int64_t b_count, int64_t diag, comp_t comp) {
  return merge_path_search<bounds, int64_t>(a, a_count, b, b_count, diag, comp);
}

// Serial merge of vt outputs starting at (a_begin, b_begin) of LDS keys, with
// the range checks of merge_predicate (cta_merge.hxx:41-50).  keys_shared
// must be readable one element past the ranges.  With checked = false the
// caller guarantees that neither range runs out within the VT steps (both
// hold at least VT keys), and the range tests are dropped.
template <int VT, bool checked = true, typename key_t, typename comp_t>
__device__ inline void serial_merge(const key_t* keys_shared, int a_begin, int a_end,
                                    int b_begin, int b_end, key_t (&keys)[VT],
                                    int (&indices)[VT], comp_t comp) {
  key_t a_key = keys_shared[a_begin];
  key_t b_key = keys_shared[b_begin];
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const bool p = checked ? (a_begin < a_end && (b_begin >= b_end || !comp(b_key, a_key)))
                           : !comp(b_key, a_key);
    const int index = p ? a_begin : b_begin;
    keys[i] = p ? a_key : b_key;
    indices[i] = index;
    const key_t c_key = keys_shared[index + 1];
    a_begin = p ? index + 1 : a_begin;
    b_begin = p ? b_begin : index + 1;
    a_key = p ? c_key : a_key;
    b_key = p ? b_key : c_key;
  }
}

// serial_merge with the range tests dropped when every lane of the wave has
// at least VT keys left in both ranges (a wave-uniform choice, so the wave
// runs one of the two loops, never both).
template <int VT, typename key_t, typename comp_t>
__device__ inline void serial_merge_wave(const key_t* keys_shared, int a_begin, int a_end,
                                         int b_begin, int b_end, key_t (&keys)[VT],
                                         int (&indices)[VT], comp_t comp) {
  const bool roomy = a_end - a_begin >= VT && b_end - b_begin >= VT;
  if (__all(roomy))
    serial_merge<VT, false>(keys_shared, a_begin, a_end, b_begin, b_end, keys, indices, comp);
  else
    serial_merge<VT, true>(keys_shared, a_begin, a_end, b_begin, b_end, keys, indices, comp);
}

// Partition kernel: mp[i] = merge_path(a, b, min(i * spacing, a_count + b_count))
// (vk/partition.hxx:9-20); mp_t is int64_t here, int for the reference's
// int partition arrays.
template <int NT, bounds_t bounds = bounds_lower, typename key_t, typename mp_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_merge_partition(
    const key_t* a, int64_t a_count, const key_t* b, int64_t b_count, int64_t spacing,
    int64_t num_partitions, mp_t* mp, comp_t comp) {
  int64_t index = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (index >= num_partitions) return;
  int64_t diag = min_<int64_t>(spacing * index, a_count + b_count);
  mp[index] = (mp_t)merge_path_global<bounds>(a, a_count, b, b_count, diag, comp);
}

// Partitions of a merge into `spacing`-output tiles: the reference's
// num_merge_partitions (common/cta_merge.hxx:6-8).
constexpr int64_t num_merge_partitions(int64_t count, int64_t spacing) {
  return (count + spacing - 1) / spacing + 1;
}

// vk::merge_path_partitions<bounds>(cmd, a, a_count, b, b_count, mp_data,
// spacing, comp) (vk/partition.hxx:9-20) with the stream in place of the
// command buffer: mp_data[i] = the merge path of diagonal
// min(i * spacing, a_count + b_count) for i < num_merge_partitions(a_count +
// b_count, spacing).  No aux.  mp_data is int (as in the reference: totals
// below 2^31) or int64_t.
template <bounds_t bounds, typename key_t, typename mp_t, typename comp_t>
mgpu_status_t merge_path_partitions(hipStream_t stream, const key_t* a, int64_t a_count,
                                    const key_t* b, int64_t b_count, mp_t* mp_data,
                                    int64_t spacing, comp_t comp) {
  static_assert(std::is_same_v<mp_t, int> || std::is_same_v<mp_t, int64_t>,
                "partitions are int or int64_t");
  if (a_count < 0 || b_count < 0 || spacing <= 0) return MGPU_INVALID_ARG;
  if (std::is_same_v<mp_t, int> && a_count + b_count >= ((int64_t)1 << 31))
    return MGPU_UNSUPPORTED;
  if (!mp_data || (a_count && !a) || (b_count && !b)) return MGPU_INVALID_ARG;
  const int64_t num_partitions = num_merge_partitions(a_count + b_count, spacing);
  constexpr int PNT = 256;
  return hip_status(launch_kernel("mgpu_merge_partition",
                                  kernel_merge_partition<PNT, bounds, key_t, mp_t, comp_t>,
                                  dim3((unsigned)div_up(num_partitions, (int64_t)PNT)), dim3(PNT),
                                  0, stream, a, a_count, b, b_count, spacing, num_partitions,
                                  mp_data, comp));
}

// LDS footprint of a tile merge.
template <int NT, int VT, typename key_t, typename val_t>
struct merge_storage_t {
  static constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  key_t keys[NT * VT + 1];
  typename std::conditional<has_values, val_t, char>::type vals[has_values ? NT * VT + 1 : 1];
};

// Merge A[a_begin, a_begin+a_cnt) and B[b_begin, b_begin+b_cnt) (a_cnt+b_cnt <=
// NT*VT) into C[c_begin, ...), keys and values.  Called by the whole block.
template <int NT, int VT, typename key_t, typename val_t, typename comp_t>
__device__ inline void cta_merge_tile(
    const key_t* a_keys, const val_t* a_vals, int64_t a_begin, int a_cnt,
    const key_t* b_keys, const val_t* b_vals, int64_t b_begin, int b_cnt,
    key_t* c_keys, val_t* c_vals, int64_t c_begin, int c_count, comp_t comp,
    merge_storage_t<NT, VT, key_t, val_t>& s) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  const int tid = threadIdx.x;
  // a_cnt + b_cnt == c_count <= NT*VT for valid partitions; the clamps keep a
  // corrupt partition from reaching outside the tile.
  const int total = min_(min_(a_cnt + b_cnt, c_count), NT * VT);
  a_cnt = min_(a_cnt, total);
  b_cnt = total - a_cnt;

  if (total <= 0) return;

  // Load both streams into LDS (coalesced strided loads).  Each load reads a
  // clamped index from A or B chosen by select, with no branch, so that all
  // loads are in flight together (a guarded load costs one wait each).
  {
    key_t x[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int j = min_(i * NT + tid, total - 1);
      const bool in_a = j < a_cnt;
      x[i] = (in_a ? a_keys : b_keys)[in_a ? a_begin + j : b_begin + (j - a_cnt)];
    }
#pragma unroll
    for (int i = 0; i < VT; ++i) s.keys[i * NT + tid] = x[i];
  }
  if constexpr (has_values) {
    val_t x[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int j = min_(i * NT + tid, total - 1);
      const bool in_a = j < a_cnt;
      x[i] = (in_a ? a_vals : b_vals)[in_a ? a_begin + j : b_begin + (j - a_cnt)];
    }
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[i * NT + tid] = x[i];
  }
  __syncthreads();

  int diag = min_(VT * tid, total);
  int mp = merge_path_lds(s.keys, a_cnt, s.keys + a_cnt, b_cnt, diag, comp);
  key_t keys[VT];
  int indices[VT];
  serial_merge_wave<VT>(s.keys, mp, a_cnt, a_cnt + diag - mp, total, keys, indices, comp);
  val_t vals[VT];
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) vals[i] = s.vals[indices[i]];
  }
  __syncthreads();

  // Thread order -> LDS -> strided coalesced stores.
#pragma unroll
  for (int i = 0; i < VT; ++i) s.keys[VT * tid + i] = keys[i];
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[VT * tid + i] = vals[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const int j = i * NT + tid;
    const key_t x = s.keys[j];
    if (j < total) c_keys[c_begin + j] = x;
  }
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int j = i * NT + tid;
      const val_t x = s.vals[j];
      if (j < total) c_vals[c_begin + j] = x;
    }
  }
}

template <int NT, int VT, typename key_t, typename val_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_merge(
    const int64_t* mp, const key_t* a_keys, const val_t* a_vals, int64_t a_count,
    const key_t* b_keys, const val_t* b_vals, int64_t b_count, key_t* c_keys, val_t* c_vals,
    comp_t comp) {
  constexpr int NV = NT * VT;
  __shared__ merge_storage_t<NT, VT, key_t, val_t> s;
  const int64_t cta = blockIdx.x;
  const int64_t diag0 = cta * NV;
  const int64_t diag1 = min_<int64_t>(a_count + b_count, diag0 + NV);
  const int64_t mp0 = mp[cta], mp1 = mp[cta + 1];
  cta_merge_tile<NT, VT>(a_keys, a_vals, mp0, (int)(mp1 - mp0), b_keys, b_vals, diag0 - mp0,
                         (int)((diag1 - mp1) - (diag0 - mp0)), c_keys, c_vals, diag0,
                         (int)(diag1 - diag0), comp, s);
}

// Stable merge of (a_keys, a_vals) and (b_keys, b_vals) into (c_keys, c_vals);
// val_t = empty_t for key-only.  inc/mgpu/vk/merge.hxx:12-59.  The default
// tile is the one measured fastest for 32-bit keys on MI355X (512 x 23); VT
// is odd because the thread-order LDS transposes read word VT * tid + i,
// which a power-of-two VT puts in few banks.
template <int NT = 512, int VT = 23, typename key_t, typename val_t, typename comp_t>
mgpu_status_t merge(void* aux_data, size_t& aux_size, hipStream_t stream,
                    const key_t* a_keys, const val_t* a_vals, int64_t a_count,
                    const key_t* b_keys, const val_t* b_vals, int64_t b_count,
                    key_t* c_keys, val_t* c_vals, comp_t comp) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr int NV = NT * VT;
  if (a_count < 0 || b_count < 0) return MGPU_INVALID_ARG;
  const int64_t total = a_count + b_count;
  if (total == 0) return MGPU_OK;
  const int64_t num_tiles = div_up(total, (int64_t)NV);
  if (num_tiles >= ((int64_t)1 << 31)) return MGPU_UNSUPPORTED;
  const int64_t num_partitions = num_tiles + 1;
  size_t need = aux_bytes<int64_t>(num_partitions);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if ((a_count && !a_keys) || (b_count && !b_keys) || !c_keys) return MGPU_INVALID_ARG;
  if (has_values && ((a_count && !a_vals) || (b_count && !b_vals) || !c_vals))
    return MGPU_INVALID_ARG;
  if (mgpu_status_t st = check_aux(aux_data, aux_size, need)) return st;
  void* p = aux_data;
  int64_t* mp = advance_pointer<int64_t>(p, num_partitions);
  constexpr int PNT = 256;
  hipError_t e = launch_kernel("mgpu_merge_partition",
                               kernel_merge_partition<PNT, bounds_lower, key_t, int64_t, comp_t>,
                               dim3((unsigned)div_up(num_partitions, (int64_t)PNT)),
                               dim3(PNT), 0, stream, a_keys, a_count, b_keys, b_count,
                               (int64_t)NV, num_partitions, mp, comp);
  if (e != hipSuccess) return hip_status(e);
  e = launch_kernel("mgpu_merge", kernel_merge<NT, VT, key_t, val_t, comp_t>,
                    dim3((unsigned)num_tiles), dim3(NT), 0, stream, (const int64_t*)mp, a_keys,
                    a_vals, a_count, b_keys, b_vals, b_count, c_keys, c_vals, comp);
  return hip_status(e);
}

// Key-only merge (inc/mgpu/vk/merge.hxx:49-59; an explicit comparator).
template <int NT = 512, int VT = 23, typename key_t, typename comp_t>
mgpu_status_t merge(void* aux_data, size_t& aux_size, hipStream_t stream,
                    const key_t* a_keys, int64_t a_count, const key_t* b_keys,
                    int64_t b_count, key_t* c_keys, comp_t comp) {
  return merge<NT, VT>(aux_data, aux_size, stream, a_keys, (const empty_t*)nullptr, a_count,
                       b_keys, (const empty_t*)nullptr, b_count, c_keys, (empty_t*)nullptr,
                       comp);
}

}  // namespace hip
}  // namespace mgpu
