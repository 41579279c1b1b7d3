// Merge-path mergesort for gfx950.
//
// Structure as in the reference (inc/mgpu/vk/mergesort.hxx:10-84):
//   1. blocksort: each nv-tile is sorted by a stable odd-even transposition
//      network over vt registers (common/sort_networks.hxx:11-39, with the
//      partial-tile mask of common/cta_mergesort.hxx:7-15) and log2(nt) merge
//      passes in LDS (common/cta_mergesort.hxx:87-133);
//   2. ceil(log2(tiles)) levels of merging.  The reference runs one global
//      pass per level: partition the coop groups of 2<<pass tiles with merge
//      path (vk/mergesort.hxx:65-72), then merge one output tile per block
//      (common/kernel_mergesort.hxx:72-121).  Large sorts here merge two
//      levels per pass (four runs into one, below), so 2^28 pairs take 8
//      global passes instead of 15; small ones keep one level per pass.
// The ping-pong parity lands the result in the caller's buffers
// (vk/mergesort.hxx:47-48).  Every comparison is strict (swap only when
// comp(b, a)), and merges take the earlier input on ties, so the sort is
// stable and its output equals std::stable_sort.
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

// Loads of one tile (tile_count >= 1) into registers in strided order (slot
// i of thread tid is element i * NT + tid).  Slots past the end hold 0 or a
// copy of the last element; the out-of-range mask keeps them out of the sort.
// All loads are in flight together.
template <int NT, int VT, bool load_values, typename key_t, typename val_t>
__device__ inline void blocksort_load(const key_t* keys_in, const val_t* vals_in, int tile_count,
                                      key_t (&kx)[VT], val_t (&vx)[VT]) {
  tile_load_strided<NT, VT>(keys_in, tile_count, kx);
  if constexpr (load_values) tile_load_strided<NT, VT>(vals_in, tile_count, vx);
}

// Writes a tile loaded by blocksort_load to LDS, in the same strided order.
template <int NT, int VT, bool load_values, typename key_t, typename val_t>
__device__ inline void blocksort_stage(const key_t (&kx)[VT], const val_t (&vx)[VT],
                                       blocksort_storage_t<NT, VT, key_t, val_t>& s) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < VT; ++i) s.keys[i * NT + tid] = kx[i];
  if constexpr (load_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[i * NT + tid] = vx[i];
  }
}

// The blocksort of a tile of tile_count (1 <= tile_count <= NT*VT) keys (and
// values) that blocksort_stage has written to LDS: sort, then write the tile
// out.  With sort_indices the values are index_base + position
// (common/kernel_mergesort.hxx:43-49) and none were staged.
template <int NT, int VT, bool sort_indices, typename key_t, typename val_t, typename comp_t>
__device__ inline void cta_blocksort_staged(key_t* keys_out, val_t* vals_out, int tile_count,
                                            int64_t index_base, comp_t comp,
                                            blocksort_storage_t<NT, VT, key_t, val_t>& s) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr int NV = NT * VT;
  const int tid = threadIdx.x;
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
  {
    key_t x[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i) x[i] = s.keys[i * NT + tid];
    tile_store_strided<NT, VT>(keys_out, tile_count, x);
  }
  if constexpr (has_values) {
    val_t x[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i) x[i] = s.vals[i * NT + tid];
    tile_store_strided<NT, VT>(vals_out, tile_count, x);
  }
}

// Sort tile_count (<= NT*VT) keys (and values) of one tile in LDS and write
// them out: the body of the blocksort kernel (common/kernel_mergesort.hxx:
// 12-67).  keys_in may equal keys_out (the tile is fully read before it is
// written).
template <int NT, int VT, bool sort_indices, typename key_t, typename val_t, typename comp_t>
__device__ inline void cta_blocksort(const key_t* keys_in, const val_t* vals_in, key_t* keys_out,
                                     val_t* vals_out, int tile_count, int64_t index_base,
                                     comp_t comp, blocksort_storage_t<NT, VT, key_t, val_t>& s) {
  constexpr bool stage_values = !std::is_same_v<val_t, empty_t> && !sort_indices;
  if (tile_count <= 0) return;
  key_t kx[VT];
  val_t vx[VT];
  blocksort_load<NT, VT, stage_values>(keys_in, vals_in, tile_count, kx, vx);
  blocksort_stage<NT, VT, stage_values>(kx, vx, s);
  cta_blocksort_staged<NT, VT, sort_indices>(keys_out, vals_out, tile_count, index_base, comp, s);
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

////////////////////////////////////////////////////////////////////////////////
// Four-way passes: one pass merges groups of four runs of `run` keys into one
// run of 4 * run, so a sort takes ceil(levels / 2) global passes instead of
// `levels` (2^28 pairs in 9728-key tiles: 8 passes instead of 15).  Each
// output tile's slices of the four runs come from merge4_path; the tile is
// merged in LDS in two rounds, (r0, r1) and (r2, r3) in the first, the two
// results in the second.  Stability: merge(merge(r0, r1), merge(r2, r3)) with
// every tie to the earlier input is the stable merge of r0 r1 r2 r3.

// The partitions of a four-way pass come from two kernels, so that no search
// nests a full-length binary search inside another:
//  1. samples: every S-th key of X = merge(r0, r1) and of Y = merge(r2, r3)
//     (S = NV / merge4_sample_div) with its merge path in its pair and its
//     position in the group's output, X[jS] at jS + (keys of Y before it),
//     each a plain binary search;
//  2. partitions: per tile boundary, the samples before it (one search over
//     each sample array) fix the answer of merge4_search to one interval of
//     S positions and its inner merge paths to the keys between two samples.
// One sample per tile (S = NV) measured best: 2^28 pairs, samples +
// partitions 0.18 ms per pass at S = NV, 0.22 at NV / 4, 0.49 at NV / 16;
// without samples (merge4_path alone, one nested search per tile) 0.39 ms.
constexpr int merge4_sample_div = 1;

// Four-way passes pay off once a pass moves enough bytes to outweigh the
// extra partition kernels.  Keys only: per-pass keys >= 256 MiB (2^25 u32
// keys, 134 MB: 0.78 ms two-way vs 1.03 four-way; 201 MB: 1.34 vs 1.47; 2^28
// keys: 7.4 vs 6.0).  Keys and values: >= 168 MiB per pass (u32 pairs,
// 168 MB: 1.12 vs 1.14 ms, 201 MB: 1.32 vs 1.26, 2^28 pairs 13.5 vs 9.2; u64
// key + u32 value pairs, 176 MB: 1.14 vs 1.09, 2^27 pairs 9.75 vs 6.5).
constexpr int64_t merge4_min_bytes = (int64_t)1 << 28;
constexpr int64_t merge4_min_bytes_kv = (int64_t)168 << 20;

// Pass structure of the merge sort (a process-wide setting): 0 = automatic
// (four-way passes once a pass moves merge4_min_bytes of keys, or
// merge4_min_bytes_kv of keys and values), 2 = two-way passes only (the
// reference's structure), 4 = four-way passes at every size.  The
// sorted result is the same; the sizing call covers every structure.
inline std::atomic<int>& mergesort_passes_setting() {
  static std::atomic<int> a{0};
  return a;
}

template <typename key_t, typename val_t>
inline bool mergesort_four_way(int64_t count) {
  const int setting = mergesort_passes_setting().load(std::memory_order_relaxed);
  if (setting == 2 || setting == 4) return setting == 4;
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  return count * (int64_t)(sizeof(key_t) + (has_values ? sizeof(val_t) : 0)) >=
         (has_values ? merge4_min_bytes_kv : merge4_min_bytes);
}

// Samples per side (X or Y) of one group of four runs of `run` keys.
__host__ __device__ inline int64_t merge4_samples_per_group(int64_t run, int64_t spacing) {
  return 2 * run / spacing;
}

// Sample t of a four-way pass: t < per_side is position jS of X = merge(r0,
// r1) of group g (t = g * SG + j), t >= per_side the same of Y = merge(r2,
// r3).  Returns (its merge path within its pair, its position in the group's
// output), or (0, INT32_MAX) past the end of X or Y.
template <typename key_t, typename comp_t>
__host__ __device__ inline int2 merge4_sample(const key_t* keys, int64_t count, int64_t run,
                                              int64_t spacing, int64_t per_side, int64_t t,
                                              comp_t comp) {
  const bool y_side = t >= per_side;
  const int64_t u = y_side ? t - per_side : t;
  const int64_t sg = merge4_samples_per_group(run, spacing);
  const int64_t g0 = u / sg * 4 * run, pos = u % sg * spacing;
  int64_t rb[5];
  for (int r = 0; r < 5; ++r) rb[r] = min_(count, g0 + r * run);
  // This side's pair (p, q) and the other side's pair (o1, o2).
  const key_t* p = keys + (y_side ? rb[2] : rb[0]);
  const key_t* q = keys + (y_side ? rb[3] : rb[1]);
  const int64_t np = y_side ? rb[3] - rb[2] : rb[1] - rb[0];
  const int64_t nq = y_side ? rb[4] - rb[3] : rb[2] - rb[1];
  const key_t* o1 = keys + (y_side ? rb[0] : rb[2]);
  const key_t* o2 = keys + (y_side ? rb[1] : rb[3]);
  const int64_t no1 = y_side ? rb[1] - rb[0] : rb[3] - rb[2];
  const int64_t no2 = y_side ? rb[2] - rb[1] : rb[4] - rb[3];
  if (pos >= np + nq) return make_int2(0, INT32_MAX);
  const int64_t a = kary_search(max_<int64_t>(0, pos - nq), min_(pos, np),
                                [&](int64_t i) { return !comp(q[pos - 1 - i], p[i]); });
  const bool from_p = a < np && (pos - a >= nq || !comp(q[pos - a], p[a]));
  const key_t x = from_p ? p[a] : q[pos - a];
  // Keys of the other pair before x (Y's keys come after equal X keys); the
  // two searches run interleaved.
  int64_t b1 = 0, b2 = 0;
  kary_search2(
      b1, no1, [&](int64_t i) { return y_side ? !comp(x, o1[i]) : comp(o1[i], x); }, b2, no2,
      [&](int64_t i) { return y_side ? !comp(x, o2[i]) : comp(o2[i], x); });
  return make_int2((int)a, (int)(pos + b1 + b2));
}

// Number of samples (ranks ascending, INT32_MAX past the end) ranked below d.
__host__ __device__ inline int64_t merge4_samples_below(const int2* s, int64_t n, int64_t d) {
  int64_t begin = 0, end = n;
  while (begin < end) {
    const int64_t mid = (begin + end) >> 1;
    if ((int64_t)s[mid].y < d) begin = mid + 1;
    else end = mid;
  }
  return begin;
}

// Partition c of a four-way pass: the slices of the four runs of the group
// of position min(c * spacing, count) that precede it.
template <typename key_t, typename comp_t>
__host__ __device__ inline int4 merge4_partition(const key_t* keys, int64_t count, int64_t run,
                                                 int64_t spacing, int64_t c, const int2* samples,
                                                 int64_t per_side, int64_t sample_spacing,
                                                 comp_t comp) {
  const int64_t pos = min_(c * spacing, count);
  const int64_t g = pos / (4 * run), g0 = g * 4 * run;
  int64_t rb[5];
  for (int r = 0; r < 5; ++r) rb[r] = min_(count, g0 + r * run);
  const int64_t n0 = rb[1] - rb[0], n1 = rb[2] - rb[1], n2 = rb[3] - rb[2], n3 = rb[4] - rb[3];
  const int64_t d = pos - g0, nx = n0 + n1, ny = n2 + n3;
  int64_t split[4] = {0, 0, 0, 0};
  if (d > 0) {
    const int64_t sg = merge4_samples_per_group(run, sample_spacing), S = sample_spacing;
    const int2* xs = samples + g * sg;
    const int2* ys = samples + per_side + g * sg;
    const int64_t J = merge4_samples_below(xs, sg, d), K = merge4_samples_below(ys, sg, d);
    const int64_t nvx = div_up(nx, S), nvy = div_up(ny, S);
    merge4_state_t st;
    st.xlm = J ? (J - 1) * S : 0;
    st.xla = J ? xs[J - 1].x : 0;
    st.xhm = J < nvx ? J * S : nx;
    st.xha = J < nvx ? xs[J].x : n0;
    st.yln = K ? (K - 1) * S : 0;
    st.ylb = K ? ys[K - 1].x : 0;
    st.yhn = K < nvy ? K * S : ny;
    st.yhb = K < nvy ? ys[K].x : n2;
    // Sample J - 1 of X is among the first d outputs and sample J is not
    // (likewise K of Y): i is in (xlm, xhm] and d - i in (yln, yhn].
    st.lo = max_(max_<int64_t>(0, d - ny), max_(J ? st.xlm + 1 : (int64_t)0, d - st.yhn));
    st.hi = min_(min_(d, nx), min_(st.xhm, d - (K ? st.yln + 1 : (int64_t)0)));
    merge4_search(keys + rb[0], n0, keys + rb[1], n1, keys + rb[2], n2, keys + rb[3], n3, d, comp,
                  st, split);
  }
  return make_int4((int)split[0], (int)split[1], (int)split[2], (int)split[3]);
}

template <int NT, typename key_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_mergesort_samples4(
    const key_t* keys, int64_t count, int64_t run, int64_t spacing, int64_t per_side,
    int2* samples, comp_t comp) {
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t < 2 * per_side) samples[t] = merge4_sample(keys, count, run, spacing, per_side, t, comp);
}

// mp[c] = the slices of the four runs of tile c's group that precede tile c
// (offsets within each run).  One thread per tile.
template <int NT, typename key_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_mergesort_partition4(
    const key_t* keys, int64_t count, int64_t run, int64_t spacing, int64_t num_partitions,
    const int2* samples, int64_t per_side, int64_t sample_spacing, int4* mp, comp_t comp) {
  const int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (c < num_partitions)
    mp[c] = merge4_partition(keys, count, run, spacing, c, samples, per_side, sample_spacing,
                             comp);
}

// Merge path over A = p0 p2 and B = p1 p3 (held in LDS) in which every key of
// p0 and p1 (segment 0) precedes every key of p2 and p3 (segment 1): the
// merge of A and B is then merge(p0, p1) followed by merge(p2, p3).  a_split
// and b_split are the segment boundaries within A and B.
template <typename key_t, typename comp_t>
__device__ inline int merge_path_seg(const key_t* a, int a_count, int a_split, const key_t* b,
                                     int b_count, int b_split, int diag, comp_t comp) {
  int begin = max_(0, diag - b_count), end = min_(diag, a_count);
  while (begin < end) {
    const int mid = (begin + end) >> 1;
    const int j = diag - 1 - mid;
    const bool sa = mid >= a_split, sb = j >= b_split;
    // a[mid] precedes b[j] unless b[j] is in an earlier segment or, in the
    // same segment, strictly less.
    const bool pred = sa == sb ? !comp(b[j], a[mid]) : sa < sb;
    if (pred) begin = mid + 1;
    else end = mid;
  }
  return begin;
}

// serial_merge over the segmented order of merge_path_seg; indices are LDS
// positions (A in [a_begin, a_end) with segment boundary a_split, B in
// [b_begin, b_end) with boundary b_split).
template <int VT, typename key_t, typename comp_t>
__device__ inline void serial_merge_seg(const key_t* keys_shared, int a_begin, int a_end,
                                        int a_split, int b_begin, int b_end, int b_split,
                                        key_t (&keys)[VT], int (&indices)[VT], comp_t comp) {
  key_t a_key = keys_shared[a_begin];
  key_t b_key = keys_shared[b_begin];
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const bool sa = a_begin >= a_split, sb = b_begin >= b_split;
    const bool p = a_begin < a_end &&
                   (b_begin >= b_end || (sa == sb ? !comp(b_key, a_key) : sa < sb));
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

// Group offset of LDS slot j of a four-way tile (regions p0 p2 p1 p3 ending
// at q1, q2, q3, each with its own offset o).
__device__ inline uint32_t merge4_slot_offset(int j, int q1, int q2, int q3, int o0, int o2,
                                              int o1, int o3) {
  return (uint32_t)(j + (j < q1 ? o0 : j < q2 ? o2 : j < q3 ? o1 : o3));
}

// Waves per SIMD that a kernel of NT threads whose blocks hold `lds` bytes of
// LDS can keep resident (160 KB of LDS per CU, four SIMDs, at most eight
// waves each): the register budget its launch bounds ask for.
constexpr int lds_waves_per_simd(int nt, size_t lds) {
  const int blocks = (int)((160 * 1024) / lds);
  const int waves = blocks * nt / 256;
  return waves < 1 ? 1 : waves > 8 ? 8 : waves;
}

// One output tile of a four-way pass.  Tile c of group g (tiles of 4 * run
// keys) takes, from each run r of the group, the keys between mp[c].r and
// mp[c + 1].r (the run's end for the group's last tile).  The launch bounds
// hold its registers to what the LDS-limited occupancy allows (key-only
// 32-bit tiles: four blocks per CU, 64 VGPRs).
template <int NT, int VT, typename key_t, typename val_t, typename comp_t>
__global__ __launch_bounds__(NT, lds_waves_per_simd(NT, sizeof(merge_storage_t<NT, VT, key_t, val_t>)))
void kernel_mergesort_pass4(
    const int4* mp, const key_t* keys_in, const val_t* vals_in, key_t* keys_out,
    val_t* vals_out, int64_t count, int64_t run, comp_t comp) {
  constexpr int NV = NT * VT;
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  __shared__ merge_storage_t<NT, VT, key_t, val_t> s;
  const int tid = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * NV;
  const int64_t g0 = c0 / (4 * run) * (4 * run);
  int64_t rb[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) rb[r] = min_(count, g0 + r * run);
  const int4 lo = mp[blockIdx.x];
  const int4 hi = c0 + NV >= rb[4]
                      ? make_int4((int)(rb[1] - rb[0]), (int)(rb[2] - rb[1]),
                                  (int)(rb[3] - rb[2]), (int)(rb[4] - rb[3]))
                      : mp[blockIdx.x + 1];
  // Slice sizes, clamped so that a corrupt partition cannot leave the tile.
  const int tile = (int)min_<int64_t>(NV, count - c0);
  const int q0 = min_(max_(hi.x - lo.x, 0), tile);
  const int q2 = min_(max_(hi.z - lo.z, 0), tile - q0);
  const int q1 = min_(max_(hi.y - lo.y, 0), tile - q0 - q2);
  const int q3 = min_(max_(hi.w - lo.w, 0), tile - q0 - q2 - q1);
  const int total = q0 + q1 + q2 + q3;
  if (total <= 0) return;
  // LDS order p0 p2 p1 p3: A = [0, Q2) (segment boundary Q1), B = [Q2, total)
  // (segment boundary Q3).  Slot j loads from group offset j + o[region of
  // j]: 32-bit offsets from the group's first key (a group holds fewer than
  // 2^31 keys), one per slot, shared by the key and value loads.
  const int Q1 = q0, Q2 = q0 + q2, Q3 = Q2 + q1;
  const int o0 = (int)(rb[0] - g0) + lo.x, o2 = (int)(rb[2] - g0) + lo.z - Q1,
            o1 = (int)(rb[1] - g0) + lo.y - Q2, o3 = (int)(rb[3] - g0) + lo.w - Q3;
  const key_t* kbase = keys_in + g0;
  key_t x[VT];
  if constexpr (has_values) {
    uint32_t off[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i)
      off[i] = merge4_slot_offset(min_(i * NT + tid, total - 1), Q1, Q2, Q3, o0, o2, o1, o3);
#pragma unroll
    for (int i = 0; i < VT; ++i) x[i] = kbase[off[i]];
#pragma unroll
    for (int i = 0; i < VT; ++i) s.keys[i * NT + tid] = x[i];
    const val_t* vbase = vals_in + g0;
    val_t y[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i) y[i] = vbase[off[i]];
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[i * NT + tid] = y[i];
  } else {
#pragma unroll
    for (int i = 0; i < VT; ++i)
      x[i] = kbase[merge4_slot_offset(min_(i * NT + tid, total - 1), Q1, Q2, Q3, o0, o2, o1, o3)];
#pragma unroll
    for (int i = 0; i < VT; ++i) s.keys[i * NT + tid] = x[i];
  }
  __syncthreads();

  const int diag = min_(VT * tid, total);
  key_t k[VT];
  val_t v[VT];
  int idx[VT];
  // Round 1: X = merge(p0, p1) at [0, q0 + q1), Y = merge(p2, p3) after it.
  // A thread whose outputs all lie in X (or all in Y) merges p0 with p1 (or
  // p2 with p3) directly; only a wave holding the thread that straddles the
  // two needs the segmented comparison.
  {
    const int nx = q0 + q1;
    const int m = merge_path_seg(s.keys, Q2, Q1, s.keys + Q2, total - Q2, q1, diag, comp);
    const bool straddle = diag < nx && diag + VT > nx;
    if (__any(straddle)) {
      serial_merge_seg<VT>(s.keys, m, Q2, Q1, Q2 + diag - m, total, Q3, k, idx, comp);
    } else {
      const bool in_x = diag < nx;
      serial_merge<VT>(s.keys, m, in_x ? Q1 : Q2, Q2 + diag - m, in_x ? Q3 : total, k, idx,
                       comp);
    }
    if constexpr (has_values) {
#pragma unroll
      for (int i = 0; i < VT; ++i) v[i] = s.vals[idx[i]];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < VT; ++i) s.keys[VT * tid + i] = k[i];
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[VT * tid + i] = v[i];
  }
  __syncthreads();
  // Round 2: merge(X, Y).
  {
    const int nx = q0 + q1;
    const int m = merge_path_lds(s.keys, nx, s.keys + nx, total - nx, diag, comp);
    serial_merge_wave<VT>(s.keys, m, nx, nx + diag - m, total, k, idx, comp);
    if constexpr (has_values) {
#pragma unroll
      for (int i = 0; i < VT; ++i) v[i] = s.vals[idx[i]];
    }
  }
  __syncthreads();
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
    if (j < total) keys_out[c0 + j] = x;
  }
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int j = i * NT + tid;
      const val_t x = s.vals[j];
      if (j < total) vals_out[c0 + j] = x;
    }
  }
}

// Samples of the largest four-way pass of a sort of `count` keys in tiles of
// nv (both sides, all groups), over both schedules: runs that start at a
// tile, and runs that start at half a tile (mergesort_half_tiles).
inline int64_t mergesort_max_samples(int64_t count, int64_t nv) {
  int64_t most = 0;
  for (const int64_t run0 : {nv, nv / 2}) {
    const int levels = find_log2(div_up(count, run0), true);
    for (int level = levels & 1; level < levels; level += 2) {
      const int64_t run = run0 << level;
      const int64_t per_side =
          div_up(count, 4 * run) * merge4_samples_per_group(run, nv / merge4_sample_div);
      most = max_(most, 2 * per_side);
    }
  }
  return most;
}

// Whether a sort with four-way passes starts from runs of half a tile: when
// runs of a whole tile leave an odd number of merge levels (one two-way pass
// plus four-way passes), a blocksort of NT / 2 x VT makes the number even
// (four-way passes only, the same number of passes).  The smaller blocksort
// is the cheaper one (2^28 u32 pairs: 256 x 19 1.69 ms vs 512 x 19 2.08).
template <int NT>
inline bool mergesort_half_tiles(bool four_way, int64_t count, int64_t nv) {
  if (NT < 2 * warp_size || !four_way) return false;
  const int levels = find_log2(div_up(count, nv), true);
  const int half_levels = find_log2(div_up(count, nv / 2), true);
  return (levels & 1) && !(half_levels & 1);
}

// Bytes of aux a mergesort call needs (vk/mergesort.hxx:35-41): the
// partitions (four 32-bit offsets per tile for the four-way passes; the
// two-way pass reads the same array as 64-bit merge paths), the four-way
// passes' samples, and the ping-pong keys and values.
template <int NT, int VT, typename key_t, typename val_t>
inline size_t mergesort_aux_bytes(int64_t count) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr int NV = NT * VT;
  int64_t num_ctas = div_up(count, (int64_t)NV);
  if (find_log2(num_ctas, true) == 0) return 0;
  size_t need = aux_bytes<int4>(num_ctas + 1) + aux_bytes<int2>(mergesort_max_samples(count, NV)) +
                aux_bytes<key_t>(count);
  if (has_values) need += aux_bytes<val_t>(count);
  return need;
}

template <int NT, int VT, bool sort_indices, typename key_t, typename val_t, typename comp_t>
mgpu_status_t mergesort_kv_impl(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                                val_t* vals, int64_t count, comp_t comp) {
  static_assert(!sort_indices || std::is_same_v<val_t, int32_t>, "indices are int32");
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr int NV = NT * VT;
  static_assert(NV % merge4_sample_div == 0, "samples must divide the tile");
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
  // Levels of two-way merging: all as two-way passes for small sorts, else
  // four-way passes of two levels each, after one two-way pass or from runs
  // of half a tile when the number of levels is odd.
  const bool four_way = mergesort_four_way<key_t, val_t>(count);
  const bool half = mergesort_half_tiles<NT>(four_way, count, NV);
  const int64_t run0 = half ? NV / 2 : NV;
  const int levels = half ? find_log2(div_up(count, run0), true) : num_passes;
  const int two_way_passes = four_way ? levels & 1 : levels;
  const int data_passes = two_way_passes + (levels - two_way_passes) / 2;
  void* p = aux_data;
  const int64_t num_partitions = num_ctas + 1;
  int4* mp4 = advance_pointer<int4>(p, num_partitions);
  int64_t* mp = (int64_t*)mp4;
  int2* samples = advance_pointer<int2>(p, mergesort_max_samples(count, NV));
  key_t* keys2 = advance_pointer<key_t>(p, count);
  val_t* vals2 = has_values ? advance_pointer<val_t>(p, count) : nullptr;
  key_t* keys_bs = (1 & data_passes) ? keys2 : keys;
  val_t* vals_bs = (1 & data_passes) ? vals2 : vals;
  if constexpr (NT >= 2 * warp_size) {
    if (half)
      e = launch_kernel("mgpu_blocksort",
                        kernel_blocksort<NT / 2, VT, sort_indices, key_t, val_t, comp_t>,
                        dim3((unsigned)div_up(count, run0)), dim3(NT / 2), 0, stream,
                        (const key_t*)keys, (const val_t*)vals, keys_bs, vals_bs, count, comp);
  }
  if (!half)
    e = launch_kernel("mgpu_blocksort", kernel_blocksort<NT, VT, sort_indices, key_t, val_t, comp_t>,
                      dim3((unsigned)num_ctas), dim3(NT), 0, stream, (const key_t*)keys,
                      (const val_t*)vals, keys_bs, vals_bs, count, comp);
  if (e != hipSuccess) return hip_status(e);
  key_t* k_in = keys_bs;
  val_t* v_in = vals_bs;
  key_t* k_out = (1 & data_passes) ? keys : keys2;
  val_t* v_out = (1 & data_passes) ? vals : vals2;
  // Partition searches are chains of dependent loads, one per thread: blocks of
  // one wave spread them over every CU (256-thread blocks left most CUs idle).
  constexpr int PNT = 64;
  const unsigned partition_grid = (unsigned)div_up(num_partitions, (int64_t)PNT);
  int level = 0;
  for (; level < two_way_passes; ++level) {
    const int64_t coop = (int64_t)2 << level;
    e = launch_kernel("mgpu_mergesort_partition", kernel_mergesort_partition<PNT, key_t, comp_t>,
                      dim3(partition_grid), dim3(PNT), 0, stream, (const key_t*)k_in, count, coop,
                      (int64_t)NV, num_partitions, mp, comp);
    if (e != hipSuccess) return hip_status(e);
    e = launch_kernel("mgpu_mergesort_pass", kernel_mergesort_pass<NT, VT, key_t, val_t, comp_t>,
                      dim3((unsigned)num_ctas), dim3(NT), 0, stream, (const int64_t*)mp,
                      (const key_t*)k_in, (const val_t*)v_in, k_out, v_out, count, coop, comp);
    if (e != hipSuccess) return hip_status(e);
    std::swap(k_in, k_out);
    std::swap(v_in, v_out);
  }
  for (; level < levels; level += 2) {
    const int64_t run = run0 << level;
    constexpr int64_t S = NV / merge4_sample_div;
    const int64_t per_side = div_up(count, 4 * run) * merge4_samples_per_group(run, S);
    e = launch_kernel("mgpu_mergesort_samples", kernel_mergesort_samples4<PNT, key_t, comp_t>,
                      dim3((unsigned)div_up(2 * per_side, (int64_t)PNT)), dim3(PNT), 0, stream,
                      (const key_t*)k_in, count, run, S, per_side, samples, comp);
    if (e != hipSuccess) return hip_status(e);
    e = launch_kernel("mgpu_mergesort_partition", kernel_mergesort_partition4<PNT, key_t, comp_t>,
                      dim3(partition_grid), dim3(PNT), 0, stream, (const key_t*)k_in, count, run,
                      (int64_t)NV, num_partitions, (const int2*)samples, per_side, S, mp4, comp);
    if (e != hipSuccess) return hip_status(e);
    e = launch_kernel("mgpu_mergesort_pass", kernel_mergesort_pass4<NT, VT, key_t, val_t, comp_t>,
                      dim3((unsigned)num_ctas), dim3(NT), 0, stream, (const int4*)mp4,
                      (const key_t*)k_in, (const val_t*)v_in, k_out, v_out, count, run, comp);
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
