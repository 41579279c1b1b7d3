// Batched segmented sort for gfx950: thousands of uneven segments per call.
//
// No counterpart in the reference (it has no segmented primitives); the call
// shape follows the mgpu family's segmented_sort: keys (+ values) of `count`
// elements, `segments[i]` = first element of segment i (segments[0] = 0, the
// last segment ends at count), one stable sort per segment by `comp`.
//
// Every step is a device kernel whose sizes come from device memory, so the
// whole call is one stream-ordered sequence with no host synchronisation and
// can be captured in a graph:
//   1. classify: segment i goes to the first size bucket whose capacity holds
//      it (blocksort shapes NT x VT, chosen per bucket by the autotuner), to
//      the `large` bucket beyond the largest capacity, nowhere if len <= 1;
//      flags[b * S + i] = 1 for that bucket;
//   2. an exclusive scan of the flags gives each segment's slot in one list
//      per bucket (segments in index order), and the bucket bounds;
//   3. per bucket one persistent kernel: each block sorts whole segments in
//      LDS (the blocksort of mergesort.hxx: odd-even network + in-LDS merge
//      passes) and takes the next segment of its bucket;
//   4. large segments: a segmented LSD radix sort over 8192-key tiles that
//      never straddle segments.  Counts are laid out [segment][digit][tile],
//      so one exclusive scan gives every key's position in the concatenation
//      of the large segments; position - (start of its segment in that
//      concatenation) + (segment start) is its place in the array.
//      Descending order uses inverted bits; float keys map -0.0 to +0.0 so
//      ties follow comparator semantics.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <utility>

#include "mergesort.hxx"
#include "radix.hxx"
#include "scan.hxx"

namespace mgpu {
namespace hip {

// Blocksort shapes available to the buckets.  VT is odd: the blocksort reads
// and writes LDS in thread order (word VT * tid + i), and an odd stride
// spreads a wave's accesses over every bank, where a power-of-two VT puts
// them in a few (measured on MI355X: 2048 tiles of 8192 key-value pairs take
// 0.51 ms at 512 x 16, 0.15 ms per 2^24 pairs at 512 x 19).  Indices 0-7 are
// the original table; 8-15 were added for buckets spaced 1.5x apart.
struct segsort_shape_t { int nt, vt; };
constexpr int segsort_num_shapes = 16;
constexpr segsort_shape_t segsort_shapes[segsort_num_shapes] = {
    {64, 5},   {64, 9},   {128, 9},  {64, 17},  {256, 9},  {256, 17}, {512, 9},  {512, 17},
    {64, 3},   {64, 7},   {128, 7},  {128, 13}, {128, 17}, {256, 13}, {512, 13}, {1024, 17}};
constexpr int segsort_max_buckets = 13;
// Large-segment radix tiles.
constexpr int segsort_radix_nt = 512, segsort_radix_vt = 16, segsort_radix_cnt = 1024;
constexpr int segsort_radix_nv = segsort_radix_nt * segsort_radix_vt;

// Whether shape i's blocksort tile fits the 160 KB of LDS for these key and
// value types (1024 x 17 holds 4-byte keys with values, or 8-byte keys alone).
template <typename key_t, typename val_t, int... I>
constexpr std::array<bool, sizeof...(I)> segsort_fit_table(std::integer_sequence<int, I...>) {
  return {{(sizeof(blocksort_storage_t<segsort_shapes[I].nt, segsort_shapes[I].vt, key_t,
                                       val_t>) <= 160 * 1024)...}};
}
template <typename key_t, typename val_t>
inline bool segsort_shape_fits(int shape) {
  constexpr std::array<bool, segsort_num_shapes> fit =
      segsort_fit_table<key_t, val_t>(std::make_integer_sequence<int, segsort_num_shapes>());
  return shape >= 0 && shape < segsort_num_shapes && fit[shape];
}

// A bucket table: buckets hold segments of length <= cap[b] (ascending); each
// bucket is sorted by blocksort shape `shape[b]` (index into segsort_shapes,
// capacity >= cap[b]).  Longer segments go to the large bucket.
struct segsort_tuning_t {
  int num_buckets;
  int cap[segsort_max_buckets];
  int shape[segsort_max_buckets];
};

// Capacities 1.5x apart, so that a segment fills about four fifths of its
// tile on average (at 2x apart, about two thirds), up to 16384 keys where
// the 1024 x 17 tile fits the LDS and 8192 otherwise.
inline segsort_tuning_t segsort_tuning_for(bool tile_16k) {
  segsort_tuning_t t;
  const int caps[13] = {128, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192, 16384};
  const int shapes[13] = {8, 0, 9, 1, 10, 2, 11, 4, 13, 5, 14, 7, 15};
  t.num_buckets = tile_16k ? 13 : 12;
  for (int b = 0; b < segsort_max_buckets; ++b) {
    t.cap[b] = b < t.num_buckets ? caps[b] : 0;
    t.shape[b] = b < t.num_buckets ? shapes[b] : 0;
  }
  return t;
}
template <typename key_t, typename val_t>
inline segsort_tuning_t segsort_default_tuning() {
  return segsort_tuning_for(segsort_shape_fits<key_t, val_t>(15));
}

struct segsort_bucket_caps_t { int cap[segsort_max_buckets]; int num_buckets; };

// A read of device data that no kernel of the call writes while it runs,
// through the constant address space: where the index is uniform over the
// wave the load becomes a scalar load, which is counted apart from the vector
// memory operations, so waiting for it does not also wait for the stores and
// loads in flight before it.
template <typename T>
__device__ inline T read_const(const T* p, int64_t i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Segment i is [begin, end) with both ends clamped to [0, count] and
// end >= begin: a start at or beyond count gives an empty segment, as in the
// segmented scan, and no start can place a segment outside the arrays.
__device__ inline void segment_bounds(const int32_t* segments, int64_t num_segments, int64_t count,
                                      int64_t i, int64_t& begin, int64_t& end) {
  begin = min_<int64_t>(max_<int64_t>(read_const(segments, i), 0), count);
  const int64_t next = i + 1 < num_segments ? (int64_t)read_const(segments, i + 1) : count;
  end = min_<int64_t>(max_<int64_t>(next, begin), count);
}

// 1. flags[b * S + i] for b in [0, num_buckets] (num_buckets = large), and
//    large_tiles[i] = tiles of segment i if it is large, else 0.  The grid
//    also zeroes clear[0, clear_words) (the status words of the two scans
//    that follow), which saves those scans their own clearing launches.
template <int NT>
__global__ __launch_bounds__(NT) void kernel_segsort_classify(
    const int32_t* segments, int64_t num_segments, int64_t count, segsort_bucket_caps_t caps,
    uint32_t* flags, uint64_t* large_tiles, uint32_t* clear, int64_t clear_words) {
  for (int64_t j = (int64_t)blockIdx.x * NT + threadIdx.x; j < clear_words;
       j += (int64_t)gridDim.x * NT)
    clear[j] = 0;
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (i >= num_segments) return;
  int64_t b0, b1;
  segment_bounds(segments, num_segments, count, i, b0, b1);
  const int64_t len = b1 - b0;
  int bucket = -1;
  if (len > 1) {
    bucket = caps.num_buckets;
    for (int b = caps.num_buckets - 1; b >= 0; --b)
      if (len <= caps.cap[b]) bucket = b;
  }
  for (int b = 0; b <= caps.num_buckets; ++b)
    flags[b * num_segments + i] = bucket == b ? 1u : 0u;
  large_tiles[i] = bucket == caps.num_buckets ? (uint64_t)div_up(len, (int64_t)segsort_radix_nv) : 0u;
}

// 2. list[pos[k]] = segment of flag k; bounds[b] = {start, end} of bucket b's
//    list; large_tasks[tile_pos[i] + t] = (i, t) for every tile t of a large
//    segment i; *num_tasks = total large tiles.
//    Segments whose starts ascend (the contract) have at most max_tiles tiles
//    in total.  Starts that do not ascend make segments overlap and can
//    exceed it: a segment whose tiles do not all fit below max_tiles is then
//    left unsorted (its slots below max_tiles get the task (-1, 0), which the
//    radix kernels skip), so no write leaves the aux arrays.
template <int NT>
__global__ __launch_bounds__(NT) void kernel_segsort_compact(
    const uint32_t* flags, const uint32_t* pos, int64_t num_segments, int num_buckets,
    int32_t* list, int32_t* bounds, const uint64_t* large_tiles, const uint64_t* tile_pos,
    int2* large_tasks, int32_t* num_tasks, int64_t max_tiles) {
  const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
  const int64_t total = (int64_t)(num_buckets + 1) * num_segments;
  if (k < total && flags[k]) list[pos[k]] = (int32_t)(k % num_segments);
  if (k <= num_buckets) {
    const int b = (int)k;
    bounds[2 * b] = (int32_t)pos[(int64_t)b * num_segments];
    bounds[2 * b + 1] = b == num_buckets ? (int32_t)(pos[total - 1] + flags[total - 1])
                                         : (int32_t)pos[(int64_t)(b + 1) * num_segments];
  }
  if (k < num_segments) {
    const uint64_t n = large_tiles[k], first = tile_pos[k];
    const bool fits = first + n <= (uint64_t)max_tiles;
    for (uint64_t t = 0; t < n && first + t < (uint64_t)max_tiles; ++t)
      large_tasks[first + t] = fits ? make_int2((int)k, (int)t) : make_int2(-1, 0);
    if (k == num_segments - 1) *num_tasks = (int32_t)min_<uint64_t>(first + n, (uint64_t)max_tiles);
  }
}

// 3. Sort whole segments of one bucket in LDS, block b taking list entries
//    bounds[0] + b, + grid, ...
template <int NT, int VT, typename key_t, typename val_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_segsort_blocks(
    key_t* keys, val_t* vals, int64_t count, const int32_t* segments, int64_t num_segments,
    const int32_t* list, const int32_t* bounds, comp_t comp) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  __shared__ blocksort_storage_t<NT, VT, key_t, val_t> s;
  const int64_t first = read_const(bounds, 0), last = read_const(bounds, 1);
  for (int64_t e = first + (int64_t)blockIdx.x; e < last; e += gridDim.x) {
    int64_t b0, b1;
    segment_bounds(segments, num_segments, count, read_const(list, e), b0, b1);
    val_t* v = has_values ? vals + b0 : vals;
    cta_blocksort<NT, VT, false>(keys + b0, v, keys + b0, v, (int)(b1 - b0), 0, comp, s);
    __syncthreads();
  }
}

// Sort bits for the large-segment radix: radix order with -0.0 tied to +0.0
// for floats (comparator order), inverted for descending comparators.
template <typename key_t, bool descending>
struct segsort_bits_t {
  typedef typename radix_permute_t<key_t>::unsigned_type type;
  __device__ type operator()(key_t k) const {
    if constexpr (std::is_floating_point_v<key_t>) {
      if (k == key_t(0)) k = key_t(0);
    }
    type b = radix_permute_t<key_t>::to_radix_bits(k);
    return descending ? (type)~b : b;
  }
};

// 4a. Counts of one digit for every large-segment tile task:
//     counts[256 * tile_pos[seg] + d * tiles(seg) + t].
template <typename key_t, bool descending>
__global__ __launch_bounds__(segsort_radix_cnt) void kernel_segsort_radix_count(
    const key_t* keys, int64_t count, const int32_t* segments, int64_t num_segments,
    const int2* tasks, const int32_t* num_tasks, const uint64_t* large_tiles,
    const uint64_t* tile_pos, int shift, uint32_t* counts, uint32_t* clear, int64_t clear_words) {
  __shared__ uint32_t h[radix_num_bins * 16];
  // Also zero the status words of the counts scan that follows.
  for (int64_t i = (int64_t)blockIdx.x * segsort_radix_cnt + threadIdx.x; i < clear_words;
       i += (int64_t)gridDim.x * segsort_radix_cnt)
    clear[i] = 0;
  const int32_t n = *num_tasks;
  for (int64_t e = blockIdx.x; e < n; e += gridDim.x) {
    const int2 task = tasks[e];
    if (task.x < 0) continue;  // a segment left unsorted (see kernel_segsort_compact)
    int64_t b0, b1;
    segment_bounds(segments, num_segments, count, task.x, b0, b1);
    const int64_t base = b0 + (int64_t)task.y * segsort_radix_nv;
    const int tile_count = (int)min_<int64_t>(segsort_radix_nv, b1 - base);
    const uint32_t tiles = (uint32_t)large_tiles[task.x];
    cta_radix_count<segsort_radix_cnt, segsort_radix_nv>(
        keys + base, tile_count, shift, segsort_bits_t<key_t, descending>(),
        counts + 256ull * tile_pos[task.x] + task.y, tiles, h);
    __syncthreads();
  }
}

// 4b. Scatter of one digit for every large-segment tile task.
template <typename key_t, bool has_values, bool descending>
__global__ __launch_bounds__(segsort_radix_nt) void kernel_segsort_radix_scatter(
    const key_t* keys_in, key_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
    int64_t count, const int32_t* segments, int64_t num_segments, const int2* tasks,
    const int32_t* num_tasks, const uint64_t* large_tiles, const uint64_t* tile_pos, int shift,
    const uint32_t* offsets) {
  __shared__ radix_scatter_storage_t<segsort_radix_nt, segsort_radix_vt, key_t> s;
  const int tid = threadIdx.x;
  const int32_t n = *num_tasks;
  for (int64_t e = blockIdx.x; e < n; e += gridDim.x) {
    const int2 task = tasks[e];
    if (task.x < 0) continue;  // uniform over the block: one task per block
    int64_t b0, b1;
    segment_bounds(segments, num_segments, count, task.x, b0, b1);
    const int64_t base = b0 + (int64_t)task.y * segsort_radix_nv;
    const int tile_count = (int)min_<int64_t>(segsort_radix_nv, b1 - base);
    const uint32_t tiles = (uint32_t)large_tiles[task.x];
    const uint64_t cbase = 256ull * tile_pos[task.x];
    // Position of the segment's first key in the concatenation of the large
    // segments (digit 0, tile 0 of this segment), mapped back to b0.
    const uint32_t seg_first = offsets[cbase];
    uint32_t goff = 0;
    if (tid < radix_num_bins)
      goff = offsets[cbase + (uint64_t)tid * tiles + task.y] - seg_first + (uint32_t)b0;
    cta_radix_scatter<segsort_radix_nt, segsort_radix_vt, key_t, has_values>(
        keys_in + base, has_values ? vals_in + base : vals_in, keys_out, vals_out, tile_count,
        shift, segsort_bits_t<key_t, descending>(), goff, count, s);
    __syncthreads();
  }
}

// Branches of one call (see segmented_sort).
constexpr int segsort_branches = 4;

// Persistent grids: enough blocks to fill the chip, never more than the work.
// The blocksort buckets launch segsort_grid_threads / NT blocks (8192 blocks
// of 64 threads: small shapes need many resident blocks to hide latency).
constexpr int segsort_grid_blocks = 2048;
constexpr int segsort_grid_threads = 256 * 2048;

// Upper bound of large-segment tiles: at most count / (largest capacity + 1)
// large segments, each adding at most one partial tile.
inline int64_t segsort_max_tiles(int64_t count, const segsort_tuning_t& t) {
  return div_up(count, (int64_t)segsort_radix_nv) + count / (t.cap[t.num_buckets - 1] + 1) + 1;
}

// Status words of the flag scan (one 64-bit word per 256 x 16 tile of the
// B * S flags, rounded up to an even count) followed by those of the 64-bit
// tile-position scan (two per scan64 tile of the S segments).
inline int64_t segsort_flag_scan_words(int64_t S, int64_t B) {
  return 2 * div_up(div_up(B * S, (int64_t)(256 * 16)), (int64_t)2);
}
inline int64_t segsort_scan_status_words(int64_t S, int64_t B) {
  return segsort_flag_scan_words(S, B) + 2 * div_up(S, scan64_tile_size(S));
}

template <typename key_t, typename val_t>
inline size_t segmented_sort_aux_bytes(int64_t count, int64_t num_segments,
                                       const segsort_tuning_t& tuning) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  const int64_t S = num_segments, B = tuning.num_buckets + 1;
  const int64_t max_tiles = segsort_max_tiles(count, tuning);
  const int64_t max_counts = 256 * max_tiles;
  size_t need = 0;
  need += aux_bytes<uint32_t>(B * S) * 2;         // flags, pos
  need += aux_bytes<int32_t>(B * S);              // lists
  need += aux_bytes<int32_t>(2 * B);              // bounds
  need += aux_bytes<uint64_t>(S) * 2;             // large tiles, tile_pos
  need += aux_bytes<int2>(max_tiles);             // tasks
  need += aux_bytes<int32_t>(1);                  // number of tasks
  need += aux_bytes<uint32_t>(max_counts);        // counts / offsets
  need += aux_bytes<uint64_t>(segsort_scan_status_words(S, B));  // status of the two scans
  need += aux_bytes<key_t>(count);                // radix ping-pong keys
  if (has_values) need += aux_bytes<uint32_t>(count);
  size_t scan_need = 0;  // the status words of the large-segment counts scan
  scan<256, 16>(nullptr, scan_need, nullptr, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                max_counts, MGPU_SCAN_EXCLUSIVE, 0u);
  return need + scan_need;
}

template <int SHAPE, typename key_t, typename val_t, typename comp_t>
inline hipError_t segsort_launch_bucket(key_t* keys, val_t* vals, int64_t count,
                                        const int32_t* segments, int64_t num_segments,
                                        const int32_t* list, const int32_t* bounds, comp_t comp,
                                        hipStream_t stream) {
  constexpr int NT = segsort_shapes[SHAPE].nt, VT = segsort_shapes[SHAPE].vt;
  if constexpr (sizeof(blocksort_storage_t<NT, VT, key_t, val_t>) <= 160 * 1024) {
    const unsigned grid = (unsigned)min_<int64_t>(num_segments, segsort_grid_threads / NT);
    return launch_kernel("mgpu_segsort_blocks",
                         kernel_segsort_blocks<NT, VT, key_t, val_t, comp_t>, dim3(grid), dim3(NT),
                         0, stream, keys, vals, count, segments, num_segments, list, bounds, comp);
  } else {
    return hipErrorInvalidValue;  // segmented_sort rejects such a table first
  }
}

template <typename key_t, typename val_t, typename comp_t>
inline hipError_t segsort_dispatch_shape(int shape, key_t* keys, val_t* vals, int64_t count,
                                         const int32_t* segments, int64_t num_segments,
                                         const int32_t* list, const int32_t* bounds,
                                         comp_t comp, hipStream_t stream) {
#define MGPU_SEGSORT_SHAPE(i)                                                             \
  case i:                                                                                 \
    return segsort_launch_bucket<i>(keys, vals, count, segments, num_segments, list,     \
                                    bounds, comp, stream);
  switch (shape) {
    MGPU_SEGSORT_SHAPE(0) MGPU_SEGSORT_SHAPE(1) MGPU_SEGSORT_SHAPE(2) MGPU_SEGSORT_SHAPE(3)
    MGPU_SEGSORT_SHAPE(4) MGPU_SEGSORT_SHAPE(5) MGPU_SEGSORT_SHAPE(6) MGPU_SEGSORT_SHAPE(7)
    MGPU_SEGSORT_SHAPE(8) MGPU_SEGSORT_SHAPE(9) MGPU_SEGSORT_SHAPE(10) MGPU_SEGSORT_SHAPE(11)
    MGPU_SEGSORT_SHAPE(12) MGPU_SEGSORT_SHAPE(13) MGPU_SEGSORT_SHAPE(14) MGPU_SEGSORT_SHAPE(15)
  }
#undef MGPU_SEGSORT_SHAPE
  return hipErrorInvalidValue;
}

// Stable per-segment sort; val_t = uint32_t or empty_t.  descending selects
// the large-segment bit order matching comp (less: false, greater: true).
template <bool descending, typename key_t, typename val_t, typename comp_t>
mgpu_status_t segmented_sort(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                             val_t* vals, int64_t count, const int32_t* segments,
                             int64_t num_segments, comp_t comp, const segsort_tuning_t& tuning) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  static_assert(!has_values || std::is_same_v<val_t, uint32_t>, "values are uint32");
  if (count < 0 || num_segments < 0) return MGPU_INVALID_ARG;
  if (tuning.num_buckets < 1 || tuning.num_buckets > segsort_max_buckets) return MGPU_INVALID_ARG;
  for (int b = 0; b < tuning.num_buckets; ++b) {
    if (!segsort_shape_fits<key_t, val_t>(tuning.shape[b])) return MGPU_INVALID_ARG;
    const segsort_shape_t sh = segsort_shapes[tuning.shape[b]];
    if (tuning.cap[b] < 2 || tuning.cap[b] > sh.nt * sh.vt) return MGPU_INVALID_ARG;
    if (b && tuning.cap[b] <= tuning.cap[b - 1]) return MGPU_INVALID_ARG;
  }
  if (count == 0 || num_segments == 0) return MGPU_OK;
  if (count >= ((int64_t)1 << 31) || num_segments >= ((int64_t)1 << 31) / (segsort_max_buckets + 1))
    return MGPU_UNSUPPORTED;
  const size_t need = segmented_sort_aux_bytes<key_t, val_t>(count, num_segments, tuning);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if (!keys || !segments || (has_values && !vals)) return MGPU_INVALID_ARG;
  if (mgpu_status_t st = check_aux(aux_data, aux_size, need)) return st;

  const int NB = tuning.num_buckets;
  const int64_t S = num_segments, B = NB + 1;
  const int64_t max_tiles = segsort_max_tiles(count, tuning);
  const int64_t max_counts = 256 * max_tiles;
  void* p = aux_data;
  uint32_t* flags = advance_pointer<uint32_t>(p, B * S);
  uint32_t* pos = advance_pointer<uint32_t>(p, B * S);
  int32_t* list = advance_pointer<int32_t>(p, B * S);
  int32_t* bounds = advance_pointer<int32_t>(p, 2 * B);
  uint64_t* large_tiles = advance_pointer<uint64_t>(p, S);
  uint64_t* tile_pos = advance_pointer<uint64_t>(p, S);
  int2* tasks = advance_pointer<int2>(p, max_tiles);
  int32_t* num_tasks = advance_pointer<int32_t>(p, 1);
  uint32_t* counts = advance_pointer<uint32_t>(p, max_counts);
  uint64_t* prefix_status = advance_pointer<uint64_t>(p, segsort_scan_status_words(S, B));
  uint64_t* scan64_status = prefix_status + segsort_flag_scan_words(S, B);
  key_t* keys2 = advance_pointer<key_t>(p, count);
  uint32_t* vals2 = has_values ? advance_pointer<uint32_t>(p, count) : nullptr;

  segsort_bucket_caps_t caps;
  caps.num_buckets = NB;
  for (int b = 0; b < segsort_max_buckets; ++b) caps.cap[b] = b < NB ? tuning.cap[b] : 0;

  hipError_t e = launch_kernel("mgpu_segsort_classify", kernel_segsort_classify<256>,
                               dim3((unsigned)div_up(S, (int64_t)256)), dim3(256), 0, stream,
                               segments, S, count, caps, flags, large_tiles,
                               (uint32_t*)prefix_status, 2 * segsort_scan_status_words(S, B));
  if (e != hipSuccess) return hip_status(e);
  e = scan_launch_cleared<256, 16>(stream, (const uint32_t*)flags, pos, B * S,
                                   MGPU_SCAN_EXCLUSIVE, 0u, identity_t(), prefix_status);
  if (e != hipSuccess) return hip_status(e);
  // 64-bit tile positions: their sum cannot wrap even when overlapping
  // segments (starts that do not ascend) claim more tiles than exist.
  e = scan64_launch_cleared<uint64_t>(stream, (const uint64_t*)large_tiles, tile_pos, S,
                                      MGPU_SCAN_EXCLUSIVE, (uint64_t)0, scan64_status);
  if (e != hipSuccess) return hip_status(e);
  e = launch_kernel("mgpu_segsort_compact", kernel_segsort_compact<256>,
                    dim3((unsigned)div_up(B * S, (int64_t)256)), dim3(256), 0, stream,
                    (const uint32_t*)flags, (const uint32_t*)pos, S, NB, list, bounds,
                    (const uint64_t*)large_tiles, (const uint64_t*)tile_pos, tasks, num_tasks,
                    max_tiles);
  if (e != hipSuccess) return hip_status(e);

  // The buckets and the large segments hold disjoint segments and run on
  // branches of their own: branch 0 takes the large-segment radix (a chain of
  // 4 x (count, scan, scatter) launches, submitted first so that it overlaps
  // the buckets instead of queueing behind them), branches 1-3 the buckets in
  // three contiguous groups, largest capacity first: the largest tiles, which
  // leave a CU room only for small ones, run one after another on branch 1
  // while the middle and the smallest groups run beside them.  Four
  // branches, as a process has four hardware queues by default: kernels of
  // streams that share a queue run in submission order.
  fork_join_t fork(stream, segsort_branches);

  // Large segments: segmented LSD radix, ping-pong keys <-> keys2.
  constexpr int D = (int)sizeof(key_t);
  key_t* k_in = keys;
  key_t* k_out = keys2;
  uint32_t* v_in = has_values ? (uint32_t*)vals : nullptr;
  uint32_t* v_out = vals2;
  const unsigned grid = (unsigned)min_<int64_t>(max_tiles, segsort_grid_blocks);
  // The counts scan keeps its status words (one 64-bit word per 4096 counts)
  // at the start of the scan aux; the count kernel zeroes them.
  uint64_t* scan_status = (uint64_t*)p;
  const int64_t scan_words = 2 * div_up(max_counts, (int64_t)(256 * 16));
  const hipStream_t rs = fork.branch(0);
  for (int pass = 0; pass < D; ++pass) {
    // Every count a task owns is rewritten each pass; entries past the last
    // task only follow the real ones in the scan, so they need no clearing.
    const int shift = 8 * pass;
    e = launch_kernel("mgpu_segsort_radix_count", kernel_segsort_radix_count<key_t, descending>,
                      dim3(grid), dim3(segsort_radix_cnt), 0, rs, (const key_t*)k_in, count,
                      segments, S, (const int2*)tasks, (const int32_t*)num_tasks,
                      (const uint64_t*)large_tiles, (const uint64_t*)tile_pos, shift, counts,
                      (uint32_t*)scan_status, scan_words);
    if (e != hipSuccess) return hip_status(e);
    e = scan_launch_cleared<256, 16>(rs, (const uint32_t*)counts, counts, max_counts,
                                     MGPU_SCAN_EXCLUSIVE, 0u, identity_t(), scan_status);
    if (e != hipSuccess) return hip_status(e);
    e = launch_kernel("mgpu_segsort_radix_scatter",
                      kernel_segsort_radix_scatter<key_t, has_values, descending>, dim3(grid),
                      dim3(segsort_radix_nt), 0, rs, (const key_t*)k_in, k_out,
                      (const uint32_t*)v_in, v_out, count, segments, S, (const int2*)tasks,
                      (const int32_t*)num_tasks, (const uint64_t*)large_tiles,
                      (const uint64_t*)tile_pos, shift, (const uint32_t*)counts);
    if (e != hipSuccess) return hip_status(e);
    std::swap(k_in, k_out);
    std::swap(v_in, v_out);
  }

  // Blocksort buckets.
  for (int i = 0; i < NB; ++i) {
    const int b = NB - 1 - i;
    e = segsort_dispatch_shape(tuning.shape[b], keys, vals, count, segments, S, list,
                               bounds + 2 * b, comp, fork.branch(1 + i * (segsort_branches - 1) / NB));
    if (e != hipSuccess) return hip_status(e);
  }
  return hip_status(fork.join());
}

////////////////////////////////////////////////////////////////////////////////
// Autotuner: picks the blocksort shape of every bucket of the default table
// by coordinate descent over whole segmented-sort calls.  The buckets of one
// call run concurrently and share the CUs, so a shape that is fastest alone
// (large NT, high occupancy, many empty slots) is not the fastest in the
// call: measured on 65,536 lognormal segments, the shapes chosen by timing
// each bucket alone made the call 8% slower than the default table.  Here the
// trial is the whole call on synthetic segments (about 2^22 SplitMix64
// keys, the same number of keys in every octave of length from 2 to twice
// the largest capacity); each bucket in turn tries every shape that holds
// its capacity (up to 8x it) with the other buckets fixed and takes one
// that makes the call at least 1.5% faster (median of 5), two rounds.
// Returns the table (also usable as an explicit argument); ms_table[b][shape]
// is the call's time with that shape in bucket b (last round), -1 for shapes
// not tried.

template <typename key_t>
__global__ __launch_bounds__(256) void kernel_segsort_tune_fill(key_t* keys, uint32_t* vals,
                                                               int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)i + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    if constexpr (std::is_floating_point_v<key_t>) keys[i] = (key_t)(z >> 40) * (key_t)(1.0 / 16777216.0);
    else keys[i] = (key_t)z;
    if (vals) vals[i] = (uint32_t)i;
  }
}

constexpr double segsort_tune_margin = 0.015;

template <typename key_t, typename val_t, typename comp_t>
inline mgpu_status_t segsort_autotune(hipStream_t stream, comp_t comp, segsort_tuning_t& out,
                                      double* ms_table /* [buckets][shapes] or null */) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr bool descending = std::is_same_v<comp_t, greater_t<key_t>>;
  out = segsort_default_tuning<key_t, val_t>();
  if (ms_table)
    for (int i = 0; i < segsort_max_buckets * segsort_num_shapes; ++i) ms_table[i] = -1;
  // Synthetic layout: lengths in every octave [2^k, 2^(k+1)) from 2 up to
  // twice the largest capacity, the same number of keys in each octave
  // (lengths uniform within it), segments in a shuffled order.
  const int64_t target = (int64_t)1 << 22;
  const int top = 2 * out.cap[out.num_buckets - 1];
  uint64_t rng = 0x243F6A8885A308D3ull;
  auto next_rand = [&]() {
    rng += 0x9E3779B97F4A7C15ull;
    uint64_t z = rng;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  std::vector<int32_t> lens;
  int octaves = 0;
  for (int lo = 2; lo < top; lo *= 2) ++octaves;
  for (int lo = 2; lo < top; lo *= 2) {
    const int hi = min_(2 * lo, top);
    for (int64_t k = 0; k < target / octaves;) {
      const int len = lo + (int)(next_rand() % (uint64_t)(hi - lo));
      lens.push_back(len);
      k += len;
    }
  }
  for (size_t i = lens.size(); i > 1; --i) std::swap(lens[i - 1], lens[next_rand() % i]);
  std::vector<int32_t> h_segs(lens.size());
  int64_t n = 0;
  for (size_t i = 0; i < lens.size(); ++i) {
    h_segs[i] = (int32_t)n;
    n += lens[i];
  }
  const int64_t S = (int64_t)h_segs.size();
  key_t* keys = nullptr;
  uint32_t* vals = nullptr;
  int32_t* segs = nullptr;
  void* aux = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  mgpu_status_t status = MGPU_OK;
  auto check = [&](hipError_t e) {
    if (e != hipSuccess && status == MGPU_OK) status = hip_status(e);
    return e == hipSuccess;
  };
  // Aux for any table of this bucket count: the sizing call depends on the
  // capacities only through the largest one, which the descent keeps.
  size_t aux_size = 0;
  segmented_sort<descending>(nullptr, aux_size, stream, (key_t*)nullptr, (val_t*)nullptr, n,
                             (const int32_t*)nullptr, S, comp, out);
  if (!check(hipMalloc(&keys, n * sizeof(key_t))) ||
      !check(hipMalloc(&vals, n * sizeof(uint32_t))) ||
      !check(hipMalloc(&segs, S * sizeof(int32_t))) || !check(hipMalloc(&aux, aux_size)) ||
      !check(hipEventCreate(&e0)) || !check(hipEventCreate(&e1))) {
    (void)hipFree(keys); (void)hipFree(vals); (void)hipFree(segs); (void)hipFree(aux);
    if (e0) (void)hipEventDestroy(e0);
    return status;
  }
  check(hipMemcpy(segs, h_segs.data(), S * sizeof(int32_t), hipMemcpyHostToDevice));
  // One trial: the median of five timed calls on freshly filled keys.
  constexpr int reps = 5;
  auto trial = [&](const segsort_tuning_t& t) -> double {
    float ms[reps];
    for (int rep = 0; rep <= reps && status == MGPU_OK; ++rep) {
      hipLaunchKernelGGL(kernel_segsort_tune_fill<key_t>, dim3(1024), dim3(256), 0, stream, keys,
                         has_values ? vals : nullptr, n, (uint64_t)rep);
      check(hipEventRecord(e0, stream));
      size_t cap = aux_size;
      const mgpu_status_t st = segmented_sort<descending>(
          aux, cap, stream, keys, has_values ? (val_t*)vals : (val_t*)nullptr, n, segs, S, comp,
          t);
      if (st != MGPU_OK && status == MGPU_OK) status = st;
      check(hipEventRecord(e1, stream));
      check(hipEventSynchronize(e1));
      float x = 0;
      check(hipEventElapsedTime(&x, e0, e1));
      if (rep) ms[rep - 1] = x;  // the first call warms up
    }
    if (status != MGPU_OK) return 1e30;
    std::sort(ms, ms + reps);
    return ms[reps / 2];
  };
  double best = trial(out);
  for (int round = 0; round < 2 && status == MGPU_OK; ++round) {
    for (int b = 0; b < out.num_buckets && status == MGPU_OK; ++b) {
      const int cap = out.cap[b];
      for (int sh = 0; sh < segsort_num_shapes && status == MGPU_OK; ++sh) {
        const int capacity = segsort_shapes[sh].nt * segsort_shapes[sh].vt;
        if (capacity < cap || capacity > 8 * cap || !segsort_shape_fits<key_t, val_t>(sh))
          continue;
        double ms = best;
        if (sh != out.shape[b]) {
          segsort_tuning_t t = out;
          t.shape[b] = sh;
          ms = trial(t);
          // A shape replaces the current one only when the call is clearly
          // faster: buckets that hold few keys move the call's time by less
          // than the run-to-run spread, and noise must not pick for them.
          if (ms < best * (1.0 - segsort_tune_margin)) {
            best = ms;
            out.shape[b] = sh;
          }
        }
        if (ms_table) ms_table[b * segsort_num_shapes + sh] = ms;
      }
    }
  }
  (void)hipFree(keys); (void)hipFree(vals); (void)hipFree(segs); (void)hipFree(aux);
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  return status;
}

}  // namespace hip
}  // namespace mgpu
