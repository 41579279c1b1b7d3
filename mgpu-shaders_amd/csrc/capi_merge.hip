// C ABI: two-way merge (inc/mgpu/vk/merge.hxx:12-59 counterpart).
#include "mgpu/hip/merge.hxx"
#include "tuning.hxx"

using namespace mgpu::hip;

#define MGPU_DEF_MERGE(K, KT)                                                              \
  extern "C" mgpu_status_t mgpu_merge_keys_##K(void* aux, size_t* aux_size,                \
                                               mgpu_stream_t stream, const KT* a_keys,     \
                                               int64_t a_count, const KT* b_keys,          \
                                               int64_t b_count, KT* c_keys,                \
                                               mgpu_comp_t comp) {                         \
    if (!aux_size) return MGPU_INVALID_ARG;                                                \
    constexpr int NT = merge_tuning<KT>::nt, VT = merge_tuning<KT>::vt;                    \
    hipStream_t s = (hipStream_t)stream;                                                   \
    if (comp == MGPU_LESS)                                                                 \
      return merge<NT, VT>(aux, *aux_size, s, a_keys, a_count, b_keys, b_count, c_keys,    \
                           less_t<KT>());                                                  \
    if (comp == MGPU_GREATER)                                                              \
      return merge<NT, VT>(aux, *aux_size, s, a_keys, a_count, b_keys, b_count, c_keys,    \
                           greater_t<KT>());                                               \
    return MGPU_INVALID_ARG;                                                               \
  }                                                                                        \
  extern "C" mgpu_status_t mgpu_merge_kv_##K##_u32(                                        \
      void* aux, size_t* aux_size, mgpu_stream_t stream, const KT* a_keys,                 \
      const uint32_t* a_vals, int64_t a_count, const KT* b_keys, const uint32_t* b_vals,   \
      int64_t b_count, KT* c_keys, uint32_t* c_vals, mgpu_comp_t comp) {                   \
    if (!aux_size) return MGPU_INVALID_ARG;                                                \
    constexpr int NT = merge_tuning<KT>::nt, VT = merge_tuning<KT>::vt;                    \
    hipStream_t s = (hipStream_t)stream;                                                   \
    if (comp == MGPU_LESS)                                                                 \
      return merge<NT, VT>(aux, *aux_size, s, a_keys, a_vals, a_count, b_keys, b_vals,     \
                           b_count, c_keys, c_vals, less_t<KT>());                         \
    if (comp == MGPU_GREATER)                                                              \
      return merge<NT, VT>(aux, *aux_size, s, a_keys, a_vals, a_count, b_keys, b_vals,     \
                           b_count, c_keys, c_vals, greater_t<KT>());                      \
    return MGPU_INVALID_ARG;                                                               \
  }

MGPU_DEF_MERGE(u32, uint32_t)
MGPU_DEF_MERGE(i32, int32_t)
MGPU_DEF_MERGE(u64, uint64_t)
MGPU_DEF_MERGE(f32, float)
MGPU_DEF_MERGE(i64, int64_t)
MGPU_DEF_MERGE(f64, double)

// Merge-path partitions (inc/mgpu/vk/partition.hxx:9-20 counterpart).
template <bounds_t bounds, typename KT>
static mgpu_status_t partitions_by_comp(hipStream_t s, const KT* a, int64_t a_count, const KT* b,
                                        int64_t b_count, int64_t* mp, int64_t spacing,
                                        mgpu_comp_t comp) {
  if (comp == MGPU_LESS)
    return merge_path_partitions<bounds>(s, a, a_count, b, b_count, mp, spacing, less_t<KT>());
  if (comp == MGPU_GREATER)
    return merge_path_partitions<bounds>(s, a, a_count, b, b_count, mp, spacing, greater_t<KT>());
  return MGPU_INVALID_ARG;
}

#define MGPU_DEF_PARTITIONS(K, KT)                                                          \
  extern "C" mgpu_status_t mgpu_merge_path_partitions_##K(                                  \
      mgpu_stream_t stream, const KT* a, int64_t a_count, const KT* b, int64_t b_count,     \
      int64_t* mp, int64_t spacing, mgpu_comp_t comp, mgpu_bounds_t bounds) {               \
    hipStream_t s = (hipStream_t)stream;                                                    \
    if (bounds == MGPU_BOUNDS_LOWER)                                                        \
      return partitions_by_comp<bounds_lower>(s, a, a_count, b, b_count, mp, spacing, comp); \
    if (bounds == MGPU_BOUNDS_UPPER)                                                        \
      return partitions_by_comp<bounds_upper>(s, a, a_count, b, b_count, mp, spacing, comp); \
    return MGPU_INVALID_ARG;                                                                \
  }

MGPU_DEF_PARTITIONS(u32, uint32_t)
MGPU_DEF_PARTITIONS(i32, int32_t)
MGPU_DEF_PARTITIONS(u64, uint64_t)
MGPU_DEF_PARTITIONS(f32, float)
MGPU_DEF_PARTITIONS(i64, int64_t)
MGPU_DEF_PARTITIONS(f64, double)
