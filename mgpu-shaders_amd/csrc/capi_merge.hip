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
