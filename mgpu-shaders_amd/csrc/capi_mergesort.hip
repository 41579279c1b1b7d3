// C ABI: mergesort (inc/mgpu/vk/mergesort.hxx:10-103 counterpart).
#include "mgpu/hip/mergesort.hxx"
#include "tuning.hxx"

using namespace mgpu::hip;

#define MGPU_DEF_MERGESORT(K, KT)                                                          \
  extern "C" mgpu_status_t mgpu_mergesort_keys_##K(void* aux, size_t* aux_size,            \
                                                   mgpu_stream_t stream, KT* keys,         \
                                                   int64_t count, mgpu_comp_t comp) {      \
    if (!aux_size) return MGPU_INVALID_ARG;                                                \
    constexpr int NT = mergesort_tuning<KT>::nt, VT = mergesort_tuning<KT>::vt;            \
    hipStream_t s = (hipStream_t)stream;                                                   \
    if (comp == MGPU_LESS)                                                                 \
      return mergesort_keys<NT, VT>(aux, *aux_size, s, keys, count, less_t<KT>());         \
    if (comp == MGPU_GREATER)                                                              \
      return mergesort_keys<NT, VT>(aux, *aux_size, s, keys, count, greater_t<KT>());      \
    return MGPU_INVALID_ARG;                                                               \
  }                                                                                        \
  extern "C" mgpu_status_t mgpu_mergesort_kv_##K##_u32(void* aux, size_t* aux_size,        \
                                                       mgpu_stream_t stream, KT* keys,     \
                                                       uint32_t* vals, int64_t count,      \
                                                       mgpu_comp_t comp) {                 \
    if (!aux_size) return MGPU_INVALID_ARG;                                                \
    constexpr int NT = mergesort_tuning<KT>::nt, VT = mergesort_tuning<KT>::vt;            \
    hipStream_t s = (hipStream_t)stream;                                                   \
    if (comp == MGPU_LESS)                                                                 \
      return mergesort_kv<NT, VT, false>(aux, *aux_size, s, keys, vals, count,             \
                                         less_t<KT>());                                    \
    if (comp == MGPU_GREATER)                                                              \
      return mergesort_kv<NT, VT, false>(aux, *aux_size, s, keys, vals, count,             \
                                         greater_t<KT>());                                 \
    return MGPU_INVALID_ARG;                                                               \
  }                                                                                        \
  extern "C" mgpu_status_t mgpu_mergesort_indices_##K(void* aux, size_t* aux_size,         \
                                                      mgpu_stream_t stream, KT* keys,      \
                                                      int32_t* indices, int64_t count,     \
                                                      mgpu_comp_t comp) {                  \
    if (!aux_size) return MGPU_INVALID_ARG;                                                \
    constexpr int NT = mergesort_tuning<KT>::nt, VT = mergesort_tuning<KT>::vt;            \
    hipStream_t s = (hipStream_t)stream;                                                   \
    if (comp == MGPU_LESS)                                                                 \
      return mergesort_indices<NT, VT>(aux, *aux_size, s, keys, indices, count,            \
                                       less_t<KT>());                                      \
    if (comp == MGPU_GREATER)                                                              \
      return mergesort_indices<NT, VT>(aux, *aux_size, s, keys, indices, count,            \
                                       greater_t<KT>());                                   \
    return MGPU_INVALID_ARG;                                                               \
  }

MGPU_DEF_MERGESORT(u32, uint32_t)
MGPU_DEF_MERGESORT(i32, int32_t)
MGPU_DEF_MERGESORT(u64, uint64_t)
MGPU_DEF_MERGESORT(f32, float)
MGPU_DEF_MERGESORT(i64, int64_t)
MGPU_DEF_MERGESORT(f64, double)

extern "C" void mgpu_set_mergesort_passes(int ways) {
  mergesort_passes_setting().store(ways == 2 || ways == 4 ? ways : 0);
}
extern "C" int mgpu_mergesort_passes(void) { return mergesort_passes_setting().load(); }
