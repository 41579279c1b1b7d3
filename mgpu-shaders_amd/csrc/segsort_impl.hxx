// Instantiation of the segmented-sort C ABI for one key type (included by
// capi_segsort_<K>.hip with MGPU_SEGSORT_K / MGPU_SEGSORT_KT defined).
#include <map>
#include <mutex>
#include <tuple>

#include "mgpu/hip/segsort.hxx"

using namespace mgpu::hip;

namespace {

// Current table per (values, comparator) for this key type.
struct segsort_tables_t {
  std::mutex mutex;
  std::map<std::pair<int, int>, segsort_tuning_t> tables;
};
segsort_tables_t& seg_tables() {
  static segsort_tables_t t;
  return t;
}
segsort_tuning_t current_tuning(int has_values, int comp) {
  segsort_tables_t& t = seg_tables();
  std::lock_guard<std::mutex> lock(t.mutex);
  auto it = t.tables.find({has_values, comp});
  if (it != t.tables.end()) return it->second;
  return has_values ? segsort_default_tuning<MGPU_SEGSORT_KT, uint32_t>()
                    : segsort_default_tuning<MGPU_SEGSORT_KT, empty_t>();
}
segsort_tuning_t to_tuning(const mgpu_segsort_tuning_t* t, int has_values, int comp) {
  if (!t) return current_tuning(has_values, comp);
  segsort_tuning_t r;
  r.num_buckets = t->num_buckets;
  for (int b = 0; b < segsort_max_buckets; ++b) {
    r.cap[b] = t->cap[b];
    r.shape[b] = t->shape[b];
  }
  return r;
}

template <typename KT, typename VT>
mgpu_status_t run_segsort(void* aux, size_t* aux_size, mgpu_stream_t stream, KT* keys, VT* vals,
                          int64_t count, const int32_t* segments, int64_t num_segments,
                          mgpu_comp_t comp, const mgpu_segsort_tuning_t* t) {
  constexpr int has_values = !std::is_same_v<VT, empty_t>;
  if (!aux_size) return MGPU_INVALID_ARG;
  if (comp != MGPU_LESS && comp != MGPU_GREATER) return MGPU_INVALID_ARG;
  segsort_tuning_t tuning = to_tuning(t, has_values, (int)comp);
  hipStream_t s = (hipStream_t)stream;
  if (comp == MGPU_LESS)
    return segmented_sort<false>(aux, *aux_size, s, keys, vals, count, segments, num_segments,
                                 less_t<KT>(), tuning);
  return segmented_sort<true>(aux, *aux_size, s, keys, vals, count, segments, num_segments,
                              greater_t<KT>(), tuning);
}

template <typename KT>
mgpu_status_t run_autotune(mgpu_stream_t stream, int has_values, mgpu_comp_t comp,
                           mgpu_segsort_tuning_t* out, double* ms_table) {
  if (comp != MGPU_LESS && comp != MGPU_GREATER) return MGPU_INVALID_ARG;
  hipStream_t s = (hipStream_t)stream;
  segsort_tuning_t t;
  mgpu_status_t st;
  if (has_values) {
    st = comp == MGPU_LESS ? segsort_autotune<KT, uint32_t>(s, less_t<KT>(), t, ms_table)
                           : segsort_autotune<KT, uint32_t>(s, greater_t<KT>(), t, ms_table);
  } else {
    st = comp == MGPU_LESS ? segsort_autotune<KT, empty_t>(s, less_t<KT>(), t, ms_table)
                           : segsort_autotune<KT, empty_t>(s, greater_t<KT>(), t, ms_table);
  }
  if (st != MGPU_OK) return st;
  {
    segsort_tables_t& tb = seg_tables();
    std::lock_guard<std::mutex> lock(tb.mutex);
    tb.tables[{has_values ? 1 : 0, (int)comp}] = t;
  }
  if (out) {
    out->num_buckets = t.num_buckets;
    for (int b = 0; b < segsort_max_buckets; ++b) {
      out->cap[b] = t.cap[b];
      out->shape[b] = t.shape[b];
    }
  }
  return MGPU_OK;
}

}  // namespace

#define MGPU_CAT2(a, b) a##b
#define MGPU_CAT(a, b) MGPU_CAT2(a, b)

extern "C" mgpu_status_t MGPU_CAT(mgpu_segmented_sort_keys_, MGPU_SEGSORT_K)(
    void* aux, size_t* aux_size, mgpu_stream_t stream, MGPU_SEGSORT_KT* keys, int64_t count,
    const int32_t* segments, int64_t num_segments, mgpu_comp_t comp,
    const mgpu_segsort_tuning_t* t) {
  return run_segsort(aux, aux_size, stream, keys, (empty_t*)nullptr, count, segments,
                     num_segments, comp, t);
}

extern "C" mgpu_status_t MGPU_CAT(MGPU_CAT(mgpu_segmented_sort_kv_, MGPU_SEGSORT_K), _u32)(
    void* aux, size_t* aux_size, mgpu_stream_t stream, MGPU_SEGSORT_KT* keys, uint32_t* vals,
    int64_t count, const int32_t* segments, int64_t num_segments, mgpu_comp_t comp,
    const mgpu_segsort_tuning_t* t) {
  return run_segsort(aux, aux_size, stream, keys, vals, count, segments, num_segments, comp, t);
}

extern "C" mgpu_status_t MGPU_CAT(mgpu_segmented_sort_autotune_, MGPU_SEGSORT_K)(
    mgpu_stream_t stream, int has_values, mgpu_comp_t comp, mgpu_segsort_tuning_t* out,
    double* ms_table) {
  return run_autotune<MGPU_SEGSORT_KT>(stream, has_values, comp, out, ms_table);
}
