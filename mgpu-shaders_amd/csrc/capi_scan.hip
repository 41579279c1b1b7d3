// C ABI: scan and reduce (inc/mgpu/vk/scan.hxx:10-115 counterpart).
#include "mgpu/hip/scan.hxx"
#include "tuning.hxx"

using namespace mgpu::hip;

#define MGPU_DEF_SCAN(T, TT)                                                          \
  extern "C" mgpu_status_t mgpu_scan_exclusive_##T(void* aux, size_t* aux_size,       \
                                                   mgpu_stream_t stream, TT* data,    \
                                                   int64_t count, TT init) {          \
    if (!aux_size) return MGPU_INVALID_ARG;                                           \
    constexpr int NT = scan_tuning<TT>::nt, VT = scan_tuning<TT>::vt;                 \
    return scan<NT, VT>(aux, *aux_size, (hipStream_t)stream, (const TT*)data, data,   \
                        count, MGPU_SCAN_EXCLUSIVE, init);                            \
  }                                                                                   \
  extern "C" mgpu_status_t mgpu_scan_##T(void* aux, size_t* aux_size,                 \
                                         mgpu_stream_t stream, const TT* input,       \
                                         TT* output, int64_t count,                   \
                                         mgpu_scan_type_t type, TT init) {            \
    if (!aux_size) return MGPU_INVALID_ARG;                                           \
    constexpr int NT = scan_tuning<TT>::nt, VT = scan_tuning<TT>::vt;                 \
    return scan<NT, VT>(aux, *aux_size, (hipStream_t)stream, input, output, count,    \
                        type, init);                                                  \
  }                                                                                   \
  extern "C" mgpu_status_t mgpu_reduce_##T(void* aux, size_t* aux_size,               \
                                           mgpu_stream_t stream, const TT* input,     \
                                           int64_t count, TT* out, TT init) {         \
    if (!aux_size) return MGPU_INVALID_ARG;                                           \
    return reduce<256>(aux, *aux_size, (hipStream_t)stream, input, count, out, init); \
  }

MGPU_DEF_SCAN(i32, int32_t)
MGPU_DEF_SCAN(u32, uint32_t)
MGPU_DEF_SCAN(f32, float)

extern "C" mgpu_status_t mgpu_scan_u64(void* aux, size_t* aux_size, mgpu_stream_t stream,
                                       const uint64_t* input, uint64_t* output, int64_t count,
                                       mgpu_scan_type_t type, uint64_t init) {
  if (!aux_size) return MGPU_INVALID_ARG;
  return scan64<uint64_t>(aux, *aux_size, (hipStream_t)stream, input, output, count, type, init);
}

extern "C" mgpu_status_t mgpu_scan_exclusive_u64(void* aux, size_t* aux_size,
                                                 mgpu_stream_t stream, uint64_t* data,
                                                 int64_t count, uint64_t init) {
  if (!aux_size) return MGPU_INVALID_ARG;
  return scan64<uint64_t>(aux, *aux_size, (hipStream_t)stream, (const uint64_t*)data, data, count,
                          MGPU_SCAN_EXCLUSIVE, init);
}

extern "C" mgpu_status_t mgpu_reduce_u64(void* aux, size_t* aux_size, mgpu_stream_t stream,
                                         const uint64_t* input, int64_t count, uint64_t* out,
                                         uint64_t init) {
  if (!aux_size) return MGPU_INVALID_ARG;
  return reduce<256>(aux, *aux_size, (hipStream_t)stream, input, count, out, init);
}

// Segmented scan (no counterpart in the reference).
#include "mgpu/hip/segscan.hxx"

#define MGPU_DEF_SEGSCAN(T, TT)                                                             \
  extern "C" mgpu_status_t mgpu_segmented_scan_##T(                                         \
      void* aux, size_t* aux_size, mgpu_stream_t stream, const TT* input, TT* output,       \
      int64_t count, const int32_t* segments, int64_t num_segments, mgpu_scan_type_t type,  \
      TT init) {                                                                            \
    if (!aux_size) return MGPU_INVALID_ARG;                                                 \
    return segmented_scan<segscan_tuning<TT>::nt, segscan_tuning<TT>::vt>(                   \
        aux, *aux_size, (hipStream_t)stream, input, output, count, segments, num_segments,   \
        type, init);                                                                        \
  }

MGPU_DEF_SEGSCAN(i32, int32_t)
MGPU_DEF_SEGSCAN(u32, uint32_t)
MGPU_DEF_SEGSCAN(f32, float)
