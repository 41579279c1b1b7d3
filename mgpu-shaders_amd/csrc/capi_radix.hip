// C ABI: LSD radix sort (inc/mgpu/vk/radix.hxx:195-470 counterpart, plus values).
#include "radix_dispatch.hxx"

using namespace mgpu::hip;

#define MGPU_DEF_RADIX(K, KT)                                                               \
  extern "C" mgpu_status_t mgpu_radix_sort_keys_##K(void* aux, size_t* aux_size,            \
                                                    mgpu_stream_t stream, KT* keys,         \
                                                    int64_t count) {                        \
    if (!aux_size) return MGPU_INVALID_ARG;                                                 \
    return radix_sort_tuned(aux, *aux_size, (hipStream_t)stream, keys, (empty_t*)nullptr,      \
                            count);                                                         \
  }                                                                                         \
  extern "C" mgpu_status_t mgpu_radix_sort_kv_##K##_u32(void* aux, size_t* aux_size,        \
                                                        mgpu_stream_t stream, KT* keys,     \
                                                        uint32_t* vals, int64_t count) {    \
    if (!aux_size) return MGPU_INVALID_ARG;                                                 \
    return radix_sort_tuned(aux, *aux_size, (hipStream_t)stream, keys, vals, count);          \
  }

MGPU_DEF_RADIX(u32, uint32_t)
MGPU_DEF_RADIX(i32, int32_t)
MGPU_DEF_RADIX(u64, uint64_t)
MGPU_DEF_RADIX(f32, float)
MGPU_DEF_RADIX(i64, int64_t)
MGPU_DEF_RADIX(f64, double)

extern "C" void mgpu_set_radix_algorithm(int algorithm) {
  radix_algorithm_setting().store(algorithm < 0 || algorithm > 2 ? 0 : algorithm);
}
extern "C" int mgpu_radix_algorithm(void) { return radix_algorithm_setting().load(); }
