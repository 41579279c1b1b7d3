// C ABI: segmented sort, i32 keys.
#define MGPU_SEGSORT_K i32
#define MGPU_SEGSORT_KT int32_t
#include "segsort_impl.hxx"
