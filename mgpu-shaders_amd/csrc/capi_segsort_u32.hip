// C ABI: segmented sort, u32 keys.
#define MGPU_SEGSORT_K u32
#define MGPU_SEGSORT_KT uint32_t
#include "segsort_impl.hxx"
