// C ABI: segmented sort, u64 keys.
#define MGPU_SEGSORT_K u64
#define MGPU_SEGSORT_KT uint64_t
#include "segsort_impl.hxx"
