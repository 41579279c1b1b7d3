// C ABI: segmented sort, i64 keys.
#define MGPU_SEGSORT_K i64
#define MGPU_SEGSORT_KT int64_t
#include "segsort_impl.hxx"
