// C ABI: segmented sort, f64 keys.
#define MGPU_SEGSORT_K f64
#define MGPU_SEGSORT_KT double
#include "segsort_impl.hxx"
