// C ABI: segmented sort, f32 keys.
#define MGPU_SEGSORT_K f32
#define MGPU_SEGSORT_KT float
#include "segsort_impl.hxx"
