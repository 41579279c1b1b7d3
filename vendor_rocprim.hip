// Speed comparator for bench.py: AMD's rocPRIM (ROCm's own device-wide
// primitives, compiled here from its source headers) run on the same device
// buffers as the mgpu-hip primitives.  It is measurement only: nothing in
// the library calls it.  Each function follows the two-phase temporary
// storage protocol of rocPRIM (temp == NULL: *temp_bytes := requirement) and
// returns the hipError_t of the call.  The stable sorts (radix sort of pairs,
// merge sort of pairs, segmented radix sort of pairs) produce the unique
// stable order, so their outputs must equal mgpu-hip's bit for bit.
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_merge.hpp>
#include <rocprim/device/device_merge_sort.hpp>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>
#include <rocprim/rocprim_version.hpp>

extern "C" {

// rocPRIM version: major * 100000 + minor * 100 + patch.
int vendor_rocprim_version(void) { return ROCPRIM_VERSION; }

// Stable LSD radix sort of (key, value) pairs, all 32 key bits.  The
// double-buffer form sorts in place over two buffer pairs (the same memory
// footprint as mgpu-hip's in-place sort with its ping-pong aux); *in_alt is
// set to 1 when the result ended in the alternate buffers.
hipError_t vendor_radix_sort_pairs_u32(void* temp, size_t* temp_bytes, uint32_t* keys,
                                       uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                                       int64_t n, hipStream_t stream, int* in_alt) {
  rocprim::double_buffer<uint32_t> k(keys, keys_alt);
  rocprim::double_buffer<uint32_t> v(vals, vals_alt);
  hipError_t e = rocprim::radix_sort_pairs(temp, *temp_bytes, k, v, (size_t)n, 0u, 32u, stream);
  if (in_alt) *in_alt = k.current() == keys_alt ? 1 : 0;
  return e;
}

// The same for 64-bit keys (all 64 bits).
hipError_t vendor_radix_sort_pairs_u64(void* temp, size_t* temp_bytes, uint64_t* keys,
                                       uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                                       int64_t n, hipStream_t stream, int* in_alt) {
  rocprim::double_buffer<uint64_t> k(keys, keys_alt);
  rocprim::double_buffer<uint32_t> v(vals, vals_alt);
  hipError_t e = rocprim::radix_sort_pairs(temp, *temp_bytes, k, v, (size_t)n, 0u, 64u, stream);
  if (in_alt) *in_alt = k.current() == keys_alt ? 1 : 0;
  return e;
}

// Keys only, 32-bit, double buffer.
hipError_t vendor_radix_sort_keys_u32(void* temp, size_t* temp_bytes, uint32_t* keys,
                                      uint32_t* keys_alt, int64_t n, hipStream_t stream,
                                      int* in_alt) {
  rocprim::double_buffer<uint32_t> k(keys, keys_alt);
  hipError_t e = rocprim::radix_sort_keys(temp, *temp_bytes, k, (size_t)n, 0u, 32u, stream);
  if (in_alt) *in_alt = k.current() == keys_alt ? 1 : 0;
  return e;
}

// Exclusive scan of uint64 (modulo 2^64) with an initial value, out of place.
hipError_t vendor_exclusive_scan_u64(void* temp, size_t* temp_bytes, const uint64_t* in,
                                     uint64_t* out, int64_t n, uint64_t init,
                                     hipStream_t stream) {
  return rocprim::exclusive_scan(temp, *temp_bytes, in, out, init, (size_t)n,
                                 rocprim::plus<uint64_t>(), stream);
}

// Stable merge sort of (key, value) pairs with operator<, out of place.
hipError_t vendor_merge_sort_pairs_u32(void* temp, size_t* temp_bytes, const uint32_t* keys_in,
                                       uint32_t* keys_out, const uint32_t* vals_in,
                                       uint32_t* vals_out, int64_t n, hipStream_t stream) {
  return rocprim::merge_sort(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n,
                             rocprim::less<uint32_t>(), stream);
}

// Stable merge of two ascending key arrays (ties take the first input).
hipError_t vendor_merge_u32(void* temp, size_t* temp_bytes, const uint32_t* a, int64_t na,
                            const uint32_t* b, int64_t nb, uint32_t* c, hipStream_t stream) {
  return rocprim::merge(temp, *temp_bytes, a, b, c, (size_t)na, (size_t)nb,
                        rocprim::less<uint32_t>(), stream);
}

// Exclusive scan of int32 with plus and an initial value, out of place.
hipError_t vendor_exclusive_scan_i32(void* temp, size_t* temp_bytes, const int32_t* in,
                                     int32_t* out, int64_t n, int32_t init, hipStream_t stream) {
  return rocprim::exclusive_scan(temp, *temp_bytes, in, out, init, (size_t)n,
                                 rocprim::plus<int32_t>(), stream);
}

// Inclusive scan of float32 with plus, out of place (float32 accumulation).
hipError_t vendor_inclusive_scan_f32(void* temp, size_t* temp_bytes, const float* in, float* out,
                                     int64_t n, hipStream_t stream) {
  return rocprim::inclusive_scan(temp, *temp_bytes, in, out, (size_t)n, rocprim::plus<float>(),
                                 stream);
}

// Sum of int32 (modulo 2^32) into *out (device).
hipError_t vendor_reduce_i32(void* temp, size_t* temp_bytes, const int32_t* in, int64_t n,
                             int32_t* out, hipStream_t stream) {
  return rocprim::reduce(temp, *temp_bytes, in, out, 0, (size_t)n, rocprim::plus<int32_t>(),
                         stream);
}

// Stable radix sort of (key, value) pairs inside every segment; segment i is
// [begin[i], end[i]).
hipError_t vendor_segmented_radix_sort_pairs_u32(void* temp, size_t* temp_bytes,
                                                 const uint32_t* keys_in, uint32_t* keys_out,
                                                 const uint32_t* vals_in, uint32_t* vals_out,
                                                 int64_t n, int64_t num_segments,
                                                 const int32_t* begin, const int32_t* end,
                                                 hipStream_t stream) {
  return rocprim::segmented_radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in,
                                             vals_out, (unsigned int)n,
                                             (unsigned int)num_segments, begin, end, 0u, 32u,
                                             stream);
}

}  // extern "C"
