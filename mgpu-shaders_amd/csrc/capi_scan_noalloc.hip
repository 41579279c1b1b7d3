// C ABI: the int32 scan for callers that cannot allocate (no counterpart in
// the reference).  Same kernels and results as mgpu_scan_i32; the look-back
// status words live in a fixed device array instead of caller aux.
#include <mutex>
#include <vector>

#include "mgpu/hip/scan.hxx"
#include "tuning.hxx"

using namespace mgpu::hip;

namespace {

constexpr int SNT = scan_tuning<int32_t>::nt, SVT = scan_tuning<int32_t>::vt;
// One 64-bit status word per 20480-element tile: 1 MiB covers 131072 tiles,
// i.e. up to 2,684,354,560 elements per call.
constexpr int64_t noalloc_max_tiles = 131072;

// Scratch of the entry point: the look-back status words of one call.  Each
// GPU holds its own copy (at its own address) until the process ends.  In
// host code `scratch` only names the symbol; kernels get the device address
// that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) uint64_t scratch[noalloc_max_tiles];

struct scratch_table_t {
  std::mutex mutex;
  std::vector<uint64_t*> ptr;  // per device
};
scratch_table_t& scratch_table() {
  static scratch_table_t t;
  return t;
}

}  // namespace

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.  Called by mgpu_init and, if
// init was not called on this device, by the entry point's first call.
hipError_t mgpu_noalloc_scratch(uint64_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  scratch_table_t& t = scratch_table();
  std::lock_guard<std::mutex> lock(t.mutex);
  if ((int)t.ptr.size() <= dev) t.ptr.resize(dev + 1, nullptr);
  if (!t.ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    t.ptr[dev] = (uint64_t*)p;
  }
  *out = t.ptr[dev];
  return hipSuccess;
}

// Calls that overlap on one GPU share its scratch and race; calls on
// different GPUs do not.  The status words are cleared by a fill kernel on
// `stream` before the scan, so the previous call's contents do not matter.
extern "C" mgpu_status_t mgpu_scan_i32_noalloc(mgpu_stream_t stream, const int32_t* input,
                                               int32_t* output, int64_t count,
                                               mgpu_scan_type_t type, int32_t init) {
  if (count < 0) return MGPU_INVALID_ARG;
  if (type != MGPU_SCAN_EXCLUSIVE && type != MGPU_SCAN_INCLUSIVE) return MGPU_INVALID_ARG;
  if (count == 0) return MGPU_OK;
  if (!input || !output) return MGPU_INVALID_ARG;
  const int64_t num_tiles = div_up(count, (int64_t)(SNT * SVT));
  if (num_tiles > noalloc_max_tiles) return MGPU_UNSUPPORTED;
  uint64_t* status = nullptr;
  hipError_t e = mgpu_noalloc_scratch(&status);
  if (e != hipSuccess) return hip_status(e);
  e = fill_u32((uint32_t*)status, 2 * num_tiles, 0u, (hipStream_t)stream);
  if (e != hipSuccess) return hip_status(e);
  return hip_status(scan_launch_cleared<SNT, SVT>((hipStream_t)stream, input, output, count, type,
                                                  init, identity_t(), status,
                                                  lookback_config_existing()));
}
