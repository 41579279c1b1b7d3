// A C++ template caller that begins a graph capture (global mode, the
// strictest: any allocation or synchronous memset in the process fails the
// capture) before its first call into the library, records a scan, a
// key-value radix sort and a segmented scan, then replays the graph three
// times.  Each replay is checked against the std:: algorithms on the host.
// Must run as a fresh process: no library call may precede the capture.
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <vector>

#include "mgpu/hip/radix.hxx"
#include "mgpu/hip/scan.hxx"
#include "mgpu/hip/segscan.hxx"

using namespace mgpu::hip;

#define CHECK_HIP(x)                                                             \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      std::printf("FAIL hip %s at line %d\n", hipGetErrorString(e_), __LINE__);  \
      return 1;                                                                  \
    }                                                                            \
  } while (0)
#define CHECK_OK(...)                                                            \
  do {                                                                           \
    mgpu_status_t s_ = (__VA_ARGS__);                                            \
    if (s_ != MGPU_OK) {                                                         \
      std::printf("FAIL status %d at line %d\n", (int)s_, __LINE__);            \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

int main() {
  const int64_t n = (1 << 22) + 123;
  std::vector<int32_t> x(n);
  std::vector<uint32_t> keys(n), vals(n);
  uint64_t z = 0x9E3779B97F4A7C15ull;
  for (int64_t i = 0; i < n; ++i) {
    z = z * 6364136223846793005ull + 1442695040888963407ull;
    x[i] = (int32_t)((z >> 33) & 15);
    keys[i] = (uint32_t)(z >> 32) & 0xFFFFF;  // ties, so stability shows
    vals[i] = (uint32_t)i;
  }
  std::vector<int32_t> segs;
  for (int64_t s = 0; s < n; s += 1000 + (s % 7777)) segs.push_back((int32_t)s);

  int32_t *d_x, *d_y, *d_sy, *d_segs;
  uint32_t *d_k, *d_v;
  CHECK_HIP(hipMalloc(&d_x, 4 * n));
  CHECK_HIP(hipMalloc(&d_y, 4 * n));
  CHECK_HIP(hipMalloc(&d_sy, 4 * n));
  CHECK_HIP(hipMalloc(&d_k, 4 * n));
  CHECK_HIP(hipMalloc(&d_v, 4 * n));
  CHECK_HIP(hipMalloc(&d_segs, 4 * segs.size()));
  CHECK_HIP(hipMemcpy(d_x, x.data(), 4 * n, hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d_segs, segs.data(), 4 * segs.size(), hipMemcpyHostToDevice));

  // Sizing calls read no data and touch no device state.
  size_t scan_aux = 0, radix_aux = 0, seg_aux = 0;
  CHECK_OK(scan(nullptr, scan_aux, nullptr, (const int32_t*)d_x, d_y, n, MGPU_SCAN_EXCLUSIVE, 0));
  CHECK_OK(radix_sort(nullptr, radix_aux, nullptr, d_k, d_v, n));
  CHECK_OK(segmented_scan(nullptr, seg_aux, nullptr, (const int32_t*)d_x, d_sy, n, d_segs,
                          (int64_t)segs.size(), MGPU_SCAN_INCLUSIVE, 0));
  void *a1, *a2, *a3;
  CHECK_HIP(hipMalloc(&a1, scan_aux));
  CHECK_HIP(hipMalloc(&a2, radix_aux));
  CHECK_HIP(hipMalloc(&a3, seg_aux));
  hipStream_t s;
  CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));

  std::printf("[case] capture before any library call\n");
  std::fflush(stdout);
  CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
  CHECK_OK(scan(a1, scan_aux, s, (const int32_t*)d_x, d_y, n, MGPU_SCAN_EXCLUSIVE, 0));
  CHECK_OK(radix_sort(a2, radix_aux, s, d_k, d_v, n));
  CHECK_OK(segmented_scan(a3, seg_aux, s, (const int32_t*)d_x, d_sy, n, d_segs,
                          (int64_t)segs.size(), MGPU_SCAN_INCLUSIVE, 0));
  hipGraph_t g;
  CHECK_HIP(hipStreamEndCapture(s, &g));
  hipGraphExec_t ge;
  CHECK_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));

  std::vector<int32_t> ey(n), esy(n);
  std::exclusive_scan(x.begin(), x.end(), ey.begin(), 0);
  {
    size_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
      if (k < segs.size() && segs[k] == i) {
        esy[i] = x[i];
        ++k;
      } else {
        esy[i] = esy[i - 1] + x[i];
      }
    }
  }
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });

  std::vector<int32_t> hy(n), hsy(n);
  std::vector<uint32_t> hk(n), hv(n);
  for (int rep = 0; rep < 3; ++rep) {
    CHECK_HIP(hipMemsetAsync(d_y, 0xFF, 4 * n, s));
    CHECK_HIP(hipMemsetAsync(d_sy, 0xFF, 4 * n, s));
    CHECK_HIP(hipMemcpyAsync(d_k, keys.data(), 4 * n, hipMemcpyHostToDevice, s));
    CHECK_HIP(hipMemcpyAsync(d_v, vals.data(), 4 * n, hipMemcpyHostToDevice, s));
    CHECK_HIP(hipGraphLaunch(ge, s));
    CHECK_HIP(hipStreamSynchronize(s));
    CHECK_HIP(hipMemcpy(hy.data(), d_y, 4 * n, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(hsy.data(), d_sy, 4 * n, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(hk.data(), d_k, 4 * n, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(hv.data(), d_v, 4 * n, hipMemcpyDeviceToHost));
    if (hy != ey) { std::printf("FAIL scan replay %d\n", rep); return 1; }
    if (hsy != esy) { std::printf("FAIL segmented scan replay %d\n", rep); return 1; }
    for (int64_t i = 0; i < n; ++i)
      if (hv[i] != order[i] || hk[i] != keys[order[i]]) {
        std::printf("FAIL radix sort replay %d at %lld\n", rep, (long long)i);
        return 1;
      }
    std::printf("PASS replay %d\n", rep);
  }
  std::printf("ALL PASS\n");
  return 0;
}
