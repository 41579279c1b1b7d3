// Common helpers for the gfx950 primitives: integer helpers, comparators,
// key <-> radix-bit maps, aux-buffer carving, kernel launch with optional
// event timing, and wave64 primitives (ballot, DPP scans).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "mgpu_hip.h"

#ifndef __HIP_PLATFORM_AMD__
#error "mgpu-hip targets AMD gfx950 only"
#endif

namespace mgpu {
namespace hip {

constexpr int warp_size = 64;

struct empty_t {};

// div_up / is_pow2 / find_log2: inc/mgpu/common/meta.hxx:20-41, extended to
// 64-bit and with find_log2(0) defined as 0 (the reference's is undefined).
template <typename a_t, typename b_t>
__host__ __device__ constexpr a_t div_up(a_t x, b_t y) {
  return (x + (a_t)y - 1) / (a_t)y;
}
__host__ __device__ constexpr bool is_pow2(int64_t x) {
  return x > 0 && 0 == (x & (x - 1));
}
__host__ __device__ constexpr int find_log2(int64_t x, bool round_up = false) {
  if (x <= 0) return 0;
  int a = 63 - __builtin_clzll((unsigned long long)x);
  if (round_up) a += !is_pow2(x);
  return a;
}
template <typename T>
__host__ __device__ constexpr T min_(T a, T b) { return a < b ? a : b; }
template <typename T>
__host__ __device__ constexpr T max_(T a, T b) { return a < b ? b : a; }

// Standard comparators (the reference defaults to std::less<key_t>,
// inc/mgpu/vk/mergesort.hxx:11).
template <typename T>
struct less_t {
  __host__ __device__ bool operator()(T a, T b) const { return a < b; }
};
template <typename T>
struct greater_t {
  __host__ __device__ bool operator()(T a, T b) const { return b < a; }
};

////////////////////////////////////////////////////////////////////////////////
// Key <-> order-preserving unsigned bits, as radix_permute_t in
// inc/mgpu/common/cta_radix.hxx:6-64 (the signed flip uses a shift of the
// unsigned type, so 64-bit signed keys are well defined).

template <int size> struct uint_by_size;
template <> struct uint_by_size<4> { typedef uint32_t type; typedef int32_t stype; };
template <> struct uint_by_size<8> { typedef uint64_t type; typedef int64_t stype; };

template <typename key_t>
struct radix_permute_t {
  typedef typename uint_by_size<sizeof(key_t)>::type unsigned_type;
  typedef typename uint_by_size<sizeof(key_t)>::stype signed_type;
  static constexpr unsigned_type msb = (unsigned_type)1 << (8 * sizeof(key_t) - 1);

  __host__ __device__ static unsigned_type to_radix_bits(key_t x) {
    unsigned_type y;
    __builtin_memcpy(&y, &x, sizeof(y));
    if constexpr (std::is_unsigned_v<key_t>) {
      return y;
    } else if constexpr (std::is_integral_v<key_t>) {
      return y ^ msb;
    } else {
      unsigned_type mask = (unsigned_type)((signed_type)y >> (8 * sizeof(key_t) - 1)) | msb;
      return y ^ mask;
    }
  }
  __host__ __device__ static key_t from_radix_bits(unsigned_type y) {
    if constexpr (std::is_unsigned_v<key_t>) {
      key_t x; __builtin_memcpy(&x, &y, sizeof(x)); return x;
    } else if constexpr (std::is_integral_v<key_t>) {
      y ^= msb;
      key_t x; __builtin_memcpy(&x, &y, sizeof(x)); return x;
    } else {
      y ^= msb;
      unsigned_type mask = (unsigned_type)((signed_type)y >> (8 * sizeof(key_t) - 1)) & (msb - 1);
      y ^= mask;
      key_t x; __builtin_memcpy(&x, &y, sizeof(x)); return x;
    }
  }
};

////////////////////////////////////////////////////////////////////////////////
// Aux carving (advance_pointer, inc/mgpu/common/meta.hxx:119-124).  Every
// region is rounded up to 256 bytes so that 128-bit loads stay aligned.

constexpr size_t aux_align = 256;
__host__ inline size_t aux_round(size_t bytes) {
  return (bytes + aux_align - 1) / aux_align * aux_align;
}
template <typename T>
__host__ inline size_t aux_bytes(int64_t count) {
  return aux_round(sizeof(T) * (size_t)(count > 0 ? count : 0));
}
template <typename T>
__host__ inline T* advance_pointer(void*& p, int64_t count) {
  T* r = (T*)p;
  p = (char*)p + aux_bytes<T>(count);
  return r;
}
// The enqueueing call's check of the caller's aux: capacity first, then
// alignment.  Regions are carved at 256-byte offsets from aux and hold
// status words used with 64-bit atomics and 128-bit loads, so aux must be
// 256-byte aligned (as hipMalloc returns) whenever any bytes are needed.
__host__ inline mgpu_status_t check_aux(const void* aux, size_t aux_size, size_t need) {
  if (aux_size < need) return MGPU_AUX_TOO_SMALL;
  if (need && (uintptr_t)aux % aux_align) return MGPU_INVALID_ARG;
  return MGPU_OK;
}

////////////////////////////////////////////////////////////////////////////////
// Per-thread last HIP error (reported through mgpu_last_hip_error).

inline int& last_hip_error_slot() {
  static thread_local int e = 0;
  return e;
}
inline mgpu_status_t hip_status(hipError_t e) {
  if (e == hipSuccess) return MGPU_OK;
  last_hip_error_slot() = (int)e;
  return MGPU_HIP_ERROR;
}

////////////////////////////////////////////////////////////////////////////////
// Event timing of every launch while enabled.

struct timing_registry_t {
  std::mutex mutex;
  std::atomic<bool> enabled{false};
  struct rec_t { const char* name; hipEvent_t start, stop; };
  std::vector<rec_t> recs;
  std::vector<hipEvent_t> pool;
  std::vector<std::string> names;  // storage for collected names

  hipEvent_t get_event() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
};
inline timing_registry_t& timing_registry() {
  static timing_registry_t r;
  return r;
}

template <typename kernel_t, typename... args_t>
inline hipError_t launch_kernel(const char* name, kernel_t kernel, dim3 grid,
                                dim3 block, size_t shmem, hipStream_t stream,
                                args_t... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return hipSuccess;
  timing_registry_t& tr = timing_registry();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // Read once: a toggle from another thread must not split a launch's pair.
  const bool timed = tr.enabled.load(std::memory_order_relaxed);
  if (timed) {
    std::lock_guard<std::mutex> lock(tr.mutex);
    e0 = tr.get_event();
    e1 = tr.get_event();
    (void)hipEventRecord(e0, stream);
  }
  hipLaunchKernelGGL(kernel, grid, block, shmem, stream, args...);
  hipError_t err = hipGetLastError();
  if (timed) {
    std::lock_guard<std::mutex> lock(tr.mutex);
    (void)hipEventRecord(e1, stream);
    tr.recs.push_back({name, e0, e1});
  }
  return err;
}

////////////////////////////////////////////////////////////////////////////////
// Look-back settings.  Tiles are numbered by blockIdx.x (no atomic ticket: a
// single-address device atomic costs ~11 ns serialised on MI355X, which would
// dominate a pass).  A tile waits for a predecessor's word at most
// `spin_limit` polls; after that it computes the predecessor's contribution
// from the predecessor's input itself, so progress never depends on the order
// in which workgroups are dispatched.  `fallbacks` (device, per device) counts
// how often that happened.

struct lookback_config_t {
  uint32_t spin_limit;
  uint32_t* fallbacks;
};

struct lookback_settings_t {
  std::mutex mutex;
  uint32_t spin_limit = 512;
  std::vector<uint32_t*> counters;  // per device
};
inline lookback_settings_t& lookback_settings() {
  static lookback_settings_t s;
  return s;
}
// The look-back settings for a launch on `stream`, on the current device.
// The fallback counter of a device is allocated and zeroed on its first use
// (or by init()); a launch recorded while `stream` is being captured into a
// graph never allocates: if the counter does not exist yet, that launch runs
// without one (it only loses the count), so a caller may begin capturing
// before any other call into the library.
inline lookback_config_t lookback_config(hipStream_t stream = nullptr) {
  lookback_settings_t& s = lookback_settings();
  std::lock_guard<std::mutex> lock(s.mutex);
  int dev = 0;
  (void)hipGetDevice(&dev);
  if ((int)s.counters.size() <= dev) s.counters.resize(dev + 1, nullptr);
  if (!s.counters[dev]) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (stream && hipStreamIsCapturing(stream, &cs) == hipSuccess &&
        cs != hipStreamCaptureStatusNone)
      return lookback_config_t{s.spin_limit, nullptr};
    uint32_t* p = nullptr;
    if (hipMalloc(&p, 256) == hipSuccess && hipMemset(p, 0, 256) == hipSuccess)
      s.counters[dev] = p;
  }
  return lookback_config_t{s.spin_limit, s.counters[dev]};
}

// The look-back settings without creating anything: the current device's
// fallback counter if it exists already, else none (the launch then only
// loses the count).  For callers that must not allocate.
inline lookback_config_t lookback_config_existing() {
  lookback_settings_t& s = lookback_settings();
  std::lock_guard<std::mutex> lock(s.mutex);
  int dev = 0;
  (void)hipGetDevice(&dev);
  uint32_t* c = dev < (int)s.counters.size() ? s.counters[dev] : nullptr;
  return lookback_config_t{s.spin_limit, c};
}

////////////////////////////////////////////////////////////////////////////////
// Side streams for independent launches inside one call (the segmented
// sort's size buckets and its large-segment radix touch disjoint segments):
// fork_join_t makes up to max_side_streams streams wait for the work already
// on the caller's stream, the call spreads its independent launches over
// them, and join makes the caller's stream wait for all of them.  Under
// graph capture the branches become parallel branches of the graph.  The
// streams (non-blocking, default priority) and events are created once per
// device, outside any capture; if they do not exist yet while the caller's
// stream is being captured, every branch is the caller's stream (the same
// work, in order).  One fork-join at a time uses a device's side streams.

constexpr int max_side_streams = 8;

struct side_streams_t {
  struct device_t {
    std::mutex use;
    hipEvent_t fork = nullptr;
    hipStream_t streams[max_side_streams] = {};
    hipEvent_t joins[max_side_streams] = {};
  };
  std::mutex mutex;
  std::vector<std::unique_ptr<device_t>> devices;
};
inline side_streams_t& side_streams() {
  static side_streams_t s;
  return s;
}

// The current device's side streams, created on first use unless `origin`
// is being captured (then nullptr).
inline side_streams_t::device_t* side_streams_for(hipStream_t origin) {
  side_streams_t& ss = side_streams();
  std::lock_guard<std::mutex> lock(ss.mutex);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  if ((int)ss.devices.size() <= dev) ss.devices.resize(dev + 1);
  if (ss.devices[dev]) return ss.devices[dev].get();
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (origin && hipStreamIsCapturing(origin, &cs) == hipSuccess &&
      cs != hipStreamCaptureStatusNone)
    return nullptr;
  std::unique_ptr<side_streams_t::device_t> d(new side_streams_t::device_t);
  bool ok = hipEventCreateWithFlags(&d->fork, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; ok && i < max_side_streams; ++i)
    ok = hipStreamCreateWithFlags(&d->streams[i], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&d->joins[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) return nullptr;  // (what was created stays until the process ends)
  ss.devices[dev] = std::move(d);
  return ss.devices[dev].get();
}

struct fork_join_t {
  // Branches 0 .. n-1 start after the work on `origin` so far.
  fork_join_t(hipStream_t origin, int n) : origin(origin) {
    if (n <= 1) return;
    dev = side_streams_for(origin);
    if (!dev) return;
    lock = std::unique_lock<std::mutex>(dev->use);
    n = n < max_side_streams ? n : max_side_streams;
    if (hipEventRecord(dev->fork, origin) != hipSuccess) return;
    for (int i = 0; i < n; ++i)
      if (hipStreamWaitEvent(dev->streams[i], dev->fork, 0) != hipSuccess) return;
    active = n;
  }
  // The stream of branch i (the caller's stream when not forked).
  hipStream_t branch(int i) const { return active ? dev->streams[i % active] : origin; }
  // The caller's stream waits for every branch.
  hipError_t join() {
    hipError_t e = hipSuccess;
    for (int i = 0; i < active; ++i) {
      hipError_t r = hipEventRecord(dev->joins[i], dev->streams[i]);
      if (r == hipSuccess) r = hipStreamWaitEvent(origin, dev->joins[i], 0);
      if (r != hipSuccess && e == hipSuccess) e = r;
    }
    active = 0;
    if (lock.owns_lock()) lock.unlock();
    return e;
  }
  ~fork_join_t() { (void)join(); }
  fork_join_t(const fork_join_t&) = delete;
  fork_join_t& operator=(const fork_join_t&) = delete;

 private:
  hipStream_t origin;
  side_streams_t::device_t* dev = nullptr;
  std::unique_lock<std::mutex> lock;
  int active = 0;
};

// Creates the library's per-device state on the current device (the
// look-back fallback counter, the side streams).  Optional: every entry
// point creates what it needs on first use outside a graph capture; call it
// once before capturing if the captured launches should count look-back
// fallbacks and a captured segmented sort should branch.
inline mgpu_status_t init() {
  if (!lookback_config().fallbacks) return MGPU_HIP_ERROR;
  return side_streams_for(nullptr) ? MGPU_OK : MGPU_HIP_ERROR;
}

////////////////////////////////////////////////////////////////////////////////
// Tiles through raw buffer instructions.  A buffer resource covers
// [p, p + bytes): loads past the end read 0 and stores past the end are
// dropped by the hardware's range check, so a tile's edge needs neither a
// branch per element nor a clamped address, and the per-element offsets are
// 32-bit (constant parts fold into the instruction's offset field).  Without
// branches around the stores the compiler also knows how many memory
// operations are in flight, so waiting for a load does not wait for later
// stores.  Word 3 (0x00020000) is the gfx9 raw-buffer format word.

__device__ inline __amdgpu_buffer_rsrc_t make_buffer(const void* p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// Elements of 4 or 8 bytes go through buffers; any other size takes the
// guarded global path.
template <typename T>
constexpr bool buffer_element = sizeof(T) == 4 || sizeof(T) == 8;

template <typename T>
__device__ inline T buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t offset) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)offset, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)offset, 0, 0));
}

template <typename T>
__device__ inline void buffer_store(T x, __amdgpu_buffer_rsrc_t r, uint32_t offset) {
  typedef __attribute__((ext_vector_type(2))) unsigned int u32x2;
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), r, (int)offset, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, (int)offset, 0, 0);
}

// x[i] = p[i * NT + tid] for the tile's count (>= 1) elements.  Slots past
// the end hold 0 (buffers) or a copy of the last element (other sizes).
template <int NT, int VT, typename T>
__device__ inline void tile_load_strided(const T* p, int count, T (&x)[VT]) {
  const int tid = threadIdx.x;
  if constexpr (buffer_element<T>) {
    const __amdgpu_buffer_rsrc_t r = make_buffer(p, (int64_t)count * sizeof(T));
#pragma unroll
    for (int i = 0; i < VT; ++i) x[i] = buffer_load<T>(r, (uint32_t)((i * NT + tid) * sizeof(T)));
  } else {
#pragma unroll
    for (int i = 0; i < VT; ++i) x[i] = p[min_(i * NT + tid, count - 1)];
  }
}

// p[i * NT + tid] = x[i] for the slots below count.
template <int NT, int VT, typename T>
__device__ inline void tile_store_strided(T* p, int count, const T (&x)[VT]) {
  const int tid = threadIdx.x;
  if constexpr (buffer_element<T>) {
    const __amdgpu_buffer_rsrc_t r = make_buffer(p, (int64_t)count * sizeof(T));
#pragma unroll
    for (int i = 0; i < VT; ++i) buffer_store<T>(x[i], r, (uint32_t)((i * NT + tid) * sizeof(T)));
  } else {
#pragma unroll
    for (int i = 0; i < VT; ++i)
      if (i * NT + tid < count) p[i * NT + tid] = x[i];
  }
}

////////////////////////////////////////////////////////////////////////////////
// wave64 primitives.

__device__ inline int lane_id() { return (int)__lane_id(); }

// Number of set bits of `mask` in lanes below the calling lane.
__device__ inline uint32_t mask_count_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Inclusive wave scan of a 32-bit value with DPP: row_shr 1,2,4,8 inside each
// row of 16 lanes, then row_bcast:15 and row_bcast:31 carry across rows.
// Lanes whose DPP source is out of range read the identity 0.
__device__ inline uint32_t wave_scan_incl_u32(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);
  return x;
}
__device__ inline float wave_scan_incl_f32(float x) {
#define MGPU_DPP_F(ctrl, rmask) \
  __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), ctrl, rmask, 0xf, false))
  x += MGPU_DPP_F(0x111, 0xf);
  x += MGPU_DPP_F(0x112, 0xf);
  x += MGPU_DPP_F(0x114, 0xf);
  x += MGPU_DPP_F(0x118, 0xf);
  x += MGPU_DPP_F(0x142, 0xa);
  x += MGPU_DPP_F(0x143, 0xc);
#undef MGPU_DPP_F
  return x;
}
__device__ inline uint32_t wave_scan_incl(uint32_t x) { return wave_scan_incl_u32(x); }
__device__ inline float wave_scan_incl(float x) { return wave_scan_incl_f32(x); }

// Value of the previous lane (lane 0 receives `first`).
template <typename T>
__device__ inline T wave_shift_up1(T x, T first) {
  T y = __shfl_up(x, 1, warp_size);
  return lane_id() == 0 ? first : y;
}

// Butterfly reduction over the wave (every lane receives the total).
template <typename T>
__device__ inline T wave_reduce_all(T x) {
#pragma unroll
  for (int offset = 1; offset < warp_size; offset <<= 1)
    x += __shfl_xor(x, offset, warp_size);
  return x;
}

// Lanes of the wave whose 8-bit digit equals the caller's (match_any over
// eight ballots), restricted to lanes in `active`.
__device__ inline uint64_t match_digit8(uint32_t digit, uint64_t active) {
  uint64_t m = active;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    bool bit = (digit >> b) & 1;
    uint64_t ballot = __ballot(bit);
    m &= bit ? ballot : ~ballot;
  }
  return m;
}

// Relaxed agent-scope atomics for the look-back state words.  Each word holds
// its payload and status together, so reading a word needs no ordering with
// other memory.  The one place where order matters is the in-place scan's
// input fallback (a tile summing a predecessor's input while that tile may
// overwrite it with output): see status_store_complete / acquire_agent.
template <typename T>
__device__ inline T atomic_load_agent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline void atomic_store_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline T atomic_add_agent(T* p, T v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Owner side of the in-place look-back: a tile's status word (an agent-scope
// store, written through to the point of coherence) must be visible before
// any later store of the tile's output.  s_waitcnt(0) holds the wave until
// the store is acknowledged, and the signal fence keeps the compiler from
// moving later stores above it.  (An agent-scope release fence orders them
// too, but on gfx950 it writes back the L2, once per tile.)
__device__ inline void status_store_complete() {
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// Reader side: the loads of a predecessor's input complete, and no stale
// cache line serves a later load, before its status word is read again.  If
// any element read was already the predecessor's output, its status store
// completed before that output was written, so the re-read sees it.
__device__ inline void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Zero-fill of 32-bit words (used instead of hipMemsetAsync so the
// schedule stays capturable in a graph and replays correctly).
template <int NT>
__global__ __launch_bounds__(NT) void kernel_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) p[i] = v;
}
inline hipError_t fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = min_<int64_t>(div_up(n, (int64_t)256), 4096);
  return launch_kernel("mgpu_fill", kernel_fill_u32<256>, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

}  // namespace hip
}  // namespace mgpu
