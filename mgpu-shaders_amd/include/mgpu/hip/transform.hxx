// Kernel launch helpers for user element-wise steps.
//
// Reference: vk::launch / vk::transform (inc/mgpu/vk/transform.hxx:9-36):
// launch<nt>(num_blocks, cmd, func) calls func(tid, cta) on every thread;
// transform<nt>(count, cmd, func) calls func(gid) for gid < count.  The same
// calls here take a hipStream_t in place of the command buffer, accept
// 64-bit counts and return a status; `func` is any device-callable object
// (a __device__ lambda or a functor), captured by value.  The reference's
// second template parameter, subgroup_size, is kept: -1 (any) or 64, the
// wave width here.
#pragma once
#include "meta.hxx"

namespace mgpu {
namespace hip {

template <int NT, typename func_t>
__global__ __launch_bounds__(NT) void kernel_launch(func_t func) {
  func((int)threadIdx.x, (int)blockIdx.x);
}

template <int NT, typename func_t>
__global__ __launch_bounds__(NT) void kernel_transform(int64_t count, func_t func) {
  const int64_t gid = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (gid < count) func(gid);
}

// func(tid, cta) for tid < NT, cta < num_blocks.
template <int NT = 256, int subgroup_size = -1, typename func_t>
mgpu_status_t launch(int num_blocks, hipStream_t stream, func_t func,
                     const char* name = "mgpu_launch") {
  static_assert(subgroup_size == -1 || subgroup_size == warp_size, "waves are 64 lanes wide");
  if (num_blocks < 0) return MGPU_INVALID_ARG;
  if (num_blocks == 0) return MGPU_OK;
  return hip_status(launch_kernel(name, kernel_launch<NT, func_t>, dim3((unsigned)num_blocks),
                                  dim3(NT), 0, stream, func));
}

// func(gid) for gid < count.
template <int NT = 256, int subgroup_size = -1, typename func_t>
mgpu_status_t transform(int64_t count, hipStream_t stream, func_t func,
                        const char* name = "mgpu_transform") {
  static_assert(subgroup_size == -1 || subgroup_size == warp_size, "waves are 64 lanes wide");
  if (count < 0) return MGPU_INVALID_ARG;
  if (count == 0) return MGPU_OK;
  const int64_t blocks = div_up(count, (int64_t)NT);
  if (blocks >= ((int64_t)1 << 31)) return MGPU_UNSUPPORTED;
  return hip_status(launch_kernel(name, kernel_transform<NT, func_t>, dim3((unsigned)blocks),
                                  dim3(NT), 0, stream, count, func));
}

// A plain device-to-device copy, the practical read + write ceiling of the
// device that the memory-bound kernels are measured against (no counterpart
// in the reference).  Each block moves one tile of NT * WPT 16-byte words
// (16 KB at 256 x 4: non-temporal loads, plain stores, all loads of a full
// tile in flight together); the bytes past the last whole word are copied one
// per thread by the last block.  On MI355X 16 KB tiles were the fastest of 16,
// 32 and 64 KB, and non-temporal stores were slower than plain ones.
template <int NT, int WPT>
__global__ __launch_bounds__(NT) void kernel_copy16(const uint8_t* src, uint8_t* dst,
                                                    int64_t words, int64_t bytes) {
  typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
  const u32x4* s = (const u32x4*)src;
  u32x4* d = (u32x4*)dst;
  const int64_t base = (int64_t)blockIdx.x * (NT * WPT) + threadIdx.x;
  u32x4 x[WPT];
  if (base + (int64_t)(WPT - 1) * NT < words) {
#pragma unroll
    for (int i = 0; i < WPT; ++i) x[i] = __builtin_nontemporal_load(s + base + i * NT);
#pragma unroll
    for (int i = 0; i < WPT; ++i) d[base + i * NT] = x[i];
  } else {
#pragma unroll
    for (int i = 0; i < WPT; ++i)
      if (base + i * NT < words) d[base + i * NT] = s[base + i * NT];
  }
  if (blockIdx.x == gridDim.x - 1)
    for (int64_t b = 16 * words + threadIdx.x; b < bytes; b += NT) dst[b] = src[b];
}

// Copies `bytes` from src to dst (device memory, both 16-byte aligned, not
// overlapping) on `stream`.
inline mgpu_status_t copy_bytes(void* dst, const void* src, int64_t bytes, hipStream_t stream) {
  constexpr int NT = 256, WPT = 4;
  if (bytes < 0) return MGPU_INVALID_ARG;
  if (bytes == 0) return MGPU_OK;
  if (!dst || !src || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return MGPU_INVALID_ARG;
  const int64_t words = bytes / 16;
  const int64_t blocks = max_<int64_t>(1, div_up(words, (int64_t)(NT * WPT)));
  if (blocks >= ((int64_t)1 << 31)) return MGPU_UNSUPPORTED;
  return hip_status(launch_kernel("mgpu_copy", kernel_copy16<NT, WPT>, dim3((unsigned)blocks),
                                  dim3(NT), 0, stream, (const uint8_t*)src, (uint8_t*)dst, words,
                                  bytes));
}

}  // namespace hip
}  // namespace mgpu
