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

}  // namespace hip
}  // namespace mgpu
