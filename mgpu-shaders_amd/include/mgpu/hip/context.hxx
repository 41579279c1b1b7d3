// Device, stream and memory owner with the reference's host call surface
// (inc/mgpu/vk/context.hxx:11-134: context_t, cmd_buffer_t, memcache_t;
// src/vk/context.cxx:156-212 alloc_gpu / alloc_cpu / free, :349-404 command
// buffer, :409-439 memcache_t), so that call sites written against it port
// with their allocation code unchanged:
//
//   context_t context;
//   float* gpu = context.alloc_gpu<float>(count);
//   cmd_buffer_t cmd_buffer(context);
//   cmd_buffer.begin();
//   mergesort_keys(nullptr, aux_size, cmd_buffer, gpu, count);   // sizing
//   void* aux = context.alloc_gpu(aux_size);
//   mergesort_keys(aux, aux_size, cmd_buffer, gpu, count);       // enqueue
//   cmd_buffer.end();
//   context.submit(cmd_buffer);
//   context.wait_idle();                    // the reference's vkQueueWaitIdle
//
// Here a command buffer is the context's HIP stream: every primitive takes a
// hipStream_t, to which both context_t and cmd_buffer_t convert, and work
// runs as it is enqueued, in order.  So begin / end / reset / submit are
// no-ops kept for the call sites, host_barrier (a pipeline barrier in the
// reference) is implied by stream order, and wait_idle waits for the stream.
// Allocation failures return nullptr; last_error() keeps the hipError_t.
#pragma once
#include <array>
#include <map>
#include <mutex>

#include "meta.hxx"

namespace mgpu {
namespace hip {

struct context_t {
  // Uses `device` (the current device if negative) and, if `stream` is null,
  // a non-blocking stream of its own (default priority), destroyed with it.
  explicit context_t(int device = -1, hipStream_t stream = nullptr) {
    if (device >= 0) error = hipSetDevice(device);
    (void)hipGetDevice(&this->device);
    if (stream) {
      this->stream = stream;
    } else {
      hipError_t e = hipStreamCreateWithFlags(&this->stream, hipStreamNonBlocking);
      if (e != hipSuccess) error = e;
      owns_stream = e == hipSuccess;
    }
  }
  ~context_t() {
    (void)hipStreamSynchronize(stream);
    for (auto& b : buffers) {
      void* p = const_cast<void*>(b.first);
      (void)(b.second.cpu ? hipHostFree(p) : hipFree(p));
    }
    if (owns_stream) (void)hipStreamDestroy(stream);
  }
  context_t(const context_t&) = delete;
  context_t& operator=(const context_t&) = delete;

  operator hipStream_t() const noexcept { return stream; }

  // Device memory (at least one byte, like the reference's alloc_gpu); the
  // Vulkan usage flags of the reference's signature are accepted and unused.
  void* alloc_gpu(size_t size, uint32_t usage = 0) {
    (void)usage;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, size ? size : 1);
    return record(e, p, size, false);
  }
  template <typename type_t>
  type_t* alloc_gpu(size_t count, uint32_t usage = 0) {
    return (type_t*)alloc_gpu(sizeof(type_t) * count, usage);
  }
  // Pinned host memory, readable and writable by the host, usable as a
  // memcpy source or destination on the stream.
  void* alloc_cpu(size_t size, uint32_t usage = 0) {
    (void)usage;
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, size ? size : 1, hipHostMallocDefault);
    return record(e, p, size, true);
  }
  template <typename type_t>
  type_t* alloc_cpu(size_t count, uint32_t usage = 0) {
    return (type_t*)alloc_cpu(sizeof(type_t) * count, usage);
  }
  // Frees a buffer this context allocated (null is a no-op).  The stream's
  // work that uses it must have completed.
  void free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(mutex);
    auto it = buffers.find(p);
    if (it == buffers.end()) {
      error = hipErrorInvalidValue;
      return;
    }
    hipError_t e = it->second.cpu ? hipHostFree(p) : hipFree(p);
    if (e != hipSuccess) error = e;
    buffers.erase(it);
  }
  // Bytes of the buffer this context allocated at p (0 if none).
  size_t size_of(const void* p) {
    std::lock_guard<std::mutex> lock(mutex);
    auto it = buffers.find(p);
    return it == buffers.end() ? 0 : it->second.size;
  }

  // Copy on the stream, in order with the primitives enqueued on it.
  void memcpy(void* dest, const void* source, size_t size) {
    hipError_t e = hipMemcpyAsync(dest, source, size, hipMemcpyDefault, stream);
    if (e != hipSuccess) error = e;
  }
  void submit(hipStream_t) {}
  // Waits until everything enqueued on the stream has completed.
  hipError_t wait_idle() {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) error = e;
    return e;
  }
  hipError_t last_error() const { return error; }

  int device = 0;
  hipStream_t stream = nullptr;

 private:
  struct buffer_t { size_t size; bool cpu; };
  void* record(hipError_t e, void* p, size_t size, bool cpu) {
    if (e != hipSuccess) {
      error = e;
      return nullptr;
    }
    std::lock_guard<std::mutex> lock(mutex);
    buffers[p] = buffer_t{size, cpu};
    return p;
  }
  bool owns_stream = false;
  hipError_t error = hipSuccess;
  std::mutex mutex;
  std::map<const void*, buffer_t> buffers;
};

// The reference's command buffer: recording and submission collapse into the
// context's stream.
struct cmd_buffer_t {
  explicit cmd_buffer_t(context_t& context) : context(context) {}
  operator hipStream_t() const noexcept { return context.stream; }
  void reset() {}
  void begin() {}
  void end() {}
  void submit() { context.submit(context.stream); }
  void host_barrier() {}
  void memcpy(void* dest, const void* source, size_t size) { context.memcpy(dest, source, size); }
  context_t& context;
};

// Grow-only scratch (src/vk/context.cxx:409-439): one device buffer, replaced
// by a larger one when a request exceeds it.  A request that grows it frees
// the old buffer, so work still using that buffer must have completed.
struct memcache_t {
  explicit memcache_t(context_t& context) : context(context) {}
  ~memcache_t() { context.free(data); }
  memcache_t(const memcache_t&) = delete;
  memcache_t& operator=(const memcache_t&) = delete;

  void* allocate(size_t size) {
    if (size > capacity || !data) {
      context.free(data);
      data = context.alloc_gpu(size);
      capacity = data ? size : 0;
    }
    return data;
  }
  template <typename type_t>
  type_t* allocate(size_t count) {
    return (type_t*)allocate(sizeof(type_t) * count);
  }
  // Several regions in one buffer, each starting on a 256-byte boundary.
  void allocate(const size_t* sizes, int count, void** allocations) {
    size_t total = 0;
    for (int i = 0; i < count; ++i) total += (sizes[i] + 255) & ~(size_t)255;
    char* p = (char*)allocate(total);
    for (int i = 0; i < count; ++i) {
      allocations[i] = p;
      if (p) p += (sizes[i] + 255) & ~(size_t)255;
    }
  }
  template <size_t count>
  std::array<void*, count> allocate(std::array<size_t, count> sizes) {
    std::array<void*, count> out;
    allocate(sizes.data(), (int)count, out.data());
    return out;
  }

  context_t& context;
  void* data = nullptr;
  size_t capacity = 0;
};

}  // namespace hip
}  // namespace mgpu
