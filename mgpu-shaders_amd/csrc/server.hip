// Serving engine: continuous batching of sort / merge / scan requests.
//
// No counterpart in the reference (it is a library of primitives with no
// request loop).  One worker thread owns `slots` HIP streams.  A request is
// admitted as soon as a slot is free and its device bytes fit the HBM
// budget (first fit from the head of the queue, so small requests do not
// wait behind a large one that does not fit yet); its buffers come from a
// stream-ordered memory pool, its input is generated on the device, and its
// result stays device-resident until the client releases it.  Requests on
// different slots run concurrently on the GPU.  Completion is detected with
// HIP events; timestamps are host steady-clock milliseconds since the
// server was created.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "mgpu/hip/generate.hxx"
#include "mgpu/hip/merge.hxx"
#include "mgpu/hip/scan.hxx"
#include "radix_dispatch.hxx"
#include "tuning.hxx"

using namespace mgpu::hip;

namespace {

typedef std::chrono::steady_clock clock_t_;

struct rec_t {
  int64_t id = 0;
  mgpu_request_t req{};
  bool caller = false;            // buffers owned by the caller (bufs)
  mgpu_request_buffers_t bufs{};
  double t_submit = 0, t_start = 0, t_done = 0;
  int slot = -1;
  mgpu_status_t status = MGPU_OK;
  size_t bytes = 0;
  void* in_a = nullptr;   // keys (sort, scan input) or A keys (merge)
  void* in_b = nullptr;   // B keys (merge)
  void* out = nullptr;    // result keys / scan output
  void* vals = nullptr;   // sort / merge values (u32)
  void* vals_a = nullptr; // merge input values
  void* aux = nullptr;
  hipEvent_t done = nullptr;
  bool reported = false;
};

size_t key_bytes(int kt) { return kt == MGPU_KEY_U64 ? 8 : 4; }

}  // namespace

struct mgpu_server_st {
  std::mutex m;
  std::condition_variable cv;
  std::deque<rec_t*> pending;
  std::vector<rec_t*> slot_rec;
  std::vector<hipStream_t> streams;
  hipStream_t free_stream = nullptr;
  std::map<int64_t, rec_t*> recs;
  std::vector<rec_t*> completed;
  size_t budget = 0, in_use = 0, peak = 0;
  int64_t next_id = 0;
  int64_t in_flight = 0;
  bool stop = false;
  int keep_inputs = 1;
  int device = 0;
  clock_t_::time_point t0;
  std::thread worker;
  hipMemPool_t pool = nullptr;

  double now() const {
    return std::chrono::duration<double, std::milli>(clock_t_::now() - t0).count();
  }
};

namespace {

// Stable key-value merge of two ascending arrays of n keys, with the tuned
// tile of the key type (aux == nullptr: sizing).
template <typename K>
mgpu_status_t serve_merge(void* aux, size_t& cap, hipStream_t s, const void* a, const void* b,
                          const uint32_t* va, int64_t n, void* c, uint32_t* vc) {
  return merge<merge_tuning<K>::nt, merge_tuning<K>::vt>(
      aux, cap, s, (const K*)a, va, n, (const K*)b, va ? va + n : nullptr, n, (K*)c, vc,
      less_t<K>());
}

// Device bytes a request needs (buffers + aux), from the sizing calls.
size_t request_bytes(const mgpu_request_t& r) {
  const int64_t n = r.n;
  const size_t kb = key_bytes(r.key_type);
  size_t aux = 0;
  switch (r.op) {
    case MGPU_REQ_SORT:
      if (r.key_type == MGPU_KEY_U64)
        radix_sort_tuned(nullptr, aux, nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr, n);
      else
        radix_sort_tuned(nullptr, aux, nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, n);
      return 2 * aux_round(n * kb) + aux_round(n * 4) + aux_round(aux);
    case MGPU_REQ_MERGE:
      if (r.key_type == MGPU_KEY_U64)
        serve_merge<uint64_t>(nullptr, aux, nullptr, nullptr, nullptr, nullptr, n, nullptr, nullptr);
      else
        serve_merge<uint32_t>(nullptr, aux, nullptr, nullptr, nullptr, nullptr, n, nullptr, nullptr);
      return 2 * aux_round(n * kb) + aux_round(2 * n * 4) + aux_round(2 * n * kb) +
             aux_round(2 * n * 4) + aux_round(aux);
    case MGPU_REQ_SCAN:
      if (r.key_type == MGPU_KEY_U64)
        scan64<uint64_t>(nullptr, aux, nullptr, nullptr, nullptr, n, MGPU_SCAN_EXCLUSIVE, 0ull);
      else
        scan<scan_tuning<uint32_t>::nt, scan_tuning<uint32_t>::vt>(nullptr, aux, nullptr, (const uint32_t*)nullptr, (uint32_t*)nullptr, n,
                      MGPU_SCAN_EXCLUSIVE, 0u);
      return 2 * aux_round(n * kb) + aux_round(aux);
  }
  return 0;
}

hipError_t gen(hipStream_t s, void* out, int64_t n, int eb, int kind, uint64_t seed) {
  return generate(out, n, eb, kind, seed, 0, s) == MGPU_OK ? hipSuccess : hipErrorLaunchFailure;
}

// Aux bytes of a request on caller buffers (the only device memory the
// server allocates for it).
size_t caller_aux_bytes(const mgpu_request_t& r, const mgpu_request_buffers_t& b) {
  const int64_t n = r.n;
  const int64_t nb = b.n_b > 0 ? b.n_b : n;
  size_t aux = 0;
  switch (r.op) {
    case MGPU_REQ_SORT:
      if (r.key_type == MGPU_KEY_U64) {
        if (b.vals_out) radix_sort_tuned(nullptr, aux, nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr, n);
        else radix_sort_tuned(nullptr, aux, nullptr, (uint64_t*)nullptr, (empty_t*)nullptr, n);
      } else {
        if (b.vals_out) radix_sort_tuned(nullptr, aux, nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, n);
        else radix_sort_tuned(nullptr, aux, nullptr, (uint32_t*)nullptr, (empty_t*)nullptr, n);
      }
      break;
    case MGPU_REQ_MERGE:
      // Partitions only: the same for keys and key-value merges.
      if (r.key_type == MGPU_KEY_U64)
        merge<merge_tuning<uint64_t>::nt, merge_tuning<uint64_t>::vt>(
            nullptr, aux, nullptr, (const uint64_t*)nullptr, n, (const uint64_t*)nullptr, nb,
            (uint64_t*)nullptr, less_t<uint64_t>());
      else
        merge<merge_tuning<uint32_t>::nt, merge_tuning<uint32_t>::vt>(
            nullptr, aux, nullptr, (const uint32_t*)nullptr, n, (const uint32_t*)nullptr, nb,
            (uint32_t*)nullptr, less_t<uint32_t>());
      break;
    case MGPU_REQ_SCAN:
      if (r.key_type == MGPU_KEY_U64)
        scan64<uint64_t>(nullptr, aux, nullptr, nullptr, nullptr, n, MGPU_SCAN_EXCLUSIVE, 0ull);
      else
        scan<scan_tuning<uint32_t>::nt, scan_tuning<uint32_t>::vt>(nullptr, aux, nullptr, (const uint32_t*)nullptr,
                                               (uint32_t*)nullptr, n, MGPU_SCAN_EXCLUSIVE, 0u);
      break;
  }
  return aux_round(aux ? aux : 1);
}

template <typename K>
mgpu_status_t caller_sort(rec_t* r, hipStream_t s, size_t cap) {
  const mgpu_request_buffers_t& b = r->bufs;
  const int64_t n = r->req.n;
  hipError_t e = hipSuccess;
  if (b.out != b.in_a) e = hipMemcpyAsync(b.out, b.in_a, n * sizeof(K), hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return hip_status(e);
  if (!b.vals_out) return radix_sort_tuned(r->aux, cap, s, (K*)b.out, (empty_t*)nullptr, n);
  if (b.vals_a) {
    if (b.vals_a != b.vals_out)
      e = hipMemcpyAsync(b.vals_out, b.vals_a, n * 4, hipMemcpyDeviceToDevice, s);
  } else {
    e = gen(s, b.vals_out, n, 4, gen_iota, 0);
  }
  if (e != hipSuccess) return hip_status(e);
  return radix_sort_tuned(r->aux, cap, s, (K*)b.out, b.vals_out, n);
}

template <typename K>
mgpu_status_t caller_merge(rec_t* r, hipStream_t s, size_t cap) {
  const mgpu_request_buffers_t& b = r->bufs;
  const int64_t n = r->req.n, nb = b.n_b > 0 ? b.n_b : n;
  constexpr int NT = merge_tuning<K>::nt, VT = merge_tuning<K>::vt;
  if (!b.vals_out)
    return merge<NT, VT>(r->aux, cap, s, (const K*)b.in_a, n, (const K*)b.in_b, nb, (K*)b.out,
                         less_t<K>());
  return merge<NT, VT>(r->aux, cap, s, (const K*)b.in_a, b.vals_a, n, (const K*)b.in_b, b.vals_b,
                       nb, (K*)b.out, b.vals_out, less_t<K>());
}

// Enqueue a request on caller-owned buffers: sort copies the input into the
// output (unless they are the same) and sorts there, values from vals_a or
// the index; merge and scan write their output directly.
mgpu_status_t enqueue_caller(rec_t* r, hipStream_t s) {
  const mgpu_request_t& q = r->req;
  const mgpu_request_buffers_t& b = r->bufs;
  const int64_t n = q.n;
  const size_t cap = caller_aux_bytes(q, b);
  hipError_t e = hipMallocAsync(&r->aux, cap, s);
  if (e != hipSuccess) return hip_status(e);
  switch (q.op) {
    case MGPU_REQ_SORT:
      switch (q.key_type) {
        case MGPU_KEY_U64: return caller_sort<uint64_t>(r, s, cap);
        case MGPU_KEY_I32: return caller_sort<int32_t>(r, s, cap);
        case MGPU_KEY_F32: return caller_sort<float>(r, s, cap);
        default: return caller_sort<uint32_t>(r, s, cap);
      }
    case MGPU_REQ_MERGE:
      switch (q.key_type) {
        case MGPU_KEY_U64: return caller_merge<uint64_t>(r, s, cap);
        case MGPU_KEY_I32: return caller_merge<int32_t>(r, s, cap);
        case MGPU_KEY_F32: return caller_merge<float>(r, s, cap);
        default: return caller_merge<uint32_t>(r, s, cap);
      }
    case MGPU_REQ_SCAN: {
      size_t c = cap;
      switch (q.key_type) {
        case MGPU_KEY_U64:
          return scan64<uint64_t>(r->aux, c, s, (const uint64_t*)b.in_a, (uint64_t*)b.out, n,
                                  (mgpu_scan_type_t)b.scan_type, 0ull);
        case MGPU_KEY_F32:
          return scan<scan_tuning<float>::nt, scan_tuning<float>::vt>(r->aux, c, s, (const float*)b.in_a,
                                                        (float*)b.out, n,
                                                        (mgpu_scan_type_t)b.scan_type, 0.0f);
        case MGPU_KEY_I32:
          return scan<scan_tuning<uint32_t>::nt, scan_tuning<uint32_t>::vt>(r->aux, c, s, (const int32_t*)b.in_a,
                                                        (int32_t*)b.out, n,
                                                        (mgpu_scan_type_t)b.scan_type, 0);
        default:
          return scan<scan_tuning<uint32_t>::nt, scan_tuning<uint32_t>::vt>(r->aux, c, s, (const uint32_t*)b.in_a,
                                                        (uint32_t*)b.out, n,
                                                        (mgpu_scan_type_t)b.scan_type, 0u);
      }
    }
  }
  return MGPU_INVALID_ARG;
}


hipError_t alloc(rec_t* r, void** p, size_t bytes, hipStream_t s) {
  if (bytes == 0) bytes = 256;
  return hipMallocAsync(p, aux_round(bytes), s);
}

// Allocate, generate and enqueue one request on stream s.
mgpu_status_t enqueue(rec_t* r, hipStream_t s) {
  const mgpu_request_t& q = r->req;
  const int64_t n = q.n;
  const int kb = (int)key_bytes(q.key_type);
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t x) {
    if (x != hipSuccess && e == hipSuccess) e = x;
    return x == hipSuccess;
  };
  size_t aux = 0;
  switch (q.op) {
    case MGPU_REQ_SORT: {
      const int kind = q.key_type == MGPU_KEY_U64 ? gen_u64 : q.key_type == MGPU_KEY_F32 ? gen_f32_pm1 : gen_u32;
      if (q.key_type == MGPU_KEY_U64)
        radix_sort_tuned(nullptr, aux, s, (uint64_t*)nullptr, (uint32_t*)nullptr, n);
      else
        radix_sort_tuned(nullptr, aux, s, (uint32_t*)nullptr, (uint32_t*)nullptr, n);
      if (!ok(alloc(r, &r->in_a, n * kb, s)) || !ok(alloc(r, &r->out, n * kb, s)) ||
          !ok(alloc(r, &r->vals, n * 4, s)) || !ok(alloc(r, &r->aux, aux, s)))
        break;
      ok(gen(s, r->in_a, n, kb, kind, q.seed));
      ok(hipMemcpyAsync(r->out, r->in_a, n * kb, hipMemcpyDeviceToDevice, s));
      ok(gen(s, r->vals, n, 4, gen_iota, 0));
      if (e != hipSuccess) break;
      size_t cap = aux_round(aux ? aux : 1);
      mgpu_status_t st;
      switch (q.key_type) {
        case MGPU_KEY_U64:
          st = radix_sort_tuned(r->aux, cap, s, (uint64_t*)r->out, (uint32_t*)r->vals, n);
          break;
        case MGPU_KEY_I32:
          st = radix_sort_tuned(r->aux, cap, s, (int32_t*)r->out, (uint32_t*)r->vals, n);
          break;
        case MGPU_KEY_F32:
          st = radix_sort_tuned(r->aux, cap, s, (float*)r->out, (uint32_t*)r->vals, n);
          break;
        default:
          st = radix_sort_tuned(r->aux, cap, s, (uint32_t*)r->out, (uint32_t*)r->vals, n);
      }
      return st;
    }
    case MGPU_REQ_MERGE: {
      const int kind = q.key_type == MGPU_KEY_U64 ? gen_sorted_u64
                       : q.key_type == MGPU_KEY_F32 ? gen_sorted_f32 : gen_sorted_u32;
      if (q.key_type == MGPU_KEY_U64)
        serve_merge<uint64_t>(nullptr, aux, s, nullptr, nullptr, nullptr, n, nullptr, nullptr);
      else
        serve_merge<uint32_t>(nullptr, aux, s, nullptr, nullptr, nullptr, n, nullptr, nullptr);
      if (!ok(alloc(r, &r->in_a, n * kb, s)) || !ok(alloc(r, &r->in_b, n * kb, s)) ||
          !ok(alloc(r, &r->vals_a, 2 * n * 4, s)) || !ok(alloc(r, &r->out, 2 * n * kb, s)) ||
          !ok(alloc(r, &r->vals, 2 * n * 4, s)) || !ok(alloc(r, &r->aux, aux, s)))
        break;
      ok(gen(s, r->in_a, n, kb, kind, q.seed));
      ok(gen(s, r->in_b, n, kb, kind, q.seed + 1));
      ok(gen(s, r->vals_a, 2 * n, 4, gen_iota, 0));
      if (e != hipSuccess) break;
      size_t cap = aux_round(aux ? aux : 1);
      const uint32_t* va = (const uint32_t*)r->vals_a;
      uint32_t* vc = (uint32_t*)r->vals;
      switch (q.key_type) {
        case MGPU_KEY_U64:
          return serve_merge<uint64_t>(r->aux, cap, s, r->in_a, r->in_b, va, n, r->out, vc);
        case MGPU_KEY_F32:
          return serve_merge<float>(r->aux, cap, s, r->in_a, r->in_b, va, n, r->out, vc);
        case MGPU_KEY_I32:
          return serve_merge<int32_t>(r->aux, cap, s, r->in_a, r->in_b, va, n, r->out, vc);
        default:
          return serve_merge<uint32_t>(r->aux, cap, s, r->in_a, r->in_b, va, n, r->out, vc);
      }
    }
    case MGPU_REQ_SCAN: {
      const int kind = q.key_type == MGPU_KEY_U64 ? gen_u64 : q.key_type == MGPU_KEY_F32 ? gen_f32_01 : gen_u32;
      if (q.key_type == MGPU_KEY_U64)
        scan64<uint64_t>(nullptr, aux, s, nullptr, nullptr, n, MGPU_SCAN_EXCLUSIVE, 0ull);
      else
        scan<scan_tuning<uint32_t>::nt, scan_tuning<uint32_t>::vt>(nullptr, aux, s, (const uint32_t*)nullptr, (uint32_t*)nullptr, n,
                      MGPU_SCAN_EXCLUSIVE, 0u);
      if (!ok(alloc(r, &r->in_a, n * kb, s)) || !ok(alloc(r, &r->out, n * kb, s)) ||
          !ok(alloc(r, &r->aux, aux, s)))
        break;
      ok(gen(s, r->in_a, n, kb, kind, q.seed));
      if (e != hipSuccess) break;
      size_t cap = aux_round(aux ? aux : 1);
      switch (q.key_type) {
        case MGPU_KEY_U64:
          return scan64<uint64_t>(r->aux, cap, s, (const uint64_t*)r->in_a, (uint64_t*)r->out, n,
                                  MGPU_SCAN_EXCLUSIVE, 0ull);
        case MGPU_KEY_F32:
          return scan<scan_tuning<float>::nt, scan_tuning<float>::vt>(r->aux, cap, s, (const float*)r->in_a, (float*)r->out, n,
                               MGPU_SCAN_INCLUSIVE, 0.0f);
        default:
          return scan<scan_tuning<uint32_t>::nt, scan_tuning<uint32_t>::vt>(r->aux, cap, s, (const uint32_t*)r->in_a, (uint32_t*)r->out, n,
                               MGPU_SCAN_EXCLUSIVE, 0u);
      }
    }
  }
  return e == hipSuccess ? MGPU_INVALID_ARG : hip_status(e);
}

void free_buffers(mgpu_server_st* sv, rec_t* r) {
  for (void** p : {&r->in_a, &r->in_b, &r->out, &r->vals, &r->vals_a, &r->aux})
    if (*p && (!r->caller || p == &r->aux)) {
      (void)hipFreeAsync(*p, sv->free_stream);
      *p = nullptr;
    }
}

void worker_loop(mgpu_server_st* sv) {
  (void)hipSetDevice(sv->device);
  std::unique_lock<std::mutex> lock(sv->m);
  while (true) {
    bool progressed = false;
    // Destroy: requests not yet started are dropped (their records are freed
    // by destroy); only those in flight are waited for.  Without this a
    // pending request that no longer fits the budget would never leave.
    if (sv->stop) sv->pending.clear();
    // Completions.
    for (size_t i = 0; i < sv->slot_rec.size(); ++i) {
      rec_t* r = sv->slot_rec[i];
      if (!r) continue;
      hipError_t q = hipEventQuery(r->done);
      if (q == hipErrorNotReady) continue;
      r->t_done = sv->now();
      if (q != hipSuccess && r->status == MGPU_OK) r->status = hip_status(q);
      if (!sv->keep_inputs && !r->caller) {
        for (void** p : {&r->in_a, &r->in_b, &r->vals_a})
          if (*p) { (void)hipFreeAsync(*p, sv->free_stream); *p = nullptr; }
      }
      if (r->aux) { (void)hipFreeAsync(r->aux, sv->free_stream); r->aux = nullptr; }
      sv->completed.push_back(r);
      sv->slot_rec[i] = nullptr;
      --sv->in_flight;
      progressed = true;
      sv->cv.notify_all();
    }
    // Admissions: first fit from the head of the queue into free slots.
    for (size_t i = 0; i < sv->slot_rec.size(); ++i) {
      if (sv->slot_rec[i]) continue;
      auto it = sv->pending.begin();
      for (; it != sv->pending.end(); ++it)
        if (sv->in_use + (*it)->bytes <= sv->budget) break;
      if (it == sv->pending.end()) break;
      rec_t* r = *it;
      sv->pending.erase(it);
      sv->in_use += r->bytes;
      sv->peak = std::max(sv->peak, sv->in_use);
      r->slot = (int)i;
      r->t_start = sv->now();
      hipStream_t s = sv->streams[i];
      r->status = r->caller ? enqueue_caller(r, s) : enqueue(r, s);
      (void)hipEventRecord(r->done, s);
      sv->slot_rec[i] = r;
      ++sv->in_flight;
      progressed = true;
    }
    if (sv->stop && sv->pending.empty() && sv->in_flight == 0) break;
    if (!progressed) {
      if (sv->in_flight == 0)
        sv->cv.wait_for(lock, std::chrono::milliseconds(5));
      else {
        lock.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lock.lock();
      }
    }
  }
}

}  // namespace

extern "C" {

mgpu_status_t mgpu_server_create(int slots, size_t hbm_budget, int keep_inputs,
                                 mgpu_server_t* out) {
  if (!out || slots < 1 || slots > 64 || hbm_budget == 0) return MGPU_INVALID_ARG;
  *out = nullptr;
  mgpu_server_st* sv = new mgpu_server_st();
  sv->budget = hbm_budget;
  sv->keep_inputs = keep_inputs;
  sv->t0 = clock_t_::now();
  hipError_t e = hipGetDevice(&sv->device);
  (void)lookback_config();
  if (e == hipSuccess) e = hipDeviceGetDefaultMemPool(&sv->pool, sv->device);
  if (e == hipSuccess) {
    uint64_t threshold = ~0ull;  // keep freed memory in the pool
    e = hipMemPoolSetAttribute(sv->pool, hipMemPoolAttrReleaseThreshold, &threshold);
  }
  for (int i = 0; e == hipSuccess && i < slots; ++i) {
    hipStream_t s = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) sv->streams.push_back(s);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&sv->free_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    for (hipStream_t s : sv->streams) (void)hipStreamDestroy(s);
    delete sv;
    return hip_status(e);
  }
  sv->slot_rec.assign(slots, nullptr);
  sv->worker = std::thread(worker_loop, sv);
  *out = sv;
  return MGPU_OK;
}

double mgpu_server_now(mgpu_server_t sv) { return sv ? sv->now() : 0.0; }

size_t mgpu_server_request_bytes(const mgpu_request_t* req) {
  return req ? request_bytes(*req) : 0;
}

static mgpu_status_t submit(mgpu_server_t sv, const mgpu_request_t* req,
                            const mgpu_request_buffers_t* bufs, int64_t* id) {
  if (!sv || !req || req->n < 1) return MGPU_INVALID_ARG;
  if (req->op < MGPU_REQ_SORT || req->op > MGPU_REQ_SCAN) return MGPU_INVALID_ARG;
  if (req->key_type < MGPU_KEY_U32 || req->key_type > MGPU_KEY_F32) return MGPU_INVALID_ARG;
  if (req->n >= ((int64_t)1 << 30)) return MGPU_UNSUPPORTED;
  if (bufs) {
    if (!bufs->in_a || !bufs->out) return MGPU_INVALID_ARG;
    if (bufs->n_b < 0 || bufs->n_b >= ((int64_t)1 << 30)) return MGPU_INVALID_ARG;
    if (req->op == MGPU_REQ_MERGE) {
      if (!bufs->in_b) return MGPU_INVALID_ARG;
      if (bufs->vals_out && (!bufs->vals_a || !bufs->vals_b)) return MGPU_INVALID_ARG;
    }
    if (req->op == MGPU_REQ_SCAN && bufs->scan_type != MGPU_SCAN_EXCLUSIVE &&
        bufs->scan_type != MGPU_SCAN_INCLUSIVE)
      return MGPU_INVALID_ARG;
  }
  rec_t* r = new rec_t();
  r->req = *req;
  if (bufs) {
    r->caller = true;
    r->bufs = *bufs;
  }
  r->bytes = bufs ? caller_aux_bytes(*req, *bufs) : request_bytes(*req);
  if (r->bytes > sv->budget) {
    delete r;
    return MGPU_UNSUPPORTED;
  }
  if (hipEventCreateWithFlags(&r->done, hipEventDisableTiming) != hipSuccess) {
    delete r;
    return MGPU_HIP_ERROR;
  }
  std::lock_guard<std::mutex> lock(sv->m);
  r->id = sv->next_id++;
  r->t_submit = sv->now();
  sv->recs[r->id] = r;
  sv->pending.push_back(r);
  if (id) *id = r->id;
  sv->cv.notify_all();
  return MGPU_OK;
}

mgpu_status_t mgpu_server_submit(mgpu_server_t sv, const mgpu_request_t* req, int64_t* id) {
  return submit(sv, req, nullptr, id);
}

mgpu_status_t mgpu_server_submit_buffers(mgpu_server_t sv, const mgpu_request_t* req,
                                         const mgpu_request_buffers_t* bufs, int64_t* id) {
  if (!bufs) return MGPU_INVALID_ARG;
  return submit(sv, req, bufs, id);
}

int mgpu_server_poll(mgpu_server_t sv, mgpu_response_t* out, int max) {
  if (!sv) return 0;
  std::lock_guard<std::mutex> lock(sv->m);
  int n = 0;
  while (n < max && !sv->completed.empty()) {
    rec_t* r = sv->completed.front();
    sv->completed.erase(sv->completed.begin());
    mgpu_response_t& o = out[n++];
    o.id = r->id;
    o.status = r->status;
    o.op = r->req.op;
    o.key_type = r->req.key_type;
    o.n = r->req.n;
    if (r->caller) {
      o.keys = r->bufs.out;
      o.vals = r->bufs.vals_out;
      o.input_a = (void*)r->bufs.in_a;
      o.input_b = (void*)r->bufs.in_b;
    } else {
      o.keys = r->out;
      o.vals = r->vals;
      o.input_a = r->in_a;
      o.input_b = r->in_b;
    }
    o.t_submit_ms = r->t_submit;
    o.t_start_ms = r->t_start;
    o.t_done_ms = r->t_done;
    o.slot = r->slot;
    r->reported = true;
  }
  return n;
}

mgpu_status_t mgpu_server_drain(mgpu_server_t sv) {
  if (!sv) return MGPU_INVALID_ARG;
  std::unique_lock<std::mutex> lock(sv->m);
  // Nothing in flight and nothing pending that fits: only a release can
  // make progress, so waiting longer would never end.
  auto stuck = [&] {
    for (rec_t* r : sv->pending)
      if (sv->in_use + r->bytes <= sv->budget) return false;
    return true;
  };
  sv->cv.wait(lock, [&] { return sv->in_flight == 0 && (sv->pending.empty() || stuck()); });
  return sv->pending.empty() ? MGPU_OK : MGPU_AUX_TOO_SMALL;
}

mgpu_status_t mgpu_server_release(mgpu_server_t sv, int64_t id) {
  if (!sv) return MGPU_INVALID_ARG;
  std::lock_guard<std::mutex> lock(sv->m);
  auto it = sv->recs.find(id);
  if (it == sv->recs.end() || !it->second->reported) return MGPU_INVALID_ARG;
  rec_t* r = it->second;
  free_buffers(sv, r);
  (void)hipEventDestroy(r->done);
  sv->in_use -= r->bytes;
  sv->recs.erase(it);
  delete r;
  sv->cv.notify_all();
  return MGPU_OK;
}

void mgpu_server_stats(mgpu_server_t sv, size_t* in_use, size_t* peak, int64_t* pending,
                       int64_t* in_flight) {
  if (!sv) return;
  std::lock_guard<std::mutex> lock(sv->m);
  if (in_use) *in_use = sv->in_use;
  if (peak) *peak = sv->peak;
  if (pending) *pending = (int64_t)sv->pending.size();
  if (in_flight) *in_flight = sv->in_flight;
}

void mgpu_server_destroy(mgpu_server_t sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lock(sv->m);
    sv->stop = true;
    sv->cv.notify_all();
  }
  sv->worker.join();
  for (auto& kv : sv->recs) {
    free_buffers(sv, kv.second);
    (void)hipEventDestroy(kv.second->done);
    delete kv.second;
  }
  (void)hipStreamSynchronize(sv->free_stream);
  (void)hipMemPoolTrimTo(sv->pool, 0);  // hand the pooled memory back
  for (hipStream_t s : sv->streams) (void)hipStreamDestroy(s);
  (void)hipStreamDestroy(sv->free_stream);
  delete sv;
}

}  // extern "C"
