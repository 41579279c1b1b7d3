// C ABI: status strings, HIP error slot, event timing, synthetic generator.
#include <map>

#include "mgpu/hip/generate.hxx"
#include "mgpu/hip/meta.hxx"
#include "mgpu/hip/segsort.hxx"
#include "mgpu/hip/transform.hxx"

using namespace mgpu::hip;

// Device address of the no-allocation scan's scratch on the current device
// (capi_scan_noalloc.hip), looked up and kept per device.
hipError_t mgpu_noalloc_scratch(uint64_t** out);

extern "C" {

const char* mgpu_status_string(mgpu_status_t status) {
  switch (status) {
    case MGPU_OK: return "MGPU_OK";
    case MGPU_INVALID_ARG: return "MGPU_INVALID_ARG";
    case MGPU_AUX_TOO_SMALL: return "MGPU_AUX_TOO_SMALL";
    case MGPU_HIP_ERROR: return "MGPU_HIP_ERROR";
    case MGPU_UNSUPPORTED: return "MGPU_UNSUPPORTED";
    default: break;
  }
  return "MGPU_UNKNOWN_STATUS";
}

int mgpu_last_hip_error(void) { return last_hip_error_slot(); }

int mgpu_version(void) { return 100; }

mgpu_status_t mgpu_init(void) {
  if (mgpu_status_t st = init()) return st;
  uint64_t* scratch = nullptr;
  return hip_status(mgpu_noalloc_scratch(&scratch));
}

void mgpu_timing_enable(int enable) {
  timing_registry_t& tr = timing_registry();
  std::lock_guard<std::mutex> lock(tr.mutex);
  for (auto& r : tr.recs) {
    tr.pool.push_back(r.start);
    tr.pool.push_back(r.stop);
  }
  tr.recs.clear();
  tr.enabled = enable != 0;
}

int mgpu_timing_collect(const char** names, int* launches, double* total_ms, int capacity) {
  timing_registry_t& tr = timing_registry();
  std::lock_guard<std::mutex> lock(tr.mutex);
  std::map<std::string, std::pair<int, double>> agg;
  std::vector<std::string> order;
  for (auto& r : tr.recs) {
    (void)hipEventSynchronize(r.stop);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.start, r.stop);
    std::string key(r.name);
    auto it = agg.find(key);
    if (it == agg.end()) {
      agg[key] = {1, (double)ms};
      order.push_back(key);
    } else {
      it->second.first += 1;
      it->second.second += ms;
    }
  }
  tr.names = order;
  int n = (int)order.size();
  for (int i = 0; i < n && i < capacity; ++i) {
    if (names) names[i] = tr.names[i].c_str();
    if (launches) launches[i] = agg[order[i]].first;
    if (total_ms) total_ms[i] = agg[order[i]].second;
  }
  return n;
}

mgpu_status_t mgpu_generate(mgpu_stream_t stream, void* out, int64_t count, int elem_bytes,
                            mgpu_gen_kind_t kind, uint64_t seed, uint64_t offset) {
  if (count < 0 || (elem_bytes != 4 && elem_bytes != 8)) return MGPU_INVALID_ARG;
  if ((int)kind < 0 || (int)kind > (int)gen_sorted_f32) return MGPU_INVALID_ARG;
  if ((kind == MGPU_GEN_U64 || kind == MGPU_GEN_SORTED_U64) && elem_bytes != 8) return MGPU_INVALID_ARG;
  if (count == 0) return MGPU_OK;
  if (!out) return MGPU_INVALID_ARG;
  return generate(out, count, elem_bytes, (int)kind, seed, offset, (hipStream_t)stream);
}

mgpu_status_t mgpu_copy(mgpu_stream_t stream, void* dst, const void* src, int64_t bytes) {
  return copy_bytes(dst, src, bytes, (hipStream_t)stream);
}

void mgpu_set_lookback_spin_limit(uint32_t limit) {
  lookback_settings_t& s = lookback_settings();
  std::lock_guard<std::mutex> lock(s.mutex);
  s.spin_limit = limit;
}

uint32_t mgpu_lookback_spin_limit(void) {
  lookback_settings_t& s = lookback_settings();
  std::lock_guard<std::mutex> lock(s.mutex);
  return s.spin_limit;
}

int mgpu_segsort_shapes(int* nt, int* vt, int capacity) {
  for (int i = 0; i < segsort_num_shapes && i < capacity; ++i) {
    if (nt) nt[i] = segsort_shapes[i].nt;
    if (vt) vt[i] = segsort_shapes[i].vt;
  }
  return segsort_num_shapes;
}

struct mgpu_graph_st {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

mgpu_status_t mgpu_capture_begin(mgpu_stream_t stream) {
  if (!stream) return MGPU_INVALID_ARG;  // the legacy default stream cannot capture
  (void)init();  // per-device state (look-back counter, side streams) exists before capture
  timing_registry().enabled = false;
  return hip_status(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
}

mgpu_status_t mgpu_capture_end(mgpu_stream_t stream, mgpu_graph_t* graph) {
  if (!stream || !graph) return MGPU_INVALID_ARG;
  *graph = nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
  if (e != hipSuccess) return hip_status(e);
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_status(e);
  }
  *graph = new mgpu_graph_st{g, x};
  return MGPU_OK;
}

mgpu_status_t mgpu_graph_launch(mgpu_graph_t graph, mgpu_stream_t stream) {
  if (!graph) return MGPU_INVALID_ARG;
  return hip_status(hipGraphLaunch(graph->exec, (hipStream_t)stream));
}

void mgpu_graph_destroy(mgpu_graph_t graph) {
  if (!graph) return;
  (void)hipGraphExecDestroy(graph->exec);
  (void)hipGraphDestroy(graph->graph);
  delete graph;
}

int64_t mgpu_lookback_fallbacks(int reset) {
  lookback_config_t c = lookback_config();
  if (!c.fallbacks) return -1;
  uint32_t v = 0;
  if (hipMemcpy(&v, c.fallbacks, 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
  if (reset && hipMemset(c.fallbacks, 0, 4) != hipSuccess) return -1;
  return (int64_t)v;
}

}  // extern "C"
