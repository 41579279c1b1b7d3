// A call site written against the reference's host surface, ported to
// mgpu/hip: context_t / cmd_buffer_t / memcache_t (inc/mgpu/vk/context.hxx:
// 11-134) allocate and order the work, in the pattern of the reference's
// examples/vk_sort/vk_sort.cxx:8-49 (alloc_cpu / alloc_gpu, cmd_buffer
// memcpy, two-phase mergesort with aux from context.alloc_gpu, submit, wait).
// Also: a radix sort and a merge whose aux comes from one memcache_t, and
// merge_path_partitions<bounds> into the reference's int partition array
// (inc/mgpu/vk/partition.hxx:9-20).  Every result is checked against the
// std:: algorithm on the host.  Prints "[case] <name>" before each GPU step.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "mgpu/hip/context.hxx"
#include "mgpu/hip/merge.hxx"
#include "mgpu/hip/mergesort.hxx"
#include "mgpu/hip/radix.hxx"

using namespace mgpu::hip;

#define CHECK(cond, what)                                     \
  do {                                                        \
    if (!(cond)) {                                            \
      std::printf("FAIL %s (line %d)\n", what, __LINE__);     \
      return 1;                                               \
    }                                                         \
  } while (0)

static void banner(const char* name) {
  std::printf("[case] %s\n", name);
  std::fflush(stdout);
}

int main() {
  context_t context;
  CHECK(context.last_error() == hipSuccess, "context");

  // examples/vk_sort: float keys, rand() % 100000, key-only mergesort.
  banner("vk_sort pattern: mergesort_keys through context_t / cmd_buffer_t");
  const int count = 100000;
  float* host = context.alloc_cpu<float>(count);
  float* gpu = context.alloc_gpu<float>(count);
  CHECK(host && gpu, "alloc");
  srand(1);
  for (int i = 0; i < count; ++i) host[i] = rand() % 100000;
  std::vector<float> expect(host, host + count);
  std::stable_sort(expect.begin(), expect.end());

  cmd_buffer_t cmd_buffer(context);
  cmd_buffer.begin();
  cmd_buffer.memcpy(gpu, host, sizeof(float) * count);
  cmd_buffer.host_barrier();
  void* aux_data = nullptr;
  size_t aux_size = 0;
  CHECK(mergesort_keys(aux_data, aux_size, cmd_buffer, gpu, count) == MGPU_OK, "sizing");
  aux_data = context.alloc_gpu(aux_size);
  CHECK(mergesort_keys(aux_data, aux_size, cmd_buffer, gpu, count) == MGPU_OK, "mergesort");
  cmd_buffer.memcpy(host, gpu, sizeof(float) * count);
  cmd_buffer.host_barrier();
  cmd_buffer.end();
  context.submit(cmd_buffer);
  CHECK(context.wait_idle() == hipSuccess, "wait");
  CHECK(std::equal(expect.begin(), expect.end(), host), "mergesort result");
  context.free(aux_data);
  std::printf("PASS vk_sort pattern\n");

  // memcache_t: one grow-only buffer serves a radix sort, then a merge.
  banner("memcache_t aux for radix_sort and merge");
  memcache_t cache(context);
  const int n = 300000;
  uint32_t* hk = context.alloc_cpu<uint32_t>(n);
  uint32_t* hv = context.alloc_cpu<uint32_t>(n);
  uint32_t* dk = context.alloc_gpu<uint32_t>(n);
  uint32_t* dv = context.alloc_gpu<uint32_t>(n);
  uint32_t* dc = context.alloc_gpu<uint32_t>(n);
  uint32_t* hc = context.alloc_cpu<uint32_t>(n);
  CHECK(hk && hv && dk && dv && dc && hc, "alloc 2");
  for (int i = 0; i < n; ++i) {
    hk[i] = (uint32_t)rand() % 5000;
    hv[i] = (uint32_t)i;
  }
  std::vector<uint32_t> order(n);
  for (int i = 0; i < n; ++i) order[i] = (uint32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return hk[a] < hk[b]; });
  std::vector<uint32_t> keys0(hk, hk + n);
  context.memcpy(dk, hk, 4 * n);
  context.memcpy(dv, hv, 4 * n);
  size_t need = 0;
  CHECK(radix_sort(nullptr, need, context, dk, dv, n) == MGPU_OK, "radix sizing");
  void* aux = cache.allocate(need);
  CHECK(radix_sort(aux, need, context, dk, dv, n) == MGPU_OK, "radix");
  context.memcpy(hk, dk, 4 * n);
  context.memcpy(hv, dv, 4 * n);
  CHECK(context.wait_idle() == hipSuccess, "wait 2");
  for (int i = 0; i < n; ++i)
    CHECK(hv[i] == order[i] && hk[i] == keys0[order[i]], "radix result");
  // Merge the two sorted halves of the sorted keys back (a trivial but
  // tie-heavy merge): C == the sorted keys.
  const int na = n / 3, nb = n - na;
  need = 0;
  CHECK(merge(nullptr, need, context, (const uint32_t*)dk, na, (const uint32_t*)dk + na, nb, dc,
              less_t<uint32_t>()) == MGPU_OK, "merge sizing");
  void* aux2 = cache.allocate(need);
  CHECK(merge(aux2, need, context, (const uint32_t*)dk, na, (const uint32_t*)dk + na, nb, dc,
              less_t<uint32_t>()) == MGPU_OK, "merge");
  context.memcpy(hc, dc, 4 * n);
  CHECK(context.wait_idle() == hipSuccess, "wait 3");
  std::vector<uint32_t> em(n);
  std::merge(hk, hk + na, hk + na, hk + n, em.begin());
  CHECK(std::equal(em.begin(), em.end(), hc), "merge result");
  std::printf("PASS memcache_t\n");

  // merge_path_partitions into int partitions, both bounds.
  banner("merge_path_partitions<bounds> into int mp_data");
  const int spacing = 1000;
  const int parts = (int)num_merge_partitions(n, spacing);
  int* dmp = context.alloc_gpu<int>(parts);
  int* hmp = context.alloc_cpu<int>(parts);
  for (int upper = 0; upper < 2; ++upper) {
    mgpu_status_t st =
        upper ? merge_path_partitions<bounds_upper>(context, (const uint32_t*)dk, na,
                                                    (const uint32_t*)dk + na, nb, dmp, spacing,
                                                    less_t<uint32_t>())
              : merge_path_partitions<bounds_lower>(context, (const uint32_t*)dk, na,
                                                    (const uint32_t*)dk + na, nb, dmp, spacing,
                                                    less_t<uint32_t>());
    CHECK(st == MGPU_OK, "partitions");
    context.memcpy(hmp, dmp, sizeof(int) * parts);
    CHECK(context.wait_idle() == hipSuccess, "wait 4");
    for (int i = 0; i < parts; ++i) {
      const int diag = std::min(i * spacing, n);
      // Brute force: A elements among the first diag outputs of the merge
      // with ties to A (lower) or to B (upper).
      int ia = 0, ib = 0;
      for (int d = 0; d < diag; ++d) {
        bool take_a;
        if (ia >= na) take_a = false;
        else if (ib >= nb) take_a = true;
        else take_a = upper ? hk[ia] < hk[na + ib] : !(hk[na + ib] < hk[ia]);
        if (take_a) ++ia;
        else ++ib;
      }
      CHECK(hmp[i] == ia, upper ? "partition (upper)" : "partition (lower)");
    }
  }
  std::printf("PASS merge_path_partitions\n");
  std::printf("ALL PASS\n");
  return 0;
}
