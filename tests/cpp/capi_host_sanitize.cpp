// Host-side argument validation and aux sizing of every C-ABI entry point,
// run without a GPU: the library's translation units are compiled for the
// host only with AddressSanitizer and UndefinedBehaviorSanitizer and linked
// with this driver, which makes only calls that return before any device
// work (sizing with aux = NULL, invalid arguments, limits, settings, status
// strings).  Exits 0 and prints "ALL PASS" when every expectation holds and
// neither sanitizer reports.
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "mgpu_hip.h"

static int failures = 0;
#define EXPECT(cond, ...)                      \
  do {                                         \
    if (!(cond)) {                             \
      std::printf("FAIL line %d: ", __LINE__); \
      std::printf(__VA_ARGS__);                \
      std::printf("\n");                       \
      ++failures;                              \
    }                                          \
  } while (0)

static void* const P = (void*)(uintptr_t)256;  // a non-null pointer never dereferenced

static const int64_t kCounts[] = {0, 1, 2, 255, 256, 4095, 4096, 4097, 9727, 9728, 9729, 65536,
                                  1000003, (int64_t)1 << 24, ((int64_t)1 << 28) + 5,
                                  ((int64_t)1 << 31) - 1};

// Sizing contract: aux = NULL adds the bytes to *aux_size and returns OK for
// valid arguments; a negative count is rejected; a NULL aux_size is
// rejected; the size of count 0 is 0 except for the reduce, which still
// writes init to *out through its partials.
#define CHECK_SIZING(name, call_with) CHECK_SIZING_EMPTY(name, call_with, true)
#define CHECK_SIZING_EMPTY(name, call_with, empty_is_free)                              \
  do {                                                                                  \
    for (int64_t n : kCounts) {                                                         \
      size_t sz = 0;                                                                    \
      mgpu_status_t st = call_with(&sz, n);                                             \
      EXPECT(st == MGPU_OK || st == MGPU_UNSUPPORTED, "%s n=%lld status %d", name,      \
             (long long)n, (int)st);                                                    \
      if (n == 0 && (empty_is_free)) EXPECT(sz == 0, "%s size of an empty call %zu", name, sz); \
      size_t acc = 1000;                                                                \
      mgpu_status_t st2 = call_with(&acc, n);                                           \
      EXPECT(st2 == st, "%s status differs on a second sizing", name);                  \
      if (st == MGPU_OK) EXPECT(acc == 1000 + sz, "%s sizing does not accumulate", name); \
    }                                                                                   \
    size_t sz = 0;                                                                      \
    EXPECT(call_with(&sz, -1) == MGPU_INVALID_ARG, "%s accepts a negative count", name); \
    EXPECT(call_with(nullptr, 100) == MGPU_INVALID_ARG, "%s accepts a NULL aux_size", name); \
  } while (0)

#define SORT_SIZING(K, KT)                                                               \
  {                                                                                      \
    auto ms_keys = [](size_t* s, int64_t n) {                                            \
      return mgpu_mergesort_keys_##K(nullptr, s, nullptr, (KT*)P, n, MGPU_LESS);         \
    };                                                                                   \
    auto ms_kv = [](size_t* s, int64_t n) {                                              \
      return mgpu_mergesort_kv_##K##_u32(nullptr, s, nullptr, (KT*)P, (uint32_t*)P, n,  \
                                         MGPU_GREATER);                                  \
    };                                                                                   \
    auto ms_idx = [](size_t* s, int64_t n) {                                             \
      return mgpu_mergesort_indices_##K(nullptr, s, nullptr, (KT*)P, (int32_t*)P, n,     \
                                        MGPU_LESS);                                      \
    };                                                                                   \
    auto rx_keys = [](size_t* s, int64_t n) {                                            \
      return mgpu_radix_sort_keys_##K(nullptr, s, nullptr, (KT*)P, n);                   \
    };                                                                                   \
    auto rx_kv = [](size_t* s, int64_t n) {                                              \
      return mgpu_radix_sort_kv_##K##_u32(nullptr, s, nullptr, (KT*)P, (uint32_t*)P, n); \
    };                                                                                   \
    auto mg_keys = [](size_t* s, int64_t n) {                                            \
      return mgpu_merge_keys_##K(nullptr, s, nullptr, (const KT*)P, n / 3, (const KT*)P, \
                                 n < 0 ? n : n - n / 3, (KT*)P, MGPU_LESS);              \
    };                                                                                   \
    auto ss_kv = [](size_t* s, int64_t n) {                                              \
      return mgpu_segmented_sort_kv_##K##_u32(nullptr, s, nullptr, (KT*)P, (uint32_t*)P, \
                                              n, (const int32_t*)P, n / 100 + 1,         \
                                              MGPU_LESS, nullptr);                       \
    };                                                                                   \
    CHECK_SIZING("mergesort_keys_" #K, ms_keys);                                         \
    CHECK_SIZING("mergesort_kv_" #K, ms_kv);                                             \
    CHECK_SIZING("mergesort_indices_" #K, ms_idx);                                       \
    CHECK_SIZING("radix_sort_keys_" #K, rx_keys);                                        \
    CHECK_SIZING("radix_sort_kv_" #K, rx_kv);                                            \
    CHECK_SIZING("merge_keys_" #K, mg_keys);                                             \
    CHECK_SIZING("segmented_sort_kv_" #K, ss_kv);                                        \
    size_t sz = 0;                                                                       \
    EXPECT(mgpu_mergesort_kv_##K##_u32(nullptr, &sz, nullptr, (KT*)P, (uint32_t*)P,     \
                                       (int64_t)1 << 31, MGPU_LESS) == MGPU_UNSUPPORTED, \
           "mergesort %s accepts 2^31", #K);                                             \
    EXPECT(mgpu_mergesort_keys_##K(nullptr, &sz, nullptr, (KT*)P, 100, (mgpu_comp_t)7) == \
               MGPU_INVALID_ARG, "mergesort %s accepts comparator 7", #K);              \
    EXPECT(mgpu_merge_keys_##K(nullptr, &sz, nullptr, (const KT*)P, 5, (const KT*)P, 5,  \
                               (KT*)P, (mgpu_comp_t)-1) == MGPU_INVALID_ARG,             \
           "merge %s accepts comparator -1", #K);                                        \
    EXPECT(mgpu_radix_sort_kv_##K##_u32(nullptr, &sz, nullptr, (KT*)P, (uint32_t*)P,     \
                                        (int64_t)1 << 40) == MGPU_UNSUPPORTED,           \
           "radix %s accepts 2^40", #K);                                                 \
    EXPECT(mgpu_merge_path_partitions_##K(nullptr, (const KT*)P, 10, (const KT*)P, 10,   \
                                          (int64_t*)P, 0, MGPU_LESS,                     \
                                          MGPU_BOUNDS_LOWER) == MGPU_INVALID_ARG,        \
           "partitions %s accept spacing 0", #K);                                        \
    EXPECT(mgpu_merge_path_partitions_##K(nullptr, (const KT*)P, -1, (const KT*)P, 10,   \
                                          (int64_t*)P, 5, MGPU_LESS,                     \
                                          MGPU_BOUNDS_UPPER) == MGPU_INVALID_ARG,        \
           "partitions %s accept a negative count", #K);                                 \
  }

#define SCAN_SIZING(T, TT)                                                               \
  {                                                                                      \
    auto sc_ex = [](size_t* s, int64_t n) {                                              \
      return mgpu_scan_exclusive_##T(nullptr, s, nullptr, (TT*)P, n, (TT)0);             \
    };                                                                                   \
    auto sc = [](size_t* s, int64_t n) {                                                 \
      return mgpu_scan_##T(nullptr, s, nullptr, (const TT*)P, (TT*)P, n,                 \
                           MGPU_SCAN_INCLUSIVE, (TT)1);                                  \
    };                                                                                   \
    auto rd = [](size_t* s, int64_t n) {                                                 \
      return mgpu_reduce_##T(nullptr, s, nullptr, (const TT*)P, n, (TT*)P, (TT)0);       \
    };                                                                                   \
    CHECK_SIZING("scan_exclusive_" #T, sc_ex);                                           \
    CHECK_SIZING("scan_" #T, sc);                                                        \
    CHECK_SIZING_EMPTY("reduce_" #T, rd, false);                                         \
  }

#define SEGSCAN_SIZING(T, TT)                                                            \
  {                                                                                      \
    auto ss = [](size_t* s, int64_t n) {                                                 \
      return mgpu_segmented_scan_##T(nullptr, s, nullptr, (const TT*)P, (TT*)P, n,       \
                                     (const int32_t*)P, n / 7 + 1, MGPU_SCAN_EXCLUSIVE,  \
                                     (TT)0);                                             \
    };                                                                                   \
    CHECK_SIZING("segmented_scan_" #T, ss);                                              \
  }

int main() {
  SORT_SIZING(u32, uint32_t)
  SORT_SIZING(i32, int32_t)
  SORT_SIZING(u64, uint64_t)
  SORT_SIZING(f32, float)
  SORT_SIZING(i64, int64_t)
  SORT_SIZING(f64, double)
  SCAN_SIZING(i32, int32_t)
  SCAN_SIZING(u32, uint32_t)
  SCAN_SIZING(f32, float)
  SCAN_SIZING(u64, uint64_t)
  SEGSCAN_SIZING(i32, int32_t)
  SEGSCAN_SIZING(u32, uint32_t)
  SEGSCAN_SIZING(f32, float)

  // Sizing does not depend on the process-wide structure settings.
  for (int ways : {0, 2, 4}) {
    mgpu_set_mergesort_passes(ways);
    EXPECT(mgpu_mergesort_passes() == ways, "mergesort passes setting %d", ways);
    size_t a = 0;
    mgpu_mergesort_kv_u32_u32(nullptr, &a, nullptr, (uint32_t*)P, (uint32_t*)P, 1 << 26, MGPU_LESS);
    static size_t first = 0;
    if (!first) first = a;
    EXPECT(a == first, "mergesort sizing depends on the pass setting");
  }
  mgpu_set_mergesort_passes(9);
  EXPECT(mgpu_mergesort_passes() == 0, "bad pass setting kept");
  mgpu_set_mergesort_passes(0);
  for (int algo : {0, 1, 2}) {
    mgpu_set_radix_algorithm(algo);
    EXPECT(mgpu_radix_algorithm() == algo, "radix algorithm setting %d", algo);
  }
  mgpu_set_radix_algorithm(0);

  // The no-alloc scan rejects before any device work.
  EXPECT(mgpu_scan_i32_noalloc(nullptr, (const int32_t*)P, (int32_t*)P, -5,
                               MGPU_SCAN_EXCLUSIVE, 0) == MGPU_INVALID_ARG, "noalloc negative");
  EXPECT(mgpu_scan_i32_noalloc(nullptr, (const int32_t*)P, (int32_t*)P, 10, (mgpu_scan_type_t)3,
                               0) == MGPU_INVALID_ARG, "noalloc bad type");
  EXPECT(mgpu_scan_i32_noalloc(nullptr, (const int32_t*)P, (int32_t*)P, 0, MGPU_SCAN_EXCLUSIVE,
                               0) == MGPU_OK, "noalloc empty");
  EXPECT(mgpu_scan_i32_noalloc(nullptr, nullptr, (int32_t*)P, 10, MGPU_SCAN_EXCLUSIVE, 0) ==
             MGPU_INVALID_ARG, "noalloc null input");
  EXPECT(mgpu_scan_i32_noalloc(nullptr, (const int32_t*)P, (int32_t*)P, (int64_t)1 << 33,
                               MGPU_SCAN_INCLUSIVE, 0) == MGPU_UNSUPPORTED, "noalloc too large");

  // Segmented sort tuning tables are validated.
  {
    mgpu_segsort_tuning_t t;
    std::memset(&t, 0, sizeof(t));
    t.num_buckets = 7;
    size_t sz = 0;
    EXPECT(mgpu_segmented_sort_keys_u32(nullptr, &sz, nullptr, (uint32_t*)P, 1000,
                                        (const int32_t*)P, 10, MGPU_LESS, &t) == MGPU_INVALID_ARG,
           "segsort accepts 7 buckets");
    t.num_buckets = 2;
    t.cap[0] = 512;
    t.cap[1] = 256;  // not ascending
    EXPECT(mgpu_segmented_sort_keys_u32(nullptr, &sz, nullptr, (uint32_t*)P, 1000,
                                        (const int32_t*)P, 10, MGPU_LESS, &t) == MGPU_INVALID_ARG,
           "segsort accepts descending capacities");
    int nt[16], vt[16];
    const int shapes = mgpu_segsort_shapes(nt, vt, 16);
    EXPECT(shapes > 0 && shapes <= 16, "segsort shapes %d", shapes);
    for (int i = 0; i < shapes; ++i) EXPECT(nt[i] > 0 && vt[i] > 0, "shape %d", i);
    EXPECT(mgpu_segsort_shapes(nt, vt, 1) == shapes, "shapes with a short array");
  }

  // Status strings cover every code and do not read out of range.
  for (int s = -3; s < 16; ++s) {
    const char* str = mgpu_status_string((mgpu_status_t)s);
    EXPECT(str != nullptr && std::strlen(str) > 0, "status string %d", s);
  }
  EXPECT(mgpu_version() >= 100, "version");

  std::printf(failures ? "%d FAILURES\n" : "ALL PASS\n", failures);
  return failures ? 1 : 0;
}
