// Oracle self-check under AddressSanitizer and UndefinedBehaviorSanitizer
// (host code only; test infrastructure, built by `make -C oracle sanitize`).
// Runs every exported oracle function on small and edge-case inputs of every
// key type and checks that the restatements of the reference's schedules
// agree with the std:: algorithms they are pinned to, so that an
// out-of-bounds access or undefined operation anywhere in the oracle stops
// the run with a report.  Exit status 0 when every check passes.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {
void oracle_generate(void* out, int64_t count, int elem_bytes, int kind, uint64_t seed,
                     uint64_t offset);
void oracle_glibc_rand(int64_t n, uint32_t* out);
void oracle_stable_sort(int kt, void* keys, uint32_t* vals, int64_t n, int greater);
double oracle_stable_sort_timed(int kt, void* keys, uint32_t* vals, int64_t n, int greater);
double oracle_stable_sort_threads_timed(int kt, void* keys, uint32_t* vals, int64_t n, int greater,
                                        int threads);
void oracle_segmented_stable_sort(int kt, void* keys, uint32_t* vals, int64_t n,
                                  const int32_t* segs, int64_t num_segments, int greater);
void oracle_stable_sort_radix(int kt, void* keys, uint32_t* vals, int64_t n);
void oracle_std_sort_u32(uint32_t* keys, int64_t n);
void oracle_merge(int kt, const void* a, const uint32_t* av, int64_t na, const void* b,
                  const uint32_t* bv, int64_t nb, void* c, uint32_t* cv, int greater);
void oracle_exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t init);
void oracle_inclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t init);
void oracle_inclusive_scan_f64(const float* in, double* out, int64_t n, double init);
void oracle_exclusive_scan_f64(const float* in, double* out, int64_t n, double init);
void oracle_segmented_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, const int32_t* segs,
                               int64_t num_segments, int inclusive, uint32_t init);
void oracle_segmented_scan_f64(const float* in, double* out, int64_t n, const int32_t* segs,
                               int64_t num_segments, int inclusive, double init);
void oracle_inclusive_scan_f32(const float* in, float* out, int64_t n);
uint32_t oracle_reduce_u32(const uint32_t* in, int64_t n, uint32_t init);
double oracle_reduce_f64(const float* in, int64_t n, double init);
int64_t oracle_merge_path(int kt, const void* a, int64_t na, const void* b, int64_t nb,
                          int64_t diag, int greater, int bounds);
void oracle_mergesort_schedule(int kt, void* keys, uint32_t* vals, int64_t n, int nt, int vt,
                               int greater, int sort_indices);
void oracle_radix_schedule(int kt, void* keys, uint32_t* vals, int64_t n, int64_t nv,
                           uint32_t* counts_out, uint64_t* offsets_out);
void oracle_radix_bits(int kt, const void* keys, void* bits, int64_t n, int inverse);
void oracle_scan_schedule_u32(uint32_t* data, int64_t n, int64_t nv, uint32_t init);
}

static int failures = 0;
#define EXPECT(cond, ...)                          \
  do {                                             \
    if (!(cond)) {                                 \
      std::printf("FAIL line %d: ", __LINE__);     \
      std::printf(__VA_ARGS__);                    \
      std::printf("\n");                           \
      ++failures;                                  \
    }                                              \
  } while (0)

static int key_bytes(int kt) { return kt == 2 || kt == 4 || kt == 5 ? 8 : 4; }

// Keys of type kt with many ties: generated words folded onto 37 values.
static std::vector<uint8_t> make_keys(int kt, int64_t n, uint64_t seed) {
  const int eb = key_bytes(kt);
  std::vector<uint8_t> k((size_t)(n * eb + 8));
  std::vector<uint32_t> w((size_t)n + 1);
  oracle_generate(w.data(), n, 4, 0, seed, 0);
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t r = w[i] % 37u;
    switch (kt) {
      case 0: { uint32_t x = r * 0x10000001u; std::memcpy(&k[i * 4], &x, 4); break; }
      case 1: { int32_t x = (int32_t)r - 18; std::memcpy(&k[i * 4], &x, 4); break; }
      case 2: { uint64_t x = (uint64_t)r << 40; std::memcpy(&k[i * 8], &x, 8); break; }
      case 3: { float x = (float)r - 18.5f; std::memcpy(&k[i * 4], &x, 4); break; }
      case 4: { int64_t x = (int64_t)r - 18; std::memcpy(&k[i * 8], &x, 8); break; }
      default: { double x = (double)r * 0.25 - 4.0; std::memcpy(&k[i * 8], &x, 8); break; }
    }
  }
  return k;
}

int main() {
  const int64_t sizes[] = {0, 1, 2, 7, 100, 1000, 4097};
  for (int kt = 0; kt < 6; ++kt) {
    const int eb = key_bytes(kt);
    for (int64_t n : sizes) {
      for (int greater = 0; greater < 2; ++greater) {
        std::vector<uint8_t> keys = make_keys(kt, n, 11 + n);
        std::vector<uint32_t> vals((size_t)n + 1);
        for (int64_t i = 0; i < n; ++i) vals[i] = (uint32_t)i;
        // std::stable_sort vs the restated mergesort schedule (small tiles).
        std::vector<uint8_t> k1 = keys, k2 = keys;
        std::vector<uint32_t> v1 = vals, v2 = vals;
        oracle_stable_sort(kt, k1.data(), v1.data(), n, greater);
        oracle_mergesort_schedule(kt, k2.data(), v2.data(), n, 8, 3, greater, 0);
        EXPECT(std::memcmp(k1.data(), k2.data(), (size_t)(n * eb)) == 0 && v1 == v2,
               "mergesort schedule kt=%d n=%lld greater=%d", kt, (long long)n, greater);
        std::vector<uint8_t> k3 = keys;
        std::vector<uint32_t> v3 = vals;
        (void)oracle_stable_sort_timed(kt, k3.data(), v3.data(), n, greater);
        EXPECT(std::memcmp(k1.data(), k3.data(), (size_t)(n * eb)) == 0 && v1 == v3,
               "timed sort kt=%d n=%lld", kt, (long long)n);
        for (int threads : {1, 3, 8}) {
          std::vector<uint8_t> k8 = keys;
          std::vector<uint32_t> v8 = vals;
          (void)oracle_stable_sort_threads_timed(kt, k8.data(), v8.data(), n, greater, threads);
          EXPECT(std::memcmp(k1.data(), k8.data(), (size_t)(n * eb)) == 0 && v1 == v8,
                 "threaded sort kt=%d n=%lld threads=%d", kt, (long long)n, threads);
        }
        // Merge of the two sorted halves, and merge path at every diagonal
        // for both bounds.
        const int64_t na = n / 3, nb = n - na;
        std::vector<uint8_t> a(k1.begin(), k1.begin() + na * eb), b(k1.begin() + na * eb,
                                                                    k1.begin() + n * eb);
        std::vector<uint8_t> sa = a, sb = b;
        std::vector<uint32_t> va((size_t)na + 1, 1u), vb((size_t)nb + 1, 2u);
        oracle_stable_sort(kt, sa.data(), va.data(), na, greater);
        oracle_stable_sort(kt, sb.data(), vb.data(), nb, greater);
        std::vector<uint8_t> c((size_t)(n * eb + 8));
        std::vector<uint32_t> cv((size_t)n + 1);
        oracle_merge(kt, sa.data(), va.data(), na, sb.data(), vb.data(), nb, c.data(), cv.data(),
                     greater);
        for (int64_t d = 0; d <= n; ++d) {
          const int64_t lo = oracle_merge_path(kt, sa.data(), na, sb.data(), nb, d, greater, 0);
          const int64_t up = oracle_merge_path(kt, sa.data(), na, sb.data(), nb, d, greater, 1);
          int64_t from_a = 0;
          for (int64_t i = 0; i < d; ++i) from_a += cv[i] == 1u;
          EXPECT(lo == from_a, "merge path kt=%d n=%lld d=%lld", kt, (long long)n, (long long)d);
          EXPECT(up <= lo && up >= 0 && up <= na, "upper bound kt=%d d=%lld", kt, (long long)d);
        }
        // Radix schedule (per-tile counts and offsets) vs stable sort by bits.
        std::vector<uint8_t> k4 = keys, k5 = keys;
        std::vector<uint32_t> v4 = vals, v5 = vals;
        const int64_t nv = 64, tiles = (n + nv - 1) / nv;
        std::vector<uint32_t> counts((size_t)tiles * 256 + 1);
        std::vector<uint64_t> offs((size_t)tiles * 256 + 1);
        oracle_radix_schedule(kt, k4.data(), v4.data(), n, nv, counts.data(), offs.data());
        oracle_stable_sort_radix(kt, k5.data(), v5.data(), n);
        EXPECT(std::memcmp(k4.data(), k5.data(), (size_t)(n * eb)) == 0 && v4 == v5,
               "radix schedule kt=%d n=%lld", kt, (long long)n);
        std::vector<uint8_t> bits((size_t)(n * eb + 8)), back((size_t)(n * eb + 8));
        oracle_radix_bits(kt, keys.data(), bits.data(), n, 0);
        oracle_radix_bits(kt, bits.data(), back.data(), n, 1);
        EXPECT(std::memcmp(keys.data(), back.data(), (size_t)(n * eb)) == 0,
               "radix bits round trip kt=%d", kt);
        // Segmented sort with empty segments, starts past the end and a
        // start at the int32 limit.
        std::vector<int32_t> segs = {0, 0, (int32_t)(n / 2), (int32_t)n, (int32_t)(n + 5),
                                     2147483647};
        std::vector<uint8_t> k6 = keys;
        std::vector<uint32_t> v6 = vals;
        oracle_segmented_stable_sort(kt, k6.data(), v6.data(), n, segs.data(),
                                     (int64_t)segs.size(), greater);
        std::vector<uint8_t> k7 = keys;
        std::vector<uint32_t> v7 = vals;
        const int64_t half = n / 2;
        oracle_stable_sort(kt, k7.data(), v7.data(), half, greater);
        oracle_stable_sort(kt, k7.data() + half * eb, v7.data() + half, n - half, greater);
        EXPECT(std::memcmp(k6.data(), k7.data(), (size_t)(n * eb)) == 0 && v6 == v7,
               "segmented sort kt=%d n=%lld", kt, (long long)n);
      }
    }
  }
  // Scans, reduces and generators.
  for (int64_t n : sizes) {
    std::vector<uint32_t> x((size_t)n + 1), e((size_t)n + 1), i1((size_t)n + 1), s1((size_t)n + 1);
    oracle_generate(x.data(), n, 4, 2, 3, 0);
    oracle_exclusive_scan_u32(x.data(), e.data(), n, 5);
    oracle_inclusive_scan_u32(x.data(), i1.data(), n, 5);
    std::vector<uint32_t> sched = x;
    oracle_scan_schedule_u32(sched.data(), n, 16, 5);
    for (int64_t i = 0; i < n; ++i) {
      EXPECT(i1[i] == e[i] + x[i], "inclusive vs exclusive at %lld", (long long)i);
      EXPECT(sched[i] == e[i], "scan schedule at %lld", (long long)i);
    }
    EXPECT(oracle_reduce_u32(x.data(), n, 5) == (n ? i1[n - 1] : 5u), "reduce n=%lld", (long long)n);
    std::vector<float> f((size_t)n + 1), f32((size_t)n + 1);
    std::vector<double> fi((size_t)n + 1), fe((size_t)n + 1);
    oracle_generate(f.data(), n, 4, 3, 4, 0);
    oracle_inclusive_scan_f64(f.data(), fi.data(), n, 0.0);
    oracle_exclusive_scan_f64(f.data(), fe.data(), n, 0.0);
    oracle_inclusive_scan_f32(f.data(), f32.data(), n);
    EXPECT(n == 0 || oracle_reduce_f64(f.data(), n, 0.0) == fi[n - 1], "reduce f64");
    std::vector<int32_t> segs = {0, 3, 3, (int32_t)n, (int32_t)(n + 9)};
    std::vector<uint32_t> ss((size_t)n + 1);
    std::vector<double> sf((size_t)n + 1);
    oracle_segmented_scan_u32(x.data(), ss.data(), n, segs.data(), (int64_t)segs.size(), 0, 7);
    oracle_segmented_scan_f64(f.data(), sf.data(), n, segs.data(), (int64_t)segs.size(), 1, 0.0);
    if (n > 3) EXPECT(ss[3] == 7u && ss[0] == 7u, "segmented scan restarts");
  }
  std::vector<uint32_t> r(16);
  oracle_glibc_rand(16, r.data());
  EXPECT(r[0] == 1804289383u && r[1] == 846930886u, "glibc rand known answer");
  std::vector<uint32_t> srt = r;
  oracle_std_sort_u32(srt.data(), 16);
  for (int i = 1; i < 16; ++i) EXPECT(srt[i - 1] <= srt[i], "std::sort");
  std::printf(failures ? "oracle sanitize: %d FAILURES\n" : "oracle sanitize: ALL PASS\n",
              failures);
  return failures ? 1 : 0;
}
