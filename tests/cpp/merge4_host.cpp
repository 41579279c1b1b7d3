// Host-side check (no GPU) of the four-way merge partitions of
// mgpu/hip/merge.hxx and mgpu/hip/mergesort.hxx against brute force:
//  * merge4_path(diag) on random runs with many ties, both comparators, at
//    every diagonal: the number taken from each run among the first diag
//    outputs of the stable four-way merge (ties to the earlier run);
//  * the two-kernel form used by the four-way passes (merge4_sample for every
//    sample, then merge4_partition for every tile boundary) over whole sort
//    schedules (runs of tile << level, or of half a tile << level, partial
//    last groups), at several sample spacings.
// Prints "ALL PASS" and exits 0 when every check holds.
#include <algorithm>
#include <cstdio>
#include <functional>
#include <random>
#include <vector>

#include "mgpu/hip/mergesort.hxx"

using namespace mgpu::hip;

static long fails = 0;

// Split after d outputs of the stable merge of runs[0..3] by brute force.
template <typename comp_t>
static void brute_split(const std::vector<uint32_t>* runs, const int64_t* n, int64_t d,
                        comp_t comp, int64_t (&out)[4]) {
  int64_t i[4] = {0, 0, 0, 0};
  for (int64_t step = 0; step < d; ++step) {
    int best = -1;
    for (int r = 0; r < 4; ++r)
      if (i[r] < n[r] && (best < 0 || comp(runs[r][i[r]], runs[best][i[best]]))) best = r;
    ++i[best];
  }
  for (int r = 0; r < 4; ++r) out[r] = i[r];
}

template <typename comp_t>
static long check_path(std::mt19937_64& rng, int trial, comp_t comp) {
  int64_t n[4];
  const uint32_t range = 1 + (uint32_t)(rng() % (trial % 3 == 0 ? 3 : 1000));
  std::vector<uint32_t> runs[4];
  for (int r = 0; r < 4; ++r) {
    n[r] = rng() % 5 == 0 ? 0 : (int64_t)(rng() % (1 + (trial % 50) * 3));
    runs[r].resize(n[r] + 1);
    for (int64_t i = 0; i < n[r]; ++i) runs[r][i] = (uint32_t)(rng() % range);
    std::stable_sort(runs[r].begin(), runs[r].begin() + n[r], comp);
  }
  long checks = 0;
  const int64_t total = n[0] + n[1] + n[2] + n[3];
  for (int64_t d = 0; d <= total; ++d) {
    int64_t got[4], want[4];
    merge4_path(runs[0].data(), n[0], runs[1].data(), n[1], runs[2].data(), n[2], runs[3].data(),
                n[3], d, comp, got);
    brute_split(runs, n, d, comp, want);
    ++checks;
    if (!std::equal(got, got + 4, want) && fails++ < 10)
      std::printf("FAIL merge4_path n=%ld,%ld,%ld,%ld d=%ld\n", (long)n[0], (long)n[1],
                  (long)n[2], (long)n[3], (long)d);
  }
  return checks;
}

// With half, the first runs are half a tile long (the merge sort's schedule
// when a half-size blocksort makes the number of levels even).
template <typename comp_t>
static long check_schedule(std::mt19937_64& rng, int64_t count, int64_t nv, int div,
                           uint32_t range, comp_t comp, bool half = false) {
  std::vector<uint32_t> keys(count + 1);
  for (int64_t i = 0; i < count; ++i) keys[i] = (uint32_t)(rng() % range);
  const int64_t tiles = div_up(count, nv);
  const int64_t run0 = half ? nv / 2 : nv;
  const int levels = find_log2(div_up(count, run0), true);
  long checks = 0;
  for (int level = levels & 1; level < levels; level += 2) {
    const int64_t run = run0 << level;
    for (int64_t b = 0; b < count; b += run)
      std::stable_sort(keys.begin() + b, keys.begin() + std::min(count, b + run), comp);
    const int64_t S = nv / div;
    const int64_t per_side = div_up(count, 4 * run) * merge4_samples_per_group(run, S);
    std::vector<int2> samples(2 * per_side);
    for (int64_t t = 0; t < 2 * per_side; ++t)
      samples[t] = merge4_sample(keys.data(), count, run, S, per_side, t, comp);
    for (int64_t c = 0; c <= tiles; ++c) {
      const int4 got =
          merge4_partition(keys.data(), count, run, nv, c, samples.data(), per_side, S, comp);
      const int64_t pos = std::min(c * nv, count), g0 = pos / (4 * run) * 4 * run;
      std::vector<uint32_t> runs[4];
      int64_t n[4];
      for (int r = 0; r < 4; ++r) {
        const int64_t b = std::min(count, g0 + r * run), e = std::min(count, g0 + (r + 1) * run);
        n[r] = e - b;
        runs[r].assign(keys.begin() + b, keys.begin() + e);
      }
      int64_t want[4];
      brute_split(runs, n, pos - g0, comp, want);
      ++checks;
      if ((got.x != want[0] || got.y != want[1] || got.z != want[2] || got.w != want[3]) &&
          fails++ < 10)
        std::printf("FAIL merge4_partition count=%ld nv=%ld level=%d c=%ld\n", (long)count,
                    (long)nv, level, (long)c);
    }
  }
  return checks;
}

int main() {
  std::mt19937_64 rng(7);
  long checks = 0;
  for (int trial = 0; trial < 2000; ++trial)
    checks += trial & 1 ? check_path(rng, trial, less_t<uint32_t>())
                        : check_path(rng, trial, greater_t<uint32_t>());
  for (int trial = 0; trial < 600; ++trial) {
    const int64_t nv = 64 * (1 + trial % 3);
    const int64_t count = 1 + (int64_t)(rng() % (nv * (trial % 7 == 0 ? 300 : 40)));
    const uint32_t range = trial % 4 == 0 ? 2 : trial % 4 == 1 ? 50 : 1u << 30;
    const int div = 1 << (trial % 4);
    checks += trial & 1 ? check_schedule(rng, count, nv, div, range, less_t<uint32_t>())
                        : check_schedule(rng, count, nv, div, range, greater_t<uint32_t>());
    checks += trial & 1 ? check_schedule(rng, count, nv, div, range, greater_t<uint32_t>(), true)
                        : check_schedule(rng, count, nv, div, range, less_t<uint32_t>(), true);
  }
  std::printf("%ld checks, %ld failures\n", checks, fails);
  if (fails) return 1;
  std::printf("ALL PASS\n");
  return 0;
}
