// CPU oracle for the mgpu sort / merge / scan path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// or as the timed CPU baseline -- never as the product path.
//
// Two kinds of functions:
//   * standard oracles: std::stable_sort, std::merge, std::exclusive_scan /
//     (and a multi-threaded stable sort built from them, a separately
//     labelled CPU baseline: std::stable_sort of one chunk per thread, then
//     std::merge rounds, each merge split across the threads by merge path),
//     std::inclusive_scan in uint32 modular arithmetic, float64 prefix sums,
//     std::reduce, and their per-segment forms for the segmented sort and
//     scan (the reference's own CPU comparators: std::sort in
//     examples/vk_radix2/vk_radix2.cxx:32, std::sort before merging in
//     examples/vk_merge/vk_merge.cxx:26-27);
//   * restatements of the reference's algorithms, run sequentially on the CPU:
//       oracle_merge_path             inc/mgpu/common/cta_merge.hxx:10-31
//       oracle_mergesort_range        inc/mgpu/common/cta_mergesort.hxx:17-68
//       oracle_mergesort_schedule     inc/mgpu/vk/mergesort.hxx:10-84 with
//                                     common/kernel_mergesort.hxx:12-121,
//                                     common/cta_mergesort.hxx:70-134,
//                                     common/sort_networks.hxx:11-39,
//                                     common/cta_merge.hxx:41-162
//       oracle_radix_schedule         inc/mgpu/vk/radix.hxx:195-470 (per-tile
//                                     digit counts, (bin, tile) offsets,
//                                     stable scatter), 8-bit digits
//       oracle_radix_bits             inc/mgpu/common/cta_radix.hxx:6-64
//       oracle_scan_schedule_u32      inc/mgpu/vk/scan.hxx:10-115
//   Each restatement is pinned by tests against the standard oracles, closed
//   forms and brute force (tests/test_oracle.py).
//
// Inputs: counter-based SplitMix64 generators identical to the device ones
// (oracle_generate), and glibc rand() with its default seed for the
// reference's example inputs (oracle_glibc_rand).

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>
#include <thread>
#include <vector>

namespace {

// Planted mistakes for sensitivity tests of the checks themselves (0 = none):
//   1: merge path with bounds_upper (ties to B) everywhere -> unstable
//   2: mergesort range fix-up applied to the last CTA of a coop group too
//      (cta_mergesort.hxx:54-62 must skip it there)
//   3: radix offsets in (tile, bin) order instead of (bin, tile)
//   4: scan schedule drops the carry-in of the downsweep
int g_plant = 0;

enum { KT_U32 = 0, KT_I32 = 1, KT_U64 = 2, KT_F32 = 3, KT_I64 = 4, KT_F64 = 5 };

// ---------------------------------------------------------------- generators
uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
uint32_t word32(uint64_t seed, uint64_t stream, uint64_t i) {
  return (uint32_t)(mix64((seed << 40) + (stream << 36) + i + 0x9E3779B97F4A7C15ull) >> 32);
}

// --------------------------------------------------------------- radix bits
template <typename K> struct ubits;
template <> struct ubits<uint32_t> { typedef uint32_t u; typedef int32_t s; };
template <> struct ubits<int32_t> { typedef uint32_t u; typedef int32_t s; };
template <> struct ubits<uint64_t> { typedef uint64_t u; typedef int64_t s; };
template <> struct ubits<float> { typedef uint32_t u; typedef int32_t s; };
template <> struct ubits<int64_t> { typedef uint64_t u; typedef int64_t s; };
template <> struct ubits<double> { typedef uint64_t u; typedef int64_t s; };

template <typename K>
typename ubits<K>::u to_bits(K x) {
  typedef typename ubits<K>::u U;
  typedef typename ubits<K>::s S;
  const U msb = (U)1 << (8 * sizeof(K) - 1);
  U y;
  std::memcpy(&y, &x, sizeof(y));
  if (std::is_same<K, uint32_t>::value || std::is_same<K, uint64_t>::value) return y;
  if (std::is_same<K, int32_t>::value || std::is_same<K, int64_t>::value) return y ^ msb;
  // float / double: flip everything when negative, else only the sign bit.
  U mask = (U)((S)y >> (8 * sizeof(K) - 1)) | msb;
  return y ^ mask;
}
template <typename K>
K from_bits(typename ubits<K>::u y) {
  typedef typename ubits<K>::u U;
  typedef typename ubits<K>::s S;
  const U msb = (U)1 << (8 * sizeof(K) - 1);
  if (std::is_same<K, int32_t>::value || std::is_same<K, int64_t>::value) y ^= msb;
  if (std::is_same<K, float>::value || std::is_same<K, double>::value) {
    y ^= msb;
    U mask = (U)((S)y >> (8 * sizeof(K) - 1)) & (msb - 1);
    y ^= mask;
  }
  K x;
  std::memcpy(&x, &y, sizeof(x));
  return x;
}

template <typename K>
struct comp_t {
  int greater;
  bool operator()(K a, K b) const { return greater ? (b < a) : (a < b); }
};

// ----------------------------------------------------------- std:: oracles
// Returns the seconds std::stable_sort itself took (steady_clock), packing
// the (key, value) pairs and unpacking them excluded.
template <typename K>
double stable_sort_kv(K* keys, uint32_t* vals, int64_t n, int greater) {
  std::vector<std::pair<K, uint32_t>> v((size_t)n);
  for (int64_t i = 0; i < n; ++i) v[i] = {keys[i], vals ? vals[i] : 0u};
  comp_t<K> c{greater};
  const auto t0 = std::chrono::steady_clock::now();
  std::stable_sort(v.begin(), v.end(),
                   [&](const std::pair<K, uint32_t>& a, const std::pair<K, uint32_t>& b) {
                     return c(a.first, b.first);
                   });
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int64_t i = 0; i < n; ++i) {
    keys[i] = v[i].first;
    if (vals) vals[i] = v[i].second;
  }
  return dt;
}

// Stable merge of sorted a and b into out (ties take a), split across
// `threads` std::threads at merge-path diagonals: each thread std::merge()s
// the slices between its two diagonals.
template <typename P, typename Cmp>
void parallel_merge(const P* a, int64_t na, const P* b, int64_t nb, P* out, int threads, Cmp cmp) {
  auto split = [&](int64_t d) {
    int64_t lo = std::max<int64_t>(0, d - nb), hi = std::min(d, na);
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (!cmp(b[d - 1 - mid], a[mid])) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t) {
    const int64_t d0 = (na + nb) * t / threads, d1 = (na + nb) * (t + 1) / threads;
    const int64_t a0 = split(d0), a1 = split(d1);
    ts.emplace_back([=] { std::merge(a + a0, a + a1, b + (d0 - a0), b + (d1 - a1), out + d0, cmp); });
  }
  for (auto& t : ts) t.join();
}

// Multi-threaded stable sort of (key, value) pairs: one chunk per thread
// sorted by std::stable_sort, then rounds of parallel_merge of adjacent runs
// (the earlier run first, so ties keep their order).  Returns the seconds
// from the first chunk sort to the last merge.
template <typename K>
double stable_sort_kv_threads(K* keys, uint32_t* vals, int64_t n, int greater, int threads) {
  typedef std::pair<K, uint32_t> P;
  std::vector<P> v((size_t)n), w((size_t)n);
  for (int64_t i = 0; i < n; ++i) v[i] = {keys[i], vals ? vals[i] : 0u};
  comp_t<K> c{greater};
  auto cmp = [&](const P& a, const P& b) { return c(a.first, b.first); };
  threads = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n));
  std::vector<int64_t> bounds;
  for (int t = 0; t <= threads; ++t) bounds.push_back(n * t / threads);
  const auto t0 = std::chrono::steady_clock::now();
  {
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
      ts.emplace_back([&, t] { std::stable_sort(v.begin() + bounds[t], v.begin() + bounds[t + 1], cmp); });
    for (auto& t : ts) t.join();
  }
  P* src = v.data();
  P* dst = w.data();
  while (bounds.size() > 2) {
    std::vector<int64_t> next;
    for (size_t r = 0; r + 1 < bounds.size(); r += 2) {
      next.push_back(bounds[r]);
      if (r + 2 < bounds.size())
        parallel_merge(src + bounds[r], bounds[r + 1] - bounds[r], src + bounds[r + 1],
                       bounds[r + 2] - bounds[r + 1], dst + bounds[r], threads, cmp);
      else
        std::copy(src + bounds[r], src + bounds[r + 1], dst + bounds[r]);
    }
    next.push_back(n);
    bounds.swap(next);
    std::swap(src, dst);
  }
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int64_t i = 0; i < n; ++i) {
    keys[i] = src[i].first;
    if (vals) vals[i] = src[i].second;
  }
  return dt;
}

template <typename K>
void stable_sort_radix(K* keys, uint32_t* vals, int64_t n) {
  typedef typename ubits<K>::u U;
  std::vector<std::pair<U, uint32_t>> v((size_t)n);
  for (int64_t i = 0; i < n; ++i) v[i] = {to_bits(keys[i]), vals ? vals[i] : 0u};
  std::stable_sort(v.begin(), v.end(),
                   [](const std::pair<U, uint32_t>& a, const std::pair<U, uint32_t>& b) {
                     return a.first < b.first;
                   });
  for (int64_t i = 0; i < n; ++i) {
    keys[i] = from_bits<K>(v[i].first);
    if (vals) vals[i] = v[i].second;
  }
}

template <typename K>
void merge_kv(const K* a, const uint32_t* av, int64_t na, const K* b, const uint32_t* bv,
              int64_t nb, K* c, uint32_t* cv, int greater) {
  std::vector<std::pair<K, uint32_t>> A((size_t)na), B((size_t)nb), C((size_t)(na + nb));
  for (int64_t i = 0; i < na; ++i) A[i] = {a[i], av ? av[i] : 0u};
  for (int64_t i = 0; i < nb; ++i) B[i] = {b[i], bv ? bv[i] : 0u};
  comp_t<K> cmp{greater};
  std::merge(A.begin(), A.end(), B.begin(), B.end(), C.begin(),
             [&](const std::pair<K, uint32_t>& x, const std::pair<K, uint32_t>& y) {
               return cmp(x.first, y.first);
             });
  for (int64_t i = 0; i < na + nb; ++i) {
    c[i] = C[i].first;
    if (cv) cv[i] = C[i].second;
  }
}

// ------------------------------------------------------- merge-path restated
// inc/mgpu/common/cta_merge.hxx:10-31.  bounds 0 = lower (ties to A), 1 = upper.
template <typename K>
int64_t merge_path(const K* a, int64_t na, const K* b, int64_t nb, int64_t diag, comp_t<K> comp,
                   int bounds) {
  int64_t begin = std::max<int64_t>(0, diag - nb);
  int64_t end = std::min<int64_t>(diag, na);
  while (begin < end) {
    int64_t mid = (begin + end) / 2;
    K ak = a[mid], bk = b[diag - 1 - mid];
    bool pred = bounds ? comp(ak, bk) : !comp(bk, ak);
    if (pred) begin = mid + 1;
    else end = mid;
  }
  return begin;
}

struct range4 { int64_t a_begin, a_end, b_begin, b_end; };

// inc/mgpu/common/cta_mergesort.hxx:17-43.
range4 mergesort_range(int64_t count, int64_t partition, int64_t coop, int64_t spacing) {
  int64_t size = spacing * (coop / 2);
  int64_t start = ~(coop - 1) & partition;
  int64_t a_begin = spacing * start;
  int64_t b_begin = spacing * start + size;
  return range4{a_begin, std::min(count, a_begin + size), std::min(count, b_begin),
                std::min(count, b_begin + size)};
}
// inc/mgpu/common/cta_mergesort.hxx:45-68.
range4 mergesort_range_mp(int64_t count, int64_t partition, int64_t coop, int64_t spacing,
                          int64_t mp0, int64_t mp1) {
  range4 r = mergesort_range(count, partition, coop, spacing);
  int64_t diag = spacing * partition - r.a_begin;
  if (g_plant == 2 || coop - 1 != ((coop - 1) & partition)) {
    r.a_end = r.a_begin + mp1;
    r.b_end = std::min(count, r.b_begin + diag + spacing - mp1);
  }
  r.a_begin = r.a_begin + mp0;
  r.b_begin = std::min(count, r.b_begin + diag - mp0);
  return r;
}

// serial_merge over a shared key array (cta_merge.hxx:110-133), vt outputs,
// with merge_predicate's range checks (cta_merge.hxx:41-50); bounds_lower.
template <typename K>
void serial_merge(const std::vector<K>& keys, int64_t a_begin, int64_t a_end, int64_t b_begin,
                  int64_t b_end, int vt, comp_t<K> comp, K* out_keys, int64_t* out_idx) {
  auto at = [&](int64_t i) { return i < (int64_t)keys.size() ? keys[i] : K(); };
  K ak = at(a_begin), bk = at(b_begin);
  for (int i = 0; i < vt; ++i) {
    bool p;
    if (a_begin >= a_end) p = false;
    else if (b_begin >= b_end) p = true;
    else p = !comp(bk, ak);
    int64_t index = p ? a_begin : b_begin;
    out_keys[i] = p ? ak : bk;
    out_idx[i] = index;
    K ck = at(index + 1);
    if (p) { ak = ck; a_begin = index + 1; }
    else { bk = ck; b_begin = index + 1; }
  }
}

// Odd-even transposition sort, swaps only on strict comp (sort_networks.hxx:
// 24-39); slots >= valid never move.
template <typename K>
void odd_even_sort(K* k, uint32_t* v, int vt, int valid, comp_t<K> comp) {
  for (int I = 0; I < vt; ++I)
    for (int i = 1 & I; i < vt - 1; i += 2)
      if (i + 1 < valid && comp(k[i + 1], k[i])) {
        std::swap(k[i], k[i + 1]);
        std::swap(v[i], v[i + 1]);
      }
}

// The reference mergesort schedule, executed sequentially.
template <typename K>
void mergesort_schedule(K* keys, uint32_t* vals, int64_t count, int nt, int vt, comp_t<K> comp,
                        int sort_indices) {
  const int64_t nv = (int64_t)nt * vt;
  const int64_t num_ctas = (count + nv - 1) / nv;
  std::vector<K> k(keys, keys + count);
  std::vector<uint32_t> v((size_t)count);
  for (int64_t i = 0; i < count; ++i) v[i] = sort_indices ? (uint32_t)i : (vals ? vals[i] : 0u);

  // Blocksort (kernel_mergesort.hxx:12-67, cta_mergesort.hxx:116-133).
  for (int64_t cta = 0; cta < num_ctas; ++cta) {
    const int64_t base = cta * nv;
    const int tile = (int)std::min<int64_t>(nv, count - base);
    std::vector<K> tk(nv + 1);
    std::vector<uint32_t> tv(nv + 1);
    for (int j = 0; j < tile; ++j) { tk[j] = k[base + j]; tv[j] = v[base + j]; }
    for (int t = 0; t < nt; ++t) {
      int valid = std::max(0, std::min(vt, tile - vt * t));
      odd_even_sort(&tk[vt * t], &tv[vt * t], vt, valid, comp);
    }
    for (int pass = 0; (1 << pass) < nt; ++pass) {
      const int coop = 2 << pass;
      std::vector<K> nk(nv + 1);
      std::vector<uint32_t> nvv(nv + 1);
      for (int t = 0; t < nt; ++t) {
        range4 r = mergesort_range(tile, t, coop, vt);
        int64_t diag = (int64_t)vt * t - r.a_begin;
        int64_t mp = merge_path(tk.data() + r.a_begin, r.a_end - r.a_begin, tk.data() + r.b_begin,
                                r.b_end - r.b_begin, diag, comp, g_plant == 1);
        std::vector<K> ok(vt);
        std::vector<int64_t> oi(vt);
        serial_merge(tk, r.a_begin + mp, r.a_end, r.b_begin + diag - mp, r.b_end, vt, comp,
                     ok.data(), oi.data());
        for (int i = 0; i < vt; ++i) {
          nk[vt * t + i] = ok[i];
          nvv[vt * t + i] = tv[std::min<int64_t>(oi[i], nv)];
        }
      }
      tk.swap(nk);
      tv.swap(nvv);
    }
    for (int j = 0; j < tile; ++j) { k[base + j] = tk[j]; v[base + j] = tv[j]; }
  }

  // Global passes (vk/mergesort.hxx:61-82).
  int num_passes = 0;
  while (((int64_t)1 << num_passes) < num_ctas) ++num_passes;
  std::vector<int64_t> mp((size_t)num_ctas + 1);
  for (int pass = 0; pass < num_passes; ++pass) {
    const int64_t coop = (int64_t)2 << pass;
    for (int64_t i = 0; i <= num_ctas; ++i) {
      range4 r = mergesort_range(count, i, coop, nv);
      int64_t diag = std::min(nv * i, count) - r.a_begin;
      mp[i] = merge_path(k.data() + r.a_begin, r.a_end - r.a_begin, k.data() + r.b_begin,
                         r.b_end - r.b_begin, diag, comp, g_plant == 1);
    }
    std::vector<K> k2((size_t)count);
    std::vector<uint32_t> v2((size_t)count);
    for (int64_t cta = 0; cta < num_ctas; ++cta) {
      range4 r = mergesort_range_mp(count, cta, coop, nv, mp[cta], mp[cta + 1]);
      const int64_t base = cta * nv;
      const int64_t tile = std::min<int64_t>(nv, count - base);
      // Concatenate the A and B slices as the tile merge does in LDS.
      std::vector<K> sk;
      std::vector<uint32_t> sv;
      for (int64_t i = r.a_begin; i < r.a_end; ++i) { sk.push_back(k[i]); sv.push_back(v[i]); }
      int64_t a_cnt = (int64_t)sk.size();
      for (int64_t i = r.b_begin; i < r.b_end; ++i) { sk.push_back(k[i]); sv.push_back(v[i]); }
      int64_t total = (int64_t)sk.size();
      std::vector<K> ok((size_t)total);
      std::vector<int64_t> oi((size_t)total);
      if (total) serial_merge(sk, 0, a_cnt, a_cnt, total, (int)total, comp, ok.data(), oi.data());
      for (int64_t j = 0; j < tile && j < total; ++j) {
        k2[base + j] = ok[j];
        v2[base + j] = sv[oi[j]];
      }
    }
    k.swap(k2);
    v.swap(v2);
  }
  std::copy(k.begin(), k.end(), keys);
  if (vals)
    for (int64_t i = 0; i < count; ++i) vals[i] = v[i];
}

// LSD radix schedule with 8-bit digits and tiles of nv keys
// (vk/radix.hxx:310-467): per-tile digit counts, offsets in (bin, tile) order,
// stable scatter.  If counts_out/offsets_out are given, the first pass's
// [tile][256] counts and offsets are written there.
template <typename K>
void radix_schedule(K* keys, uint32_t* vals, int64_t count, int64_t nv, uint32_t* counts_out,
                    uint64_t* offsets_out) {
  typedef typename ubits<K>::u U;
  const int64_t tiles = (count + nv - 1) / nv;
  std::vector<U> b((size_t)count), b2((size_t)count);
  std::vector<uint32_t> v((size_t)count), v2((size_t)count);
  for (int64_t i = 0; i < count; ++i) { b[i] = to_bits(keys[i]); v[i] = vals ? vals[i] : 0u; }
  std::vector<uint64_t> cnt((size_t)tiles * 256), off((size_t)tiles * 256);
  for (int shift = 0; shift < (int)(8 * sizeof(K)); shift += 8) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t i = 0; i < count; ++i) cnt[(i / nv) * 256 + ((b[i] >> shift) & 255)]++;
    uint64_t run = 0;
    if (g_plant == 3) {
      for (int64_t t = 0; t < tiles; ++t)
        for (int d = 0; d < 256; ++d) {
          off[t * 256 + d] = run;
          run += cnt[t * 256 + d];
        }
    } else {
      for (int d = 0; d < 256; ++d)
        for (int64_t t = 0; t < tiles; ++t) {
          off[t * 256 + d] = run;
          run += cnt[t * 256 + d];
        }
    }
    if (shift == 0) {
      if (counts_out)
        for (size_t i = 0; i < cnt.size(); ++i) counts_out[i] = (uint32_t)cnt[i];
      if (offsets_out) std::copy(off.begin(), off.end(), offsets_out);
    }
    for (int64_t i = 0; i < count; ++i) {
      uint64_t& o = off[(i / nv) * 256 + ((b[i] >> shift) & 255)];
      b2[o] = b[i];
      v2[o] = v[i];
      ++o;
    }
    b.swap(b2);
    v.swap(v2);
  }
  for (int64_t i = 0; i < count; ++i) {
    keys[i] = from_bits<K>(b[i]);
    if (vals) vals[i] = v[i];
  }
}

// vk::scan (vk/scan.hxx:10-115), exclusive, in place, uint32 arithmetic:
// <= 8 tiles: one CTA loops over the tiles with a carry; otherwise tile
// reductions, a recursive scan of the partials, and a downsweep with carry-in.
void scan_schedule_u32(uint32_t* data, int64_t count, int64_t nv, uint32_t init) {
  int64_t num_ctas = (count + nv - 1) / nv;
  if (num_ctas <= 8) {
    uint32_t carry = init;
    for (int64_t i = 0; i < count; ++i) {
      uint32_t x = data[i];
      data[i] = carry;
      carry += x;
    }
    return;
  }
  std::vector<uint32_t> partials((size_t)num_ctas);
  for (int64_t c = 0; c < num_ctas; ++c) {
    uint32_t s = 0;
    for (int64_t i = c * nv; i < std::min(count, (c + 1) * nv); ++i) s += data[i];
    partials[c] = s;
  }
  scan_schedule_u32(partials.data(), num_ctas, nv, init);
  for (int64_t c = 0; c < num_ctas; ++c) {
    uint32_t carry = g_plant == 4 ? 0u : partials[c];
    for (int64_t i = c * nv; i < std::min(count, (c + 1) * nv); ++i) {
      uint32_t x = data[i];
      data[i] = carry;
      carry += x;
    }
  }
}

template <typename F>
void dispatch_key(int kt, F f) {
  switch (kt) {
    case KT_U32: f((uint32_t*)nullptr); break;
    case KT_I32: f((int32_t*)nullptr); break;
    case KT_U64: f((uint64_t*)nullptr); break;
    case KT_F32: f((float*)nullptr); break;
    case KT_I64: f((int64_t*)nullptr); break;
    case KT_F64: f((double*)nullptr); break;
  }
}

}  // namespace

extern "C" {

int oracle_version(void) { return 1; }

void oracle_set_plant(int plant) { g_plant = plant; }

void oracle_generate(void* out, int64_t count, int elem_bytes, int kind, uint64_t seed,
                     uint64_t offset) {
  for (int64_t j = 0; j < count; ++j) {
    uint64_t i = offset + (uint64_t)j, v = 0;
    switch (kind) {
      case 0: v = word32(seed, 0, i); break;
      case 1: v = word32(seed, 0, i) & word32(seed, 1, i) & word32(seed, 2, i) & word32(seed, 3, i); break;
      case 2: v = word32(seed, 0, i) >> 28; break;
      case 3: {
        float f = (float)(word32(seed, 0, i) >> 8) * (1.0f / 16777216.0f);
        uint32_t u; std::memcpy(&u, &f, 4); v = u; break;
      }
      case 4: v = ((uint64_t)word32(seed, 0, i) << 32) | word32(seed, 1, i); break;
      case 5: {
        float f = (float)(word32(seed, 0, i) >> 8) * (1.0f / 8388608.0f) - 1.0f;
        uint32_t u; std::memcpy(&u, &f, 4); v = u; break;
      }
      case 6: v = i; break;
    }
    if (elem_bytes == 8) ((uint64_t*)out)[j] = v;
    else ((uint32_t*)out)[j] = (uint32_t)v;
  }
}

// glibc rand() with the default seed (srand(1)), as the reference's examples
// use it without seeding (examples/vk_radix2/vk_radix2.cxx:25-27).
void oracle_glibc_rand(int64_t n, uint32_t* out) {
  srand(1);
  for (int64_t i = 0; i < n; ++i) out[i] = (uint32_t)rand();
}

void oracle_stable_sort(int kt, void* keys, uint32_t* vals, int64_t n, int greater) {
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    stable_sort_kv((K*)keys, vals, n, greater);
  });
}
// The same sort; returns the seconds std::stable_sort took (the CPU baseline).
double oracle_stable_sort_timed(int kt, void* keys, uint32_t* vals, int64_t n, int greater) {
  double dt = 0;
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    dt = stable_sort_kv((K*)keys, vals, n, greater);
  });
  return dt;
}

// The multi-threaded stable sort (same result); returns its seconds.
double oracle_stable_sort_threads_timed(int kt, void* keys, uint32_t* vals, int64_t n,
                                        int greater, int threads) {
  double dt = 0;
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    dt = stable_sort_kv_threads((K*)keys, vals, n, greater, threads);
  });
  return dt;
}

// One std::stable_sort per segment [segs[i], segs[i+1]) (last ends at n),
// both ends clamped to [0, n]: a start at or beyond n gives an empty segment.
void oracle_segmented_stable_sort(int kt, void* keys, uint32_t* vals, int64_t n,
                                  const int32_t* segs, int64_t num_segments, int greater) {
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    for (int64_t i = 0; i < num_segments; ++i) {
      int64_t b = std::min<int64_t>(std::max<int64_t>(segs[i], 0), n);
      int64_t e = i + 1 < num_segments ? segs[i + 1] : n;
      e = std::min<int64_t>(std::max<int64_t>(e, b), n);
      if (e > b) stable_sort_kv((K*)keys + b, vals ? vals + b : nullptr, e - b, greater);
    }
  });
}

void oracle_stable_sort_radix(int kt, void* keys, uint32_t* vals, int64_t n) {
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    stable_sort_radix((K*)keys, vals, n);
  });
}

// std::sort of uint32 keys (the reference's radix check, vk_radix2.cxx:32).
void oracle_std_sort_u32(uint32_t* keys, int64_t n) { std::sort(keys, keys + n); }

void oracle_merge(int kt, const void* a, const uint32_t* av, int64_t na, const void* b,
                  const uint32_t* bv, int64_t nb, void* c, uint32_t* cv, int greater) {
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    merge_kv((const K*)a, av, na, (const K*)b, bv, nb, (K*)c, cv, greater);
  });
}

void oracle_exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t init) {
  std::exclusive_scan(in, in + n, out, init);
}
void oracle_inclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t init) {
  std::inclusive_scan(in, in + n, out, std::plus<uint32_t>(), init);
}
// Float64 prefix sums of float32 data: exact for inputs that are multiples of
// 2^-24 with prefix below 2^29.
void oracle_inclusive_scan_f64(const float* in, double* out, int64_t n, double init) {
  double run = init;
  for (int64_t i = 0; i < n; ++i) { run += (double)in[i]; out[i] = run; }
}
void oracle_exclusive_scan_f64(const float* in, double* out, int64_t n, double init) {
  double run = init;
  for (int64_t i = 0; i < n; ++i) { out[i] = run; run += (double)in[i]; }
}
// Segmented scans: std::exclusive_scan / std::inclusive_scan from init over
// every segment [segs[i], segs[i+1]) (the last ends at n); element 0 always
// starts a segment, so [0, segs[0]) is one too.
}  // extern "C"
template <typename F>
static void for_each_segment(int64_t n, const int32_t* segs, int64_t num_segments, F f) {
  int64_t b = 0;
  for (int64_t i = 0; i <= num_segments; ++i) {
    int64_t e = i < num_segments ? std::min<int64_t>(std::max<int64_t>(segs[i], 0), n) : n;
    if (e > b) {
      f(b, e);
      b = e;
    }
  }
}
extern "C" {
void oracle_segmented_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, const int32_t* segs,
                               int64_t num_segments, int inclusive, uint32_t init) {
  for_each_segment(n, segs, num_segments, [&](int64_t b, int64_t e) {
    if (inclusive) std::inclusive_scan(in + b, in + e, out + b, std::plus<uint32_t>(), init);
    else std::exclusive_scan(in + b, in + e, out + b, init);
  });
}
void oracle_segmented_scan_f64(const float* in, double* out, int64_t n, const int32_t* segs,
                               int64_t num_segments, int inclusive, double init) {
  for_each_segment(n, segs, num_segments, [&](int64_t b, int64_t e) {
    if (inclusive) oracle_inclusive_scan_f64(in + b, out + b, e - b, init);
    else oracle_exclusive_scan_f64(in + b, out + b, e - b, init);
  });
}
// std::inclusive_scan in float32 (the CPU baseline as timed).
void oracle_inclusive_scan_f32(const float* in, float* out, int64_t n) {
  std::inclusive_scan(in, in + n, out);
}
uint32_t oracle_reduce_u32(const uint32_t* in, int64_t n, uint32_t init) {
  return std::reduce(in, in + n, init);
}
double oracle_reduce_f64(const float* in, int64_t n, double init) {
  double s = init;
  for (int64_t i = 0; i < n; ++i) s += (double)in[i];
  return s;
}
float oracle_reduce_f32(const float* in, int64_t n) { return std::reduce(in, in + n, 0.0f); }

int64_t oracle_merge_path(int kt, const void* a, int64_t na, const void* b, int64_t nb,
                          int64_t diag, int greater, int bounds) {
  int64_t r = 0;
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    r = merge_path((const K*)a, na, (const K*)b, nb, diag, comp_t<K>{greater}, bounds);
  });
  return r;
}

void oracle_mergesort_range(int64_t count, int64_t partition, int64_t coop, int64_t spacing,
                            int64_t mp0, int64_t mp1, int use_mp, int64_t* out4) {
  range4 r = use_mp ? mergesort_range_mp(count, partition, coop, spacing, mp0, mp1)
                    : mergesort_range(count, partition, coop, spacing);
  out4[0] = r.a_begin; out4[1] = r.a_end; out4[2] = r.b_begin; out4[3] = r.b_end;
}

void oracle_mergesort_schedule(int kt, void* keys, uint32_t* vals, int64_t n, int nt, int vt,
                               int greater, int sort_indices) {
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    mergesort_schedule((K*)keys, vals, n, nt, vt, comp_t<K>{greater}, sort_indices);
  });
}

void oracle_radix_schedule(int kt, void* keys, uint32_t* vals, int64_t n, int64_t nv,
                           uint32_t* counts_out, uint64_t* offsets_out) {
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    radix_schedule((K*)keys, vals, n, nv, counts_out, offsets_out);
  });
}

void oracle_radix_bits(int kt, const void* keys, void* bits, int64_t n, int inverse) {
  dispatch_key(kt, [&](auto* tag) {
    typedef std::remove_pointer_t<decltype(tag)> K;
    typedef typename ubits<K>::u U;
    for (int64_t i = 0; i < n; ++i) {
      if (inverse) ((K*)bits)[i] = from_bits<K>(((const U*)keys)[i]);
      else ((U*)bits)[i] = to_bits(((const K*)keys)[i]);
    }
  });
}

void oracle_scan_schedule_u32(uint32_t* data, int64_t n, int64_t nv, uint32_t init) {
  scan_schedule_u32(data, n, nv, init);
}

}  // extern "C"
