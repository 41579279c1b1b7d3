// Merge-path primitives and the two-way stable merge for gfx950.
//
// merge_path follows inc/mgpu/common/cta_merge.hxx:10-31 with bounds_lower
// (ties go to A): it returns the number of A elements among the first `diag`
// outputs of the stable merge.  The partition search is one binary search
// per thread in global memory, as in vk/partition.hxx:9-20.  Searches with
// more probes per round take fewer dependent rounds but fetch more scattered
// cache lines, and on MI355X the line fetches cost more than the rounds save:
// a wave-wide 64-probe search was four times the traffic and slower, and 3 or
// 7 probes per round took the merge partition of 2 x 2^27 keys from 0.043 to
// 0.061 or 0.093 ms.
//
// A tile merge (kernel_merge.hxx:9-58 and cta_merge.hxx:136-162) stages the A
// and B slices of one output tile in LDS with coalesced loads, runs one merge
// path per thread and a vt-step serial merge (cta_merge.hxx:110-133), and
// stores keys and values through LDS; values are gathered from LDS rather than
// from global memory (the reference gathers from global, cta_merge.hxx:84-105).
#pragma once
#include "meta.hxx"

namespace mgpu {
namespace hip {

// Which input wins ties on a merge-path diagonal (common/cta_merge.hxx:18-31):
// bounds_lower puts A's equal keys first (the stable merge of A then B);
// bounds_upper puts B's first.
enum bounds_t { bounds_lower, bounds_upper };

// Number of A elements among the first `diag` outputs of the merge of sorted
// a and b whose ties follow `bounds`: a binary search over the diagonal.
template <bounds_t bounds = bounds_lower, typename index_t, typename key_t, typename comp_t>
__host__ __device__ inline index_t merge_path_search(const key_t* a, index_t a_count,
                                                     const key_t* b, index_t b_count,
                                                     index_t diag, comp_t comp) {
  index_t begin = diag > b_count ? diag - b_count : 0;
  index_t end = diag < a_count ? diag : a_count;
  while (begin < end) {
    const index_t mid = (begin + end) >> 1;
    const bool pred = bounds == bounds_upper ? comp(a[mid], b[diag - 1 - mid])
                                             : !comp(b[diag - 1 - mid], a[mid]);
    if (pred) begin = mid + 1;
    else end = mid;
  }
  return begin;
}

// Split of the first `diag` outputs of the stable merge of four sorted runs
// r0..r3 (ties go to the earlier run): split[k] = the number taken from r_k.
// The merge is merge(X, Y) with X = merge(r0, r1) and Y = merge(r2, r3), so
// the search is a merge path over X and Y whose probes X[m] and Y[n] are
// themselves merge paths over (r0, r1) and (r2, r3).  a(m), the number of r0
// keys among X's first m, is non-decreasing in m and m - a(m) is too, so
// every probe evaluated so far bounds the inner search of the next: after
// probes at ml < mh, a(m) for ml <= m <= mh lies in
// [max(a(ml), a(mh) - (mh - m)), min(a(mh), a(ml) + (m - ml))]; the same holds
// for Y with b(n), the number of r2 keys among Y's first n.  The two inner
// searches of one probe run interleaved.
//
// merge4_state_t carries the search range [lo, hi] of the answer i (the
// number of X keys among the first diag) and the probes that bound a() and
// b(): any known values a(xlm) = xla, a(xhm) = xha with xlm <= i <= xhm and
// b(yln) = ylb, b(yhn) = yhb with yln <= diag - i <= yhn.
struct merge4_state_t {
  int64_t lo, hi;
  int64_t xlm, xla, xhm, xha;
  int64_t yln, ylb, yhn, yhb;
};

// The state that knows only the run lengths.
__host__ __device__ inline merge4_state_t merge4_initial(int64_t n0, int64_t n1, int64_t n2,
                                                         int64_t n3, int64_t diag) {
  const int64_t nx = n0 + n1, ny = n2 + n3;
  return merge4_state_t{max_<int64_t>(0, diag - ny), min_(diag, nx), 0, 0, nx, n0, 0, 0, ny, n2};
}

// Probes per search round.  With P > 1 a round is a (P+1)-ary search whose
// P probes are independent loads, log_{P+1}(w) dependent rounds instead of
// log2(w); on MI355X the partition kernels are bound by the number of
// scattered cache lines their probes fetch, not by the length of the chain,
// and P = 3 or 7 made them slower (2^28 pairs: partition 0.13 -> 0.20 ms,
// samples 0.05 -> 0.08 ms at P = 3), so the searches are binary.
constexpr int merge4_probes = 1;

template <typename key_t, typename comp_t, int P = merge4_probes>
__host__ __device__ inline void merge4_search(const key_t* r0, int64_t n0, const key_t* r1,
                                              int64_t n1, const key_t* r2, int64_t n2,
                                              const key_t* r3, int64_t n3, int64_t diag,
                                              comp_t comp, merge4_state_t st, int64_t (&split)[4]) {
  // Merge paths a(m[p]) of (r0, r1) and b(n[p]) of (r2, r3) for Q probes at
  // once, within the probe bounds of st: 2Q binary searches, interleaved.
  auto inner = [&](auto const (&m)[P], auto const (&n)[P], int q, int64_t (&a)[P],
                   int64_t (&b)[P]) {
    int64_t xb[P], xe[P], yb[P], ye[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      xb[p] = max_(max_(st.xla, st.xha - (st.xhm - m[p])), max_<int64_t>(0, m[p] - n1));
      xe[p] = min_(min_(st.xha, st.xla + (m[p] - st.xlm)), min_(m[p], n0));
      yb[p] = max_(max_(st.ylb, st.yhb - (st.yhn - n[p])), max_<int64_t>(0, n[p] - n3));
      ye[p] = min_(min_(st.yhb, st.ylb + (n[p] - st.yln)), min_(n[p], n2));
      if (p >= q) xe[p] = xb[p], ye[p] = yb[p];
    }
    bool busy = true;
    while (busy) {
      busy = false;
      key_t xa[P], xk[P], ya[P], yk[P];
      int64_t xm[P], ym[P];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        xm[p] = (xb[p] + xe[p]) >> 1;
        ym[p] = (yb[p] + ye[p]) >> 1;
        if (xb[p] < xe[p]) {
          xa[p] = r0[xm[p]];
          xk[p] = r1[m[p] - 1 - xm[p]];
        }
        if (yb[p] < ye[p]) {
          ya[p] = r2[ym[p]];
          yk[p] = r3[n[p] - 1 - ym[p]];
        }
      }
#pragma unroll
      for (int p = 0; p < P; ++p) {
        if (xb[p] < xe[p]) {
          if (!comp(xk[p], xa[p])) xb[p] = xm[p] + 1;
          else xe[p] = xm[p];
        }
        if (yb[p] < ye[p]) {
          if (!comp(yk[p], ya[p])) yb[p] = ym[p] + 1;
          else ye[p] = ym[p];
        }
        busy |= xb[p] < xe[p] || yb[p] < ye[p];
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      a[p] = xb[p];
      b[p] = yb[p];
    }
  };
  int64_t m[P], n[P], a[P], b[P];
  while (st.lo < st.hi) {
    const int64_t w = st.hi - st.lo;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      m[p] = st.lo + w * (p + 1) / (P + 1);  // lo <= m < hi, so m < nx
      n[p] = diag - 1 - m[p];                // and 0 <= n < ny
    }
    inner(m, n, P, a, b);
    bool before[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      // X[m] and Y[n]: the next output of each inner merge.
      const bool xr0 = a[p] < n0 && (m[p] - a[p] >= n1 || !comp(r1[m[p] - a[p]], r0[a[p]]));
      const bool yr2 = b[p] < n2 && (n[p] - b[p] >= n3 || !comp(r3[n[p] - b[p]], r2[b[p]]));
      const key_t xkey = xr0 ? r0[a[p]] : r1[m[p] - a[p]];
      const key_t ykey = yr2 ? r2[b[p]] : r3[n[p] - b[p]];
      before[p] = !comp(ykey, xkey);  // X[m] precedes Y[n]: more than m outputs from X
    }
    // Probes ascend in m and `before` is true then false along them: the last
    // true probe raises lo, the first false one lowers hi.
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (before[p]) {
        if (m[p] + 1 > st.lo) {
          st.lo = m[p] + 1;
          st.xlm = m[p];
          st.xla = a[p];
          st.yhn = n[p];
          st.yhb = b[p];
        }
      } else if (m[p] < st.hi) {
        st.hi = m[p];
        st.xhm = m[p];
        st.xha = a[p];
        st.yln = n[p];
        st.ylb = b[p];
      }
    }
  }
  m[0] = st.lo;
  n[0] = diag - st.lo;
  inner(m, n, 1, a, b);
  split[0] = a[0];
  split[1] = st.lo - a[0];
  split[2] = b[0];
  split[3] = diag - st.lo - b[0];
}

template <typename key_t, typename comp_t>
__host__ __device__ inline void merge4_path(const key_t* r0, int64_t n0, const key_t* r1,
                                            int64_t n1, const key_t* r2, int64_t n2,
                                            const key_t* r3, int64_t n3, int64_t diag,
                                            comp_t comp, int64_t (&split)[4]) {
  merge4_search(r0, n0, r1, n1, r2, n2, r3, n3, diag, comp, merge4_initial(n0, n1, n2, n3, diag),
                split);
}

// Smallest i in [begin, end) at which pred (true, then false, along the
// range) is false, with P probes per round; two such searches run
// interleaved in kary_search2.
template <int P = merge4_probes, typename pred_t>
__host__ __device__ inline int64_t kary_search(int64_t begin, int64_t end, pred_t pred) {
  while (begin < end) {
    const int64_t w = end - begin;
    int64_t mid[P];
    bool t[P];
#pragma unroll
    for (int p = 0; p < P; ++p) mid[p] = begin + w * (p + 1) / (P + 1);
#pragma unroll
    for (int p = 0; p < P; ++p) t[p] = pred(mid[p]);
    int64_t nb = begin, ne = end;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (t[p]) nb = max_(nb, mid[p] + 1);
      else ne = min_(ne, mid[p]);
    }
    begin = nb;
    end = ne;
  }
  return begin;
}

template <int P = merge4_probes, typename pred1_t, typename pred2_t>
__host__ __device__ inline void kary_search2(int64_t& b1, int64_t e1, pred1_t pred1, int64_t& b2,
                                             int64_t e2, pred2_t pred2) {
  while (b1 < e1 || b2 < e2) {
    const bool s1 = b1 < e1, s2 = b2 < e2;
    const int64_t w1 = e1 - b1, w2 = e2 - b2;
    int64_t m1[P], m2[P];
    bool t1[P], t2[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      m1[p] = b1 + w1 * (p + 1) / (P + 1);
      m2[p] = b2 + w2 * (p + 1) / (P + 1);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (s1) t1[p] = pred1(m1[p]);
      if (s2) t2[p] = pred2(m2[p]);
    }
    int64_t nb1 = b1, ne1 = e1, nb2 = b2, ne2 = e2;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (s1) {
        if (t1[p]) nb1 = max_(nb1, m1[p] + 1);
        else ne1 = min_(ne1, m1[p]);
      }
      if (s2) {
        if (t2[p]) nb2 = max_(nb2, m2[p] + 1);
        else ne2 = min_(ne2, m2[p]);
      }
    }
    b1 = nb1;
    e1 = ne1;
    b2 = nb2;
    e2 = ne2;
  }
}

// Per-thread merge path (binary search), used in LDS.
template <typename key_t, typename comp_t>
__device__ inline int merge_path_lds(const key_t* a, int a_count, const key_t* b,
                                     int b_count, int diag, comp_t comp) {
  return merge_path_search<bounds_lower, int>(a, a_count, b, b_count, diag, comp);
}

// Merge path in global memory, one binary search per thread (the upper levels
// of every search touch the same few cache lines, which stay in L2).
template <bounds_t bounds = bounds_lower, typename key_t, typename comp_t>
__device__ inline int64_t merge_path_global(const key_t* a, int64_t a_count, const key_t* b,
                                            int64_t b_count, int64_t diag, comp_t comp) {
  return merge_path_search<bounds, int64_t>(a, a_count, b, b_count, diag, comp);
}

// Serial merge of vt outputs starting at (a_begin, b_begin) of LDS keys, with
// the range checks of merge_predicate (cta_merge.hxx:41-50).  keys_shared
// must be readable one element past the ranges.  With checked = false the
// caller guarantees that neither range runs out within the VT steps (both
// hold at least VT keys), and the range tests are dropped.
template <int VT, bool checked = true, typename key_t, typename comp_t>
__device__ inline void serial_merge(const key_t* keys_shared, int a_begin, int a_end,
                                    int b_begin, int b_end, key_t (&keys)[VT],
                                    int (&indices)[VT], comp_t comp) {
  key_t a_key = keys_shared[a_begin];
  key_t b_key = keys_shared[b_begin];
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const bool p = checked ? (a_begin < a_end && (b_begin >= b_end || !comp(b_key, a_key)))
                           : !comp(b_key, a_key);
    const int index = p ? a_begin : b_begin;
    keys[i] = p ? a_key : b_key;
    indices[i] = index;
    const key_t c_key = keys_shared[index + 1];
    a_begin = p ? index + 1 : a_begin;
    b_begin = p ? b_begin : index + 1;
    a_key = p ? c_key : a_key;
    b_key = p ? b_key : c_key;
  }
}

// serial_merge with the range tests dropped when every lane of the wave has
// at least VT keys left in both ranges (a wave-uniform choice, so the wave
// runs one of the two loops, never both).
template <int VT, typename key_t, typename comp_t>
__device__ inline void serial_merge_wave(const key_t* keys_shared, int a_begin, int a_end,
                                         int b_begin, int b_end, key_t (&keys)[VT],
                                         int (&indices)[VT], comp_t comp) {
  const bool roomy = a_end - a_begin >= VT && b_end - b_begin >= VT;
  if (__all(roomy))
    serial_merge<VT, false>(keys_shared, a_begin, a_end, b_begin, b_end, keys, indices, comp);
  else
    serial_merge<VT, true>(keys_shared, a_begin, a_end, b_begin, b_end, keys, indices, comp);
}

// Partition kernel: mp[i] = merge_path(a, b, min(i * spacing, a_count + b_count))
// (vk/partition.hxx:9-20); mp_t is int64_t here, int for the reference's
// int partition arrays.
template <int NT, bounds_t bounds = bounds_lower, typename key_t, typename mp_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_merge_partition(
    const key_t* a, int64_t a_count, const key_t* b, int64_t b_count, int64_t spacing,
    int64_t num_partitions, mp_t* mp, comp_t comp) {
  int64_t index = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (index >= num_partitions) return;
  int64_t diag = min_<int64_t>(spacing * index, a_count + b_count);
  mp[index] = (mp_t)merge_path_global<bounds>(a, a_count, b, b_count, diag, comp);
}

// Partitions of a merge into `spacing`-output tiles: the reference's
// num_merge_partitions (common/cta_merge.hxx:6-8).
constexpr int64_t num_merge_partitions(int64_t count, int64_t spacing) {
  return (count + spacing - 1) / spacing + 1;
}

// vk::merge_path_partitions<bounds>(cmd, a, a_count, b, b_count, mp_data,
// spacing, comp) (vk/partition.hxx:9-20) with the stream in place of the
// command buffer: mp_data[i] = the merge path of diagonal
// min(i * spacing, a_count + b_count) for i < num_merge_partitions(a_count +
// b_count, spacing).  No aux.  mp_data is int (as in the reference: totals
// below 2^31) or int64_t.
template <bounds_t bounds, typename key_t, typename mp_t, typename comp_t>
mgpu_status_t merge_path_partitions(hipStream_t stream, const key_t* a, int64_t a_count,
                                    const key_t* b, int64_t b_count, mp_t* mp_data,
                                    int64_t spacing, comp_t comp) {
  static_assert(std::is_same_v<mp_t, int> || std::is_same_v<mp_t, int64_t>,
                "partitions are int or int64_t");
  if (a_count < 0 || b_count < 0 || spacing <= 0) return MGPU_INVALID_ARG;
  if (std::is_same_v<mp_t, int> && a_count + b_count >= ((int64_t)1 << 31))
    return MGPU_UNSUPPORTED;
  if (!mp_data || (a_count && !a) || (b_count && !b)) return MGPU_INVALID_ARG;
  const int64_t num_partitions = num_merge_partitions(a_count + b_count, spacing);
  constexpr int PNT = 64;  // one wave per block: the searches spread over every CU
  return hip_status(launch_kernel("mgpu_merge_partition",
                                  kernel_merge_partition<PNT, bounds, key_t, mp_t, comp_t>,
                                  dim3((unsigned)div_up(num_partitions, (int64_t)PNT)), dim3(PNT),
                                  0, stream, a, a_count, b, b_count, spacing, num_partitions,
                                  mp_data, comp));
}

// LDS footprint of a tile merge.
template <int NT, int VT, typename key_t, typename val_t>
struct merge_storage_t {
  static constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  key_t keys[NT * VT + 1];
  typename std::conditional<has_values, val_t, char>::type vals[has_values ? NT * VT + 1 : 1];
};

// Merge A[a_begin, a_begin+a_cnt) and B[b_begin, b_begin+b_cnt) (a_cnt+b_cnt <=
// NT*VT) into C[c_begin, ...), keys and values.  Called by the whole block.
template <int NT, int VT, typename key_t, typename val_t, typename comp_t>
__device__ inline void cta_merge_tile(
    const key_t* a_keys, const val_t* a_vals, int64_t a_begin, int a_cnt,
    const key_t* b_keys, const val_t* b_vals, int64_t b_begin, int b_cnt,
    key_t* c_keys, val_t* c_vals, int64_t c_begin, int c_count, comp_t comp,
    merge_storage_t<NT, VT, key_t, val_t>& s) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  const int tid = threadIdx.x;
  // a_cnt + b_cnt == c_count <= NT*VT for valid partitions; the clamps keep a
  // corrupt partition from reaching outside the tile.
  const int total = min_(min_(a_cnt + b_cnt, c_count), NT * VT);
  a_cnt = min_(a_cnt, total);
  b_cnt = total - a_cnt;

  if (total <= 0) return;

  // Load both streams into LDS (coalesced strided loads).  Each load reads a
  // clamped index from A or B chosen by select, with no branch, so that all
  // loads are in flight together (a guarded load costs one wait each).
  {
    key_t x[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int j = min_(i * NT + tid, total - 1);
      const bool in_a = j < a_cnt;
      x[i] = (in_a ? a_keys : b_keys)[in_a ? a_begin + j : b_begin + (j - a_cnt)];
    }
#pragma unroll
    for (int i = 0; i < VT; ++i) s.keys[i * NT + tid] = x[i];
  }
  if constexpr (has_values) {
    val_t x[VT];
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int j = min_(i * NT + tid, total - 1);
      const bool in_a = j < a_cnt;
      x[i] = (in_a ? a_vals : b_vals)[in_a ? a_begin + j : b_begin + (j - a_cnt)];
    }
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[i * NT + tid] = x[i];
  }
  __syncthreads();

  int diag = min_(VT * tid, total);
  int mp = merge_path_lds(s.keys, a_cnt, s.keys + a_cnt, b_cnt, diag, comp);
  key_t keys[VT];
  int indices[VT];
  serial_merge_wave<VT>(s.keys, mp, a_cnt, a_cnt + diag - mp, total, keys, indices, comp);
  val_t vals[VT];
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) vals[i] = s.vals[indices[i]];
  }
  __syncthreads();

  // Thread order -> LDS -> strided coalesced stores.
#pragma unroll
  for (int i = 0; i < VT; ++i) s.keys[VT * tid + i] = keys[i];
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) s.vals[VT * tid + i] = vals[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const int j = i * NT + tid;
    const key_t x = s.keys[j];
    if (j < total) c_keys[c_begin + j] = x;
  }
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int j = i * NT + tid;
      const val_t x = s.vals[j];
      if (j < total) c_vals[c_begin + j] = x;
    }
  }
}

template <int NT, int VT, typename key_t, typename val_t, typename comp_t>
__global__ __launch_bounds__(NT) void kernel_merge(
    const int64_t* mp, const key_t* a_keys, const val_t* a_vals, int64_t a_count,
    const key_t* b_keys, const val_t* b_vals, int64_t b_count, key_t* c_keys, val_t* c_vals,
    comp_t comp) {
  constexpr int NV = NT * VT;
  __shared__ merge_storage_t<NT, VT, key_t, val_t> s;
  const int64_t cta = blockIdx.x;
  const int64_t diag0 = cta * NV;
  const int64_t diag1 = min_<int64_t>(a_count + b_count, diag0 + NV);
  const int64_t mp0 = mp[cta], mp1 = mp[cta + 1];
  cta_merge_tile<NT, VT>(a_keys, a_vals, mp0, (int)(mp1 - mp0), b_keys, b_vals, diag0 - mp0,
                         (int)((diag1 - mp1) - (diag0 - mp0)), c_keys, c_vals, diag0,
                         (int)(diag1 - diag0), comp, s);
}

// Stable merge of (a_keys, a_vals) and (b_keys, b_vals) into (c_keys, c_vals);
// val_t = empty_t for key-only.  inc/mgpu/vk/merge.hxx:12-59.  The default
// tile is the one measured fastest for 32-bit keys on MI355X (512 x 23); VT
// is odd because the thread-order LDS transposes read word VT * tid + i,
// which a power-of-two VT puts in few banks.
template <int NT = 512, int VT = 23, typename key_t, typename val_t, typename comp_t>
mgpu_status_t merge(void* aux_data, size_t& aux_size, hipStream_t stream,
                    const key_t* a_keys, const val_t* a_vals, int64_t a_count,
                    const key_t* b_keys, const val_t* b_vals, int64_t b_count,
                    key_t* c_keys, val_t* c_vals, comp_t comp) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr int NV = NT * VT;
  if (a_count < 0 || b_count < 0) return MGPU_INVALID_ARG;
  const int64_t total = a_count + b_count;
  if (total == 0) return MGPU_OK;
  const int64_t num_tiles = div_up(total, (int64_t)NV);
  if (num_tiles >= ((int64_t)1 << 31)) return MGPU_UNSUPPORTED;
  const int64_t num_partitions = num_tiles + 1;
  size_t need = aux_bytes<int64_t>(num_partitions);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if ((a_count && !a_keys) || (b_count && !b_keys) || !c_keys) return MGPU_INVALID_ARG;
  if (has_values && ((a_count && !a_vals) || (b_count && !b_vals) || !c_vals))
    return MGPU_INVALID_ARG;
  if (mgpu_status_t st = check_aux(aux_data, aux_size, need)) return st;
  void* p = aux_data;
  int64_t* mp = advance_pointer<int64_t>(p, num_partitions);
  constexpr int PNT = 64;  // one wave per block: the searches spread over every CU
  hipError_t e = launch_kernel("mgpu_merge_partition",
                               kernel_merge_partition<PNT, bounds_lower, key_t, int64_t, comp_t>,
                               dim3((unsigned)div_up(num_partitions, (int64_t)PNT)),
                               dim3(PNT), 0, stream, a_keys, a_count, b_keys, b_count,
                               (int64_t)NV, num_partitions, mp, comp);
  if (e != hipSuccess) return hip_status(e);
  e = launch_kernel("mgpu_merge", kernel_merge<NT, VT, key_t, val_t, comp_t>,
                    dim3((unsigned)num_tiles), dim3(NT), 0, stream, (const int64_t*)mp, a_keys,
                    a_vals, a_count, b_keys, b_vals, b_count, c_keys, c_vals, comp);
  return hip_status(e);
}

// Key-only merge (inc/mgpu/vk/merge.hxx:49-59; an explicit comparator).
template <int NT = 512, int VT = 23, typename key_t, typename comp_t>
mgpu_status_t merge(void* aux_data, size_t& aux_size, hipStream_t stream,
                    const key_t* a_keys, int64_t a_count, const key_t* b_keys,
                    int64_t b_count, key_t* c_keys, comp_t comp) {
  return merge<NT, VT>(aux_data, aux_size, stream, a_keys, (const empty_t*)nullptr, a_count,
                       b_keys, (const empty_t*)nullptr, b_count, c_keys, (empty_t*)nullptr,
                       comp);
}

}  // namespace hip
}  // namespace mgpu
