// LSD radix sort with 8-bit digits for gfx950, keys or key-value.
//
// Reference: vk::radix_sort (inc/mgpu/vk/radix.hxx:195-470) runs, per digit,
// an upsweep of per-tile digit counts (:315-365), a recursive column scan of
// the [tile][256] counters (radix_scan_8, :13-185, 7 dispatches) and a
// downsweep scatter (:380-465); its 8-bit ranking is a 32-lane ballot match
// with per-warp histograms (common/cta_radix.hxx:212-351).  Keys are compared
// in the order-preserving bit form of radix_permute_t (common/cta_radix.hxx:
// 6-64).
//
// Here each digit pass is three launches with no inter-block waiting:
//   1. kernel_radix_count: per-tile digit counts, written digit-major
//      counts[d * tiles + t] (4 bytes read per key);
//   2. an exclusive scan of that array (the single-pass scan of scan.hxx), so
//      counts[d * tiles + t] becomes the global position of tile t's first key
//      of digit d: sum_{d'<d} total(d') + sum_{t'<t} count(t', d), the
//      (bin, tile) order of vk/radix.hxx:357-375;
//   3. kernel_radix_scatter: ranks every key within its tile (a wave64 match
//      of 64-bit lane masks, through an LDS table of per-digit masks for
//      32-bit keys and as the AND of 8 digit-bit ballots for 64-bit keys, plus
//      per-wave counters in LDS), reorders the tile by digit in LDS, and
//      writes keys and values in runs of consecutive addresses (16 bytes
//      moved per 32-bit key-value pair).
// A single-pass structure (kernel_radix_onesweep) is available by setting;
// on MI355X it is slower at every size (see radix_algorithm_setting).
// Tiles are remapped so that each XCD (blockIdx % 8) owns a contiguous tile
// range, which keeps the count/offset lines a tile touches in its XCD's L2.
// Digits go from least to most significant and every pass is stable, so the
// result is the stable sort by radix bits.
#pragma once
#include "meta.hxx"
#include "scan.hxx"

namespace mgpu {
namespace hip {

constexpr int radix_num_bits = 8;
constexpr int radix_num_bins = 256;
constexpr int num_xcds = 8;

// Tile processed by block b when blocks are dealt round-robin to the XCDs:
// XCD x = b % 8 gets the contiguous range [start(x), start(x) + share(x)).
__device__ inline int64_t xcd_tile(int64_t b, int64_t num_tiles) {
  const int64_t q = num_tiles / num_xcds, r = num_tiles % num_xcds;
  const int64_t x = b % num_xcds, k = b / num_xcds;
  return x * q + min_(x, r) + k;
}

// Key -> sort bits.  radix_bits_t is radix_permute_t's order (ascending);
// the segmented sort also uses descending bits and a float map that ties
// -0.0 with +0.0, so its order matches comparator semantics.
template <typename key_t>
struct radix_bits_t {
  typedef typename radix_permute_t<key_t>::unsigned_type type;
  __device__ type operator()(key_t k) const { return radix_permute_t<key_t>::to_radix_bits(k); }
};

// Counts of one digit over tile_count keys: out[d * stride] = count of d.
// CNT threads read the tile with 128-bit loads when it is full and aligned
// (order is irrelevant for counting).  16 sub-histograms (by lane mod 16)
// keep same-address LDS atomics at most 4 deep even when every lane holds the
// same digit.
template <int CNT, int NV, typename key_t, typename xbits_t>
__device__ inline void cta_radix_count(const key_t* keys, int tile_count, int shift, xbits_t xb,
                                       uint32_t* out, int64_t stride, uint32_t* h) {
  constexpr int COLS = 16;
  constexpr int KPV = 16 / (int)sizeof(key_t);   // keys per 128-bit vector
  constexpr int NVEC = NV / KPV;                 // vectors per tile
  static_assert(NV % (KPV * CNT) == 0, "tile must split into 128-bit loads");
  const int tid = threadIdx.x;
  for (int i = tid; i < radix_num_bins * COLS; i += CNT) h[i] = 0;
  __syncthreads();
  const int col = lane_id() % COLS;
  auto add = [&](key_t key) {
    uint32_t d = (uint32_t)(xb(key) >> shift) & 255u;
    atomicAdd(&h[d * COLS + col], 1u);
  };
  if (tile_count == NV && 0 == ((uintptr_t)keys & 15)) {
    typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
    u32x4 v[NVEC / CNT];
#pragma unroll
    for (int i = 0; i < NVEC / CNT; ++i)
      v[i] = __builtin_nontemporal_load((const u32x4*)keys + i * CNT + tid);
#pragma unroll
    for (int i = 0; i < NVEC / CNT; ++i) {
      key_t kk[KPV];
      __builtin_memcpy(kk, &v[i], 16);
#pragma unroll
      for (int e = 0; e < KPV; ++e) add(kk[e]);
    }
  } else {
    for (int j = tid; j < tile_count; j += CNT) add(keys[j]);
  }
  __syncthreads();
  for (int d = tid; d < radix_num_bins; d += CNT) {
    uint32_t sum = 0;
#pragma unroll
    for (int c = 0; c < COLS; ++c) sum += h[d * COLS + c];
    out[d * stride] = sum;
  }
}

// Per-tile counts of one digit, digit-major: counts[d * num_tiles + tile].
// The grid also zeroes clear[0, clear_words) (the status words of the scan
// that follows), which saves that scan its own clearing launch.
template <int CNT, int NV, typename key_t>
__global__ __launch_bounds__(CNT) void kernel_radix_count(const key_t* keys, int64_t count,
                                                          int shift, uint32_t* counts,
                                                          int64_t num_tiles, uint32_t* clear,
                                                          int64_t clear_words) {
  __shared__ uint32_t h[radix_num_bins * 16];
  for (int64_t i = (int64_t)blockIdx.x * CNT + threadIdx.x; i < clear_words;
       i += (int64_t)gridDim.x * CNT)
    clear[i] = 0;
  const int64_t tile = xcd_tile(blockIdx.x, num_tiles);
  const int64_t base = tile * NV;
  const int tile_count = (int)min_<int64_t>(NV, count - base);
  cta_radix_count<CNT, NV>(keys + base, tile_count, shift, radix_bits_t<key_t>(),
                           counts + tile, num_tiles, h);
}

// Per-tile counts of digit bytes written by the previous pass's scatter
// (digits[i] = digit of the key at position i), digit-major like
// kernel_radix_count; also zeroes clear[0, clear_words).
struct radix_byte_bits_t {
  __device__ uint32_t operator()(uint8_t b) const { return b; }
};
template <int NV>
__global__ __launch_bounds__(128) void kernel_radix_count_digits(const uint8_t* digits,
                                                                 int64_t count, uint32_t* counts,
                                                                 int64_t num_tiles,
                                                                 uint32_t* clear,
                                                                 int64_t clear_words) {
  __shared__ uint32_t h[radix_num_bins * 16];
  for (int64_t i = (int64_t)blockIdx.x * 128 + threadIdx.x; i < clear_words;
       i += (int64_t)gridDim.x * 128)
    clear[i] = 0;
  const int64_t tile = xcd_tile(blockIdx.x, num_tiles);
  const int64_t base = tile * NV;
  const int tile_count = (int)min_<int64_t>(NV, count - base);
  cta_radix_count<128, NV>(digits + base, tile_count, 0, radix_byte_bits_t(), counts + tile,
                           num_tiles, h);
}

template <int NT, int VT, typename key_t>
struct radix_scatter_storage_t {
  static constexpr int NW = NT / warp_size;
  union {
    key_t keys[NT * VT];                   // tile reordered by digit (then values)
    uint64_t match[NW][radix_num_bins];    // per-wave lane masks of each digit (ranking)
  };
  uint32_t hist[NW][radix_num_bins];       // per-wave counts -> tile positions of wave runs
  uint32_t hist_sink[NW][warp_size];       // write target of lanes that update nothing
  uint64_t match_sink[NW][warp_size];
  uint32_t global_off[radix_num_bins];     // output position of local slot 0 per digit
  uint32_t scan_tot[NW];
};

// Stable scatter of one tile by one digit.  Thread d < 256 passes `goff`,
// the output position of the tile's first key with digit d (relative to
// out_base); positions are checked against out_limit.
//
// `full` (tile_count == NT * VT) drops every clamp and guard.  Otherwise one
// branch-free path serves partial tiles too: the lanes past the end load the
// last key but take digit 255.  They are the
// last keys of the tile in input order, so the stable ranking places them
// after every real key of digit 255, i.e. exactly at tile positions
// [tile_count, NT * VT), and their stores are simply predicated off.
// (Divergent branches between an LDS read and its use would make the
// compiler wait for every outstanding LDS operation.)
// Where a tile's digit runs go: publish(d, n) is called by thread d < 256 as
// soon as the tile's count n of digit d is known; resolve(d, n) later returns
// the output position of the tile's first key with digit d.  With offsets
// from a scan of per-tile counts both are trivial; a single-pass sort
// publishes its counts and looks back over its predecessors in resolve.
struct radix_fixed_offsets_t {
  uint32_t goff;
  __device__ void publish(int, uint32_t) const {}
  __device__ uint32_t resolve(int, uint32_t) const { return goff; }
};

// Loads and ranks one tile: k[i] gets the keys, r[i] = digit << 16 | the
// key's rank among the same-digit keys of its wave, and s.hist[w][d] the
// tile position of wave w's first key of digit d.  Thread d < 256 gets the
// tile's count of digit d (total) and the tile position of its first key
// (off_local).  Ends with a block barrier.
template <int NT, int VT, typename key_t, bool full, typename xbits_t, typename offs_t>
__device__ inline void cta_radix_rank(const key_t* keys_in, int tile_count, int shift, xbits_t xb,
                                      offs_t& offs, radix_scatter_storage_t<NT, VT, key_t>& s,
                                      key_t (&k)[VT], uint32_t (&r)[VT], uint32_t& total,
                                      uint32_t& off_local) {
  constexpr int NW = NT / warp_size;
  static_assert(NT >= radix_num_bins && NT % radix_num_bins == 0, "NT must cover the bins");
  static_assert(sizeof(key_t) * NT * VT >= sizeof(uint64_t) * NW * radix_num_bins,
                "match table must fit in the key buffer");
  static_assert(NT * VT <= 65536, "ranks are packed in 16 bits");
  constexpr bool ballot_rank = sizeof(key_t) == 8;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int warp = tid / warp_size;
  const uint32_t last = (uint32_t)tile_count - 1;

  // Each wave owns a contiguous chunk of 64*VT keys; slot i holds keys
  // [chunk + 64 i, chunk + 64 i + 64) in lane order, so wave-local rank order
  // (slot, lane) is the input order.
  const uint32_t chunk = (uint32_t)(warp * warp_size * VT) + (uint32_t)lane;
#pragma unroll
  for (int i = 0; i < VT; ++i)
    k[i] = full ? (keys_in + chunk)[i * warp_size]  // constant offsets from one address
                : keys_in[min_(chunk + (uint32_t)(i * warp_size), last)];
  auto digit = [&](int i) -> uint32_t {
    return full || chunk + (uint32_t)(i * warp_size) <= last
               ? ((uint32_t)(xb(k[i]) >> shift) & 255u)
               : 255u;
  };
  // Wave-private counters (and match table) in LDS: the operations of one
  // wave are performed in order, so no block barrier is needed before use.
#pragma unroll
  for (int j = 0; j < radix_num_bins / warp_size; ++j) {
    s.hist[warp][lane + j * warp_size] = 0;
    if constexpr (!ballot_rank) s.match[warp][lane + j * warp_size] = 0;
  }

  // Wave-local stable rank: rank = the wave's running count of the digit
  // (per-wave counters in LDS) + the number of lower lanes with the same
  // digit in this slot.  The lowest lane of each digit advances the count;
  // the other lanes write the same value to a private sink slot, so the write
  // needs no branch.  The mask of lanes with the same digit comes from
  //  * 32-bit keys: the lanes sharing the first lane's digit from one ballot;
  //    every other lane ORs its bit into the table entry of its digit, reads
  //    the entry back, and the lowest lane clears it;
  //  * 64-bit keys: the AND over the digit's 8 bits of the ballot of that bit
  //    (or of its complement), in registers: these tiles stage 8-byte keys
  //    through LDS, so LDS is the tighter resource.
  // (Each was measured faster than the other for its key size on MI355X.)
  const uint64_t lane_bit = 1ull << lane;
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const uint32_t d = digit(i);
    // The counter read goes first so that it is in flight together with the
    // table read or the ballots (one LDS round trip per slot).
    const uint32_t c = s.hist[warp][d];
    uint64_t m;
    bool clear = false;
    if constexpr (ballot_rank) {
      m = ~0ull;
#pragma unroll
      for (int b = 0; b < radix_num_bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t ones = __ballot(bit);
        m &= bit ? ones : ~ones;
      }
    } else {
      const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
      m = __ballot(d == d0);
      clear = d != d0;
      if (clear) {
        atomicOr((unsigned long long*)&s.match[warp][d], (unsigned long long)lane_bit);
        m = s.match[warp][d];
      }
    }
    const uint32_t below = mask_count_below(m);
    r[i] = (d << 16) | (c + below);  // the digit rides along for the reorder
    const bool lead = below == 0;
    uint32_t* hp = lead ? &s.hist[warp][d] : &s.hist_sink[warp][lane];
    *hp = c + (uint32_t)__popcll(m);
    if constexpr (!ballot_rank) {
      uint64_t* mp = lead && clear ? &s.match[warp][d] : &s.match_sink[warp][lane];
      *mp = 0ull;
    }
  }
  __syncthreads();

  // Per digit: exclusive prefix over waves, tile total, tile-local offsets.
  total = 0;
  if (tid < radix_num_bins) {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      uint32_t c = s.hist[w][tid];
      s.hist[w][tid] = total;
      total += c;
    }
  }
  uint32_t incl = 0;
  if (tid < radix_num_bins) {
    offs.publish(tid, total);
    incl = wave_scan_incl_u32(total);
    if (lane == warp_size - 1) s.scan_tot[warp] = incl;
  }
  __syncthreads();
  off_local = 0;
  if (tid < radix_num_bins) {
    off_local = incl - total;
    for (int w = 0; w < warp; ++w) off_local += s.scan_tot[w];
    // hist[w][d] becomes the tile position of wave w's first key of digit d.
#pragma unroll
    for (int w = 0; w < NW; ++w) s.hist[w][tid] += off_local;
  }
  __syncthreads();
}

template <int NT, int VT, typename key_t, bool has_values, bool full = false, bool emit = false,
          typename xbits_t, typename offs_t>
__device__ inline void cta_radix_scatter(const key_t* keys_in, const uint32_t* vals_in,
                                         key_t* keys_out, uint32_t* vals_out, int tile_count,
                                         int shift, xbits_t xb, offs_t& offs, int64_t out_limit,
                                         radix_scatter_storage_t<NT, VT, key_t>& s,
                                         uint8_t* digits_out = nullptr, int next_shift = 0) {
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int warp = tid / warp_size;
  const uint32_t last = (uint32_t)tile_count - 1;
  const uint32_t chunk = (uint32_t)(warp * warp_size * VT) + (uint32_t)lane;
  key_t k[VT];
  uint32_t r[VT];
  uint32_t total, off_local;
  cta_radix_rank<NT, VT, key_t, full>(keys_in, tile_count, shift, xb, offs, s, k, r, total,
                                      off_local);

  // Values are loaded only now: their latency overlaps the key reorder and
  // stores, and they are not live during the ranking.
  // The index base goes through an empty asm so that the compiler does not
  // keep the keys' load offsets alive (in registers) to reuse them here.
  uint32_t v[VT];
  if constexpr (has_values) {
    uint32_t vchunk = chunk;
    asm volatile("" : "+v"(vchunk));
#pragma unroll
    for (int i = 0; i < VT; ++i)
      v[i] = full ? (vals_in + vchunk)[i * warp_size]
                  : vals_in[min_(vchunk + (uint32_t)(i * warp_size), last)];
  }

  // Reorder the tile by digit in LDS.
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const uint32_t d = r[i] >> 16;
    const uint32_t pos = s.hist[warp][d] + (r[i] & 0xFFFFu);
    s.keys[pos] = k[i];
    r[i] = pos;
  }
  if (tid < radix_num_bins) s.global_off[tid] = offs.resolve(tid, total) - off_local;
  __syncthreads();

  uint32_t dst[VT];
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const uint32_t p = (uint32_t)(i * NT + tid);
    const key_t key = s.keys[p];
    const uint32_t d = (uint32_t)(xb(key) >> shift) & 255u;
    dst[i] = s.global_off[d] + p;
    if ((full || p <= last) && dst[i] < out_limit) {
      keys_out[dst[i]] = key;
      if constexpr (emit) digits_out[dst[i]] = (uint8_t)((uint32_t)(xb(key) >> next_shift) & 255u);
    }
  }
  if constexpr (has_values) {
    __syncthreads();
    uint32_t* s_vals = (uint32_t*)s.keys;
#pragma unroll
    for (int i = 0; i < VT; ++i) s_vals[r[i]] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const uint32_t p = (uint32_t)(i * NT + tid);
      if ((full || p <= last) && dst[i] < out_limit) vals_out[dst[i]] = s_vals[p];
    }
  }
}

template <int NT, int VT, typename key_t, bool has_values, bool full = false, bool emit = false,
          typename xbits_t>
__device__ inline void cta_radix_scatter(const key_t* keys_in, const uint32_t* vals_in,
                                         key_t* keys_out, uint32_t* vals_out, int tile_count,
                                         int shift, xbits_t xb, uint32_t goff, int64_t out_limit,
                                         radix_scatter_storage_t<NT, VT, key_t>& s,
                                         uint8_t* digits_out = nullptr, int next_shift = 0) {
  radix_fixed_offsets_t offs{goff};
  cta_radix_scatter<NT, VT, key_t, has_values, full, emit>(keys_in, vals_in, keys_out, vals_out,
                                                           tile_count, shift, xb, offs,
                                                           out_limit, s, digits_out, next_shift);
}

// Scatter of one digit pass.  offsets[d * num_tiles + t] is the global
// position of tile t's first key with digit d.  Blocks [0, num_full) take
// the full tiles (dealt XCD-contiguously); a block num_full, if launched,
// takes the trailing partial tile.  With emit, each key's digit of the next
// pass is also written as one byte at its output position, so that the next
// pass counts 1 byte per key instead of reading the keys.
// Two blocks per CU (the LDS allows two): at most 128 VGPRs.
template <int NT, int VT, typename key_t, bool has_values, bool emit>
__global__ __launch_bounds__(NT, 2 * NT / 256) void kernel_radix_scatter(
    const key_t* keys_in, key_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
    int64_t count, int shift, const uint32_t* offsets, int64_t num_tiles, int64_t num_full,
    uint8_t* digits_out) {
  constexpr int NV = NT * VT;
  __shared__ radix_scatter_storage_t<NT, VT, key_t> s;
  const int tid = threadIdx.x;
  const bool full = blockIdx.x < num_full;
  const int64_t tile = full ? xcd_tile(blockIdx.x, num_full) : num_full;
  const int64_t base = tile * NV;
  const int tile_count = (int)min_<int64_t>(NV, count - base);
  uint32_t goff = 0;
  if (tid < radix_num_bins) goff = offsets[tid * num_tiles + tile];
  const key_t* kin = keys_in + base;
  const uint32_t* vin = has_values ? vals_in + base : vals_in;
  if (full)
    cta_radix_scatter<NT, VT, key_t, has_values, true, emit>(
        kin, vin, keys_out, vals_out, tile_count, shift, radix_bits_t<key_t>(), goff, count, s,
        digits_out, shift + radix_num_bits);
  else
    cta_radix_scatter<NT, VT, key_t, has_values, false, emit>(
        kin, vin, keys_out, vals_out, tile_count, shift, radix_bits_t<key_t>(), goff, count, s,
        digits_out, shift + radix_num_bits);
}

////////////////////////////////////////////////////////////////////////////////
// Single-pass digit passes (Adinets & Merrill's onesweep structure): one read
// of the keys counts every digit of every pass up front; then each pass is a
// single kernel in which a tile publishes its per-digit counts right after
// ranking and looks back over its predecessors for the exclusive prefix of
// each digit, so there is no count kernel and no scan per pass.
// Status words [tile][digit]: 2-bit flag (aggregate / inclusive) + 30-bit
// count, hence count < 2^30.  Each digit thread reads 16 predecessor words
// per round (coalesced across digits), spins on an unpublished word at most
// lb.spin_limit times and then counts that predecessor's digit from its
// input, which a ping-pong pass never overwrites.
// Used only when selected (radix_algorithm_setting() == 2): on MI355X the
// look-back over status words exchanged across XCDs costs more per pass than
// the count kernel and scan it removes, at every size measured.

constexpr uint32_t radix_word_agg = 1u << 30, radix_word_inc = 2u << 30;
constexpr uint32_t radix_word_value = (1u << 30) - 1;

template <int NV, typename key_t, typename xbits_t>
struct radix_lookback_offsets_t {
  const uint32_t* bases;  // output position of the first key of each digit
  uint32_t* status;       // [tile][digit] words of this pass
  int64_t tile;
  const key_t* keys_in;   // the pass input (for the fallback)
  int shift;
  xbits_t xb;
  lookback_config_t lb;

  __device__ void publish(int d, uint32_t n) const {
    atomic_store_agent(status + tile * radix_num_bins + d,
                       (tile == 0 ? radix_word_inc : radix_word_agg) | n);
  }
  __device__ uint32_t resolve(int d, uint32_t n) const {
    uint32_t excl = 0;
    if (tile > 0) {
      excl = lookback(d);
      atomic_store_agent(status + tile * radix_num_bins + d, radix_word_inc | (excl + n));
    }
    return bases[d] + excl;
  }
  __device__ uint32_t lookback(int d) const {
    constexpr int K = 16;
    uint32_t excl = 0, spins = 0;
    int64_t j = tile - 1;
    while (true) {
      uint32_t w[K];
#pragma unroll
      for (int k = 0; k < K; ++k)
        w[k] = j - k >= 0 ? atomic_load_agent(status + (j - k) * radix_num_bins + d)
                          : radix_word_inc;  // before tile 0: an empty inclusive prefix
      int k = 0;
      bool done = false;
      for (; k < K; ++k) {
        const uint32_t f = w[k] & ~radix_word_value;
        if (f == 0) break;
        excl += w[k] & radix_word_value;
        if (f == radix_word_inc) {
          done = true;
          break;
        }
      }
      if (done) return excl;
      if (k == K) {
        j -= K;
        spins = 0;
        continue;
      }
      j -= k;  // the nearest predecessor whose word is not published yet
      if (++spins <= lb.spin_limit) {
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
      const key_t* t = keys_in + j * NV;  // predecessors are full tiles
      uint32_t c = 0;
      for (int e = 0; e < NV; ++e) c += ((uint32_t)(xb(t[e]) >> shift) & 255u) == (uint32_t)d;
      excl += c;
      if (lb.fallbacks) atomic_add_agent(lb.fallbacks, 1u);
      --j;
      spins = 0;
    }
  }
};

// counts[p * 256 + d] += number of keys whose digit p is d, for every pass at
// once (a pass permutes the keys, so these counts hold for every pass input).
// Sub-histograms by lane (16 for 32-bit keys, 8 for 64-bit) keep same-address
// LDS atomics shallow when digits repeat.
template <int NT, typename key_t>
__global__ __launch_bounds__(NT) void kernel_radix_histogram(const key_t* keys, int64_t count,
                                                             uint32_t* counts) {
  constexpr int P = (int)sizeof(key_t);
  constexpr int COLS = P == 4 ? 16 : 8;
  constexpr int KPV = 16 / P;  // keys per 128-bit vector
  constexpr int U = 8;         // vectors in flight per thread
  __shared__ uint32_t h[P * radix_num_bins * COLS];
  const int tid = threadIdx.x;
  for (int i = tid; i < P * radix_num_bins * COLS; i += NT) h[i] = 0;
  __syncthreads();
  const int64_t g = blockIdx.x, G = gridDim.x;
  const int64_t begin = 0, end = count;
  const int col = lane_id() % COLS;
  const radix_bits_t<key_t> xb;
  auto add = [&](key_t k) {
    const auto b = xb(k);
#pragma unroll
    for (int p = 0; p < P; ++p)
      atomicAdd(&h[(p * radix_num_bins + (int)((b >> (8 * p)) & 255u)) * COLS + col], 1u);
  };
  int64_t tail = begin;
  if (end > begin && 0 == ((uintptr_t)(keys + begin) & 15)) {
    typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
    const u32x4* v = (const u32x4*)(keys + begin);
    const int64_t nvec = (end - begin) / KPV;
    for (int64_t c = g * NT * U; c < nvec; c += G * NT * U) {
      u32x4 r[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        r[u] = __builtin_nontemporal_load(v + min_<int64_t>(c + u * NT + tid, nvec - 1));
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (c + u * NT + tid < nvec) {
          key_t kk[KPV];
          __builtin_memcpy(kk, &r[u], 16);
#pragma unroll
          for (int e = 0; e < KPV; ++e) add(kk[e]);
        }
      }
    }
    tail = begin + nvec * KPV;
  }
  for (int64_t i = tail + g * NT + tid; i < end; i += G * NT) add(keys[i]);
  __syncthreads();
  for (int i = tid; i < P * radix_num_bins; i += NT) {
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < COLS; ++k) c += h[i * COLS + k];
    if (c) atomic_add_agent(counts + i, c);
  }
}

// bases[p * 256 + d] = exclusive prefix of counts[p * 256 + 0..d): the output
// position of the first key of digit d in pass p.  One block per pass.
template <int P>
__global__ __launch_bounds__(radix_num_bins) void kernel_radix_bases(const uint32_t* counts,
                                                                     uint32_t* bases) {
  __shared__ uint32_t wave_tot[radix_num_bins / warp_size];
  const int p = blockIdx.x, d = threadIdx.x, warp = d / warp_size;
  const uint32_t c = counts[p * radix_num_bins + d];
  const uint32_t incl = wave_scan_incl_u32(c);
  if (lane_id() == warp_size - 1) wave_tot[warp] = incl;
  __syncthreads();
  uint32_t off = incl - c;
  for (int w = 0; w < warp; ++w) off += wave_tot[w];
  bases[p * radix_num_bins + d] = off;
}

// One single-pass digit pass: tile = blockIdx.x (the look-back waits on
// lower tiles, which are dispatched first, so there is no XCD remapping).
template <int NT, int VT, typename key_t, bool has_values>
__global__ __launch_bounds__(NT, 2 * NT / 256) void kernel_radix_onesweep(
    const key_t* keys_in, key_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
    int64_t count, int shift, const uint32_t* bases, uint32_t* status, lookback_config_t lb) {
  constexpr int NV = NT * VT;
  __shared__ radix_scatter_storage_t<NT, VT, key_t> s;
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * NV;
  const int tile_count = (int)min_<int64_t>(NV, count - base);
  radix_lookback_offsets_t<NV, key_t, radix_bits_t<key_t>> offs{
      bases, status, tile, keys_in, shift, radix_bits_t<key_t>(), lb};
  const key_t* kin = keys_in + base;
  const uint32_t* vin = has_values ? vals_in + base : vals_in;
  if (tile_count == NV)
    cta_radix_scatter<NT, VT, key_t, has_values, true>(kin, vin, keys_out, vals_out, tile_count,
                                                       shift, radix_bits_t<key_t>(), offs, count,
                                                       s);
  else
    cta_radix_scatter<NT, VT, key_t, has_values, false>(kin, vin, keys_out, vals_out,
                                                        tile_count, shift, radix_bits_t<key_t>(),
                                                        offs, count, s);
}

// Which digit-pass structure radix_sort uses (a process-wide setting):
// 0 = automatic, 1 = count + scan + scatter, 2 = single pass where eligible
// (count < 2^30).  Automatic is count + scan + scatter at every size: on
// MI355X the single pass's look-back costs 1.42 ms per pass at 2^28 uint32
// pairs against 0.88 + 0.19 ms for scatter + count, and small sorts are
// faster with count + scan + scatter over small tiles (1M keys, 512 x 8
// tiles: 0.068 ms) than with the single pass (0.077 ms at 512 x 8, 0.079 ms
// at 512 x 20).
inline std::atomic<int>& radix_algorithm_setting() {
  static std::atomic<int> a{0};
  return a;
}

// Whether the count + scan + scatter passes carry each key's next digit as a
// byte (written by the scatter, read by the next count): for 64-bit keys the
// count then reads 1 byte per key instead of 8 (2^27 u64 pairs: 7.13 ->
// 6.78 ms).  For 32-bit keys the scattered byte stores cost the scatter more
// than the count saves (2^28 u32 pairs: 4.35 -> 4.71 ms).
template <typename key_t>
constexpr bool radix_digit_bytes() { return sizeof(key_t) == 8; }

template <int NT, int VT, typename key_t, typename val_t>
inline size_t radix_aux_bytes(int64_t count) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  const int64_t num_tiles = div_up(count, (int64_t)(NT * VT));
  const int64_t num_counts = num_tiles * radix_num_bins;
  size_t need = aux_bytes<key_t>(count);
  if (has_values) need += aux_bytes<uint32_t>(count);
  need += aux_bytes<uint32_t>(num_counts);
  size_t scan_need = 0;
  scan<256, 16>(nullptr, scan_need, nullptr, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                num_counts, MGPU_SCAN_EXCLUSIVE, 0u);
  // The single-pass path instead needs [passes][tiles][256] status words and
  // the all-pass counts and bases; size for both so that the choice can be
  // made when the sort is enqueued.
  const size_t scan_path = aux_bytes<uint32_t>(num_counts) + scan_need +
                           (radix_digit_bytes<key_t>() ? aux_bytes<uint8_t>(count) : 0);
  const size_t one_path = aux_bytes<uint32_t>(num_counts * (int64_t)sizeof(key_t)) +
                          2 * aux_bytes<uint32_t>((int64_t)sizeof(key_t) * radix_num_bins);
  return need - aux_bytes<uint32_t>(num_counts) + max_(scan_path, one_path);
}

// Stable LSD radix sort in place; val_t = uint32_t or empty_t.  NT x VT is the
// scatter's block shape (tile = NT * VT keys); CNT threads count a tile.
template <int NT, int VT, int CNT, typename key_t, typename val_t>
mgpu_status_t radix_sort_impl(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                              val_t* vals, int64_t count) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  static_assert(!has_values || std::is_same_v<val_t, uint32_t>, "values are uint32");
  constexpr int D = (int)sizeof(key_t);
  constexpr int NV = NT * VT;
  if (count < 0) return MGPU_INVALID_ARG;
  if (count == 0) return MGPU_OK;
  if (count >= ((int64_t)1 << 32) - NV) return MGPU_UNSUPPORTED;
  const int64_t num_tiles = div_up(count, (int64_t)NV);
  const int64_t num_counts = num_tiles * radix_num_bins;
  const size_t need = radix_aux_bytes<NT, VT, key_t, val_t>(count);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if (!keys || (has_values && !vals)) return MGPU_INVALID_ARG;
  if (mgpu_status_t st = check_aux(aux_data, aux_size, need)) return st;
  void* p = aux_data;
  key_t* keys2 = advance_pointer<key_t>(p, count);
  uint32_t* vals2 = has_values ? advance_pointer<uint32_t>(p, count) : nullptr;

  key_t* k_in = keys;
  key_t* k_out = keys2;
  uint32_t* v_in = has_values ? (uint32_t*)vals : nullptr;
  uint32_t* v_out = vals2;
  hipError_t e;
  const int algo = radix_algorithm_setting().load(std::memory_order_relaxed);
  const bool single_pass = algo == 2 && count < ((int64_t)1 << 30);
  if (single_pass) {
    uint32_t* status = advance_pointer<uint32_t>(p, num_counts * D);
    uint32_t* all_counts = advance_pointer<uint32_t>(p, (int64_t)D * radix_num_bins);
    uint32_t* bases = advance_pointer<uint32_t>(p, (int64_t)D * radix_num_bins);
    const lookback_config_t lb = lookback_config(stream);
    e = fill_u32(status, num_counts * D, 0u, stream);
    if (e == hipSuccess) e = fill_u32(all_counts, (int64_t)D * radix_num_bins, 0u, stream);
    // At least 8 vectors per thread; at most 4096 blocks.
    const int64_t hist_blocks = min_<int64_t>(4096, div_up(count, (int64_t)256 * 8 * (16 / D)));
    if (e == hipSuccess)
      e = launch_kernel("mgpu_radix_histogram", kernel_radix_histogram<256, key_t>,
                        dim3((unsigned)hist_blocks), dim3(256), 0, stream, (const key_t*)keys,
                        count, all_counts);
    if (e == hipSuccess)
      e = launch_kernel("mgpu_radix_bases", kernel_radix_bases<D>, dim3(D), dim3(radix_num_bins), 0,
                        stream, (const uint32_t*)all_counts, bases);
    for (int pass = 0; e == hipSuccess && pass < D; ++pass) {
      e = launch_kernel("mgpu_radix_onesweep", kernel_radix_onesweep<NT, VT, key_t, has_values>,
                        dim3((unsigned)num_tiles), dim3(NT), 0, stream, (const key_t*)k_in, k_out,
                        (const uint32_t*)v_in, v_out, count, 8 * pass,
                        (const uint32_t*)bases + pass * radix_num_bins,
                        status + (int64_t)pass * num_counts, lb);
      std::swap(k_in, k_out);
      std::swap(v_in, v_out);
    }
    return hip_status(e);
  }
  constexpr bool emit = radix_digit_bytes<key_t>();
  uint8_t* digits = emit ? advance_pointer<uint8_t>(p, count) : nullptr;
  uint32_t* counts = advance_pointer<uint32_t>(p, num_counts);
  // The counts scan's status words (one 64-bit word per 4096 counts, at the
  // start of its aux) are zeroed by the count kernel of each pass.
  uint64_t* scan_status = (uint64_t*)p;
  const int64_t scan_words = 2 * div_up(num_counts, (int64_t)(256 * 16));
  const int64_t num_full = count / NV;
  for (int pass = 0; pass < D; ++pass) {
    const int shift = 8 * pass;
    // Pass 0 counts digits from the keys; with digit bytes, later passes
    // count the bytes the previous scatter wrote.
    if (pass == 0 || !emit)
      e = launch_kernel("mgpu_radix_count", kernel_radix_count<CNT, NV, key_t>,
                        dim3((unsigned)num_tiles), dim3(CNT), 0, stream, (const key_t*)k_in,
                        count, shift, counts, num_tiles, (uint32_t*)scan_status, scan_words);
    else
      e = launch_kernel("mgpu_radix_count_digits", kernel_radix_count_digits<NV>,
                        dim3((unsigned)num_tiles), dim3(128), 0, stream, (const uint8_t*)digits,
                        count, counts, num_tiles, (uint32_t*)scan_status, scan_words);
    if (e != hipSuccess) return hip_status(e);
    e = scan_launch_cleared<256, 16>(stream, (const uint32_t*)counts, counts, num_counts,
                                     MGPU_SCAN_EXCLUSIVE, 0u, identity_t(), scan_status);
    if (e != hipSuccess) return hip_status(e);
    if (emit && pass + 1 < D)
      e = launch_kernel("mgpu_radix_scatter", kernel_radix_scatter<NT, VT, key_t, has_values, emit>,
                        dim3((unsigned)num_tiles), dim3(NT), 0, stream, (const key_t*)k_in, k_out,
                        (const uint32_t*)v_in, v_out, count, shift, (const uint32_t*)counts,
                        num_tiles, num_full, digits);
    else
      e = launch_kernel("mgpu_radix_scatter", kernel_radix_scatter<NT, VT, key_t, has_values, false>,
                        dim3((unsigned)num_tiles), dim3(NT), 0, stream, (const key_t*)k_in, k_out,
                        (const uint32_t*)v_in, v_out, count, shift, (const uint32_t*)counts,
                        num_tiles, num_full, (uint8_t*)nullptr);
    if (e != hipSuccess) return hip_status(e);
    std::swap(k_in, k_out);
    std::swap(v_in, v_out);
  }
  return MGPU_OK;
}

// A tile shape this scatter and count can run: NT covers the 256 digits; the
// tile splits into the count's 128-bit loads (and, for 64-bit keys, into the
// digit-byte count's 128 x 16 bytes); the ranking's match table fits in the
// key buffer; ranks fit 16 bits; the scatter's LDS fits 160 KB.
template <int NT, int VT, int CNT, typename key_t>
constexpr bool radix_shape_ok() {
  constexpr int64_t NV = (int64_t)NT * VT;
  return NT >= radix_num_bins && NT % radix_num_bins == 0 && NT <= 1024 && VT >= 1 &&
         CNT >= 64 && CNT <= 1024 && NV % ((16 / (int)sizeof(key_t)) * CNT) == 0 &&
         (sizeof(key_t) != 8 || NV % 2048 == 0) &&
         (int64_t)sizeof(key_t) * NV >= 8 * (NT / warp_size) * radix_num_bins &&
         NV <= 65536 && sizeof(radix_scatter_storage_t<NT, VT, key_t>) <= 160 * 1024;
}

// The reference's template form radix_sort<nt, vt, num_bits> (vk/radix.hxx:
// 195-197).  num_bits (4 or 8) is kept for call compatibility: digits here are
// always 8 bits wide, which yields the same stable order.  nt and vt only tune
// (the result does not depend on them), so a shape this implementation cannot
// run, such as the reference's default 128 x 15, runs as 512 x 24.
template <int NT = 512, int VT = 24, int num_bits = 8, int CNT = 512, typename key_t,
          typename val_t>
mgpu_status_t radix_sort(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                         val_t* vals, int64_t count) {
  static_assert(num_bits == 4 || num_bits == 8, "num_bits is 4 or 8, as in the reference");
  if constexpr (radix_shape_ok<NT, VT, CNT, key_t>())
    return radix_sort_impl<NT, VT, CNT>(aux_data, aux_size, stream, keys, vals, count);
  else
    return radix_sort_impl<512, 24, 512>(aux_data, aux_size, stream, keys, vals, count);
}

template <int NT = 512, int VT = 24, int num_bits = 8, int CNT = 512, typename key_t>
mgpu_status_t radix_sort(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                         int64_t count) {
  return radix_sort<NT, VT, num_bits, CNT>(aux_data, aux_size, stream, keys, (empty_t*)nullptr,
                                           count);
}

}  // namespace hip
}  // namespace mgpu
