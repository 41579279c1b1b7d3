// LSD radix sort with 8-bit digits for gfx950, keys or key-value.
//
// Reference: vk::radix_sort (inc/mgpu/vk/radix.hxx:195-470) runs, per digit,
// an upsweep of per-tile digit counts, a recursive column scan of the
// [tile][256] counters (radix_scan_8, :13-185, 7 dispatches) and a downsweep
// scatter; its 8-bit ranking is a 32-lane ballot match with per-warp
// histograms (common/cta_radix.hxx:212-351).  Keys are compared in the
// order-preserving bit form of radix_permute_t (common/cta_radix.hxx:6-64).
//
// Here:
//   * one histogram kernel reads the keys once and counts every digit;
//   * each digit pass is one kernel ("onesweep", Adinets & Merrill 2022): a
//     tile ranks its keys with a wave64 ballot match (8 ballots, 64-bit masks)
//     and per-wave LDS counters, publishes its per-digit counts, and looks back
//     over its predecessors' counts (decoupled look-back per digit) to find
//     each digit's global offset; keys (and values) are reordered by digit in
//     LDS and written with runs of consecutive addresses.
// Digits go from least to most significant, and every pass is stable, so the
// result is the stable sort by radix bits.  The offset of digit d in tile t is
// sum_{d'<d} total(d') + sum_{t'<t} count(t', d), the (bin, tile) order of
// vk/radix.hxx:357-375.
#pragma once
#include "meta.hxx"

namespace mgpu {
namespace hip {

constexpr int radix_num_bits = 8;
constexpr int radix_num_bins = 256;

// Look-back word: 2-bit flag in bits 30..31 and a 30-bit count.
constexpr uint32_t radix_flag_aggregate = 1u << 30;
constexpr uint32_t radix_flag_inclusive = 2u << 30;
constexpr uint32_t radix_value_mask = (1u << 30) - 1;

// Counts of every digit of every key: hist[digit_index][bin].  Sub-histograms
// per lane group keep same-address LDS atomics at most 64/COLS deep.
template <int NT, typename key_t>
__global__ __launch_bounds__(NT) void kernel_radix_histogram(const key_t* keys, int64_t count,
                                                             uint32_t* hist) {
  typedef typename radix_permute_t<key_t>::unsigned_type bits_t;
  constexpr int D = (int)sizeof(key_t);
  constexpr int COLS = 64 / D;
  __shared__ uint32_t h[D * radix_num_bins * COLS];
  const int tid = threadIdx.x;
  for (int i = tid; i < D * radix_num_bins * COLS; i += NT) h[i] = 0;
  __syncthreads();
  const int col = lane_id() % COLS;
  const int64_t stride = (int64_t)gridDim.x * NT;
  int64_t i = (int64_t)blockIdx.x * NT + tid;
  auto count_key = [&](key_t key) {
    bits_t b = radix_permute_t<key_t>::to_radix_bits(key);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      uint32_t digit = (uint32_t)(b >> (8 * d)) & 255u;
      atomicAdd(&h[(d * radix_num_bins + digit) * COLS + col], 1u);
    }
  };
  for (; i + 3 * stride < count; i += 4 * stride) {
    key_t k0 = keys[i], k1 = keys[i + stride], k2 = keys[i + 2 * stride], k3 = keys[i + 3 * stride];
    count_key(k0); count_key(k1); count_key(k2); count_key(k3);
  }
  for (; i < count; i += stride) count_key(keys[i]);
  __syncthreads();
  for (int j = tid; j < D * radix_num_bins; j += NT) {
    uint32_t sum = 0;
#pragma unroll
    for (int c = 0; c < COLS; ++c) sum += h[j * COLS + c];
    if (sum) atomicAdd(&hist[j], sum);
  }
}

template <int NT, int VT, typename key_t>
struct radix_onesweep_storage_t {
  typedef typename radix_permute_t<key_t>::unsigned_type bits_t;
  static constexpr int NW = NT / warp_size;
  bits_t keys[NT * VT];                    // tile reordered by digit (then values)
  uint32_t hist[NW][radix_num_bins];       // per-wave counts -> wave prefixes
  uint32_t local_off[radix_num_bins];      // tile-local exclusive digit offsets
  uint32_t global_off[radix_num_bins];     // global position of local slot 0 per digit
  uint32_t scan_tot[2][NW];
  uint32_t tile;
};

// One digit pass.  `hist` is this digit's global histogram, `lookback` this
// pass's [num_tiles][256] words, `counter` this pass's tile ticket.
template <int NT, int VT, typename key_t, bool has_values>
__global__ __launch_bounds__(NT) void kernel_radix_onesweep(
    const key_t* keys_in, key_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
    int64_t count, int shift, const uint32_t* hist, uint32_t* lookback, uint32_t* counter) {
  typedef radix_permute_t<key_t> perm_t;
  typedef typename perm_t::unsigned_type bits_t;
  constexpr int NW = NT / warp_size;
  constexpr int NV = NT * VT;
  static_assert(NT >= radix_num_bins && NT % radix_num_bins == 0, "NT must cover the bins");
  __shared__ radix_onesweep_storage_t<NT, VT, key_t> s;

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int warp = tid / warp_size;
  if (tid == 0) s.tile = atomic_add_agent(counter, 1u);
  for (int i = tid; i < NW * radix_num_bins; i += NT) (&s.hist[0][0])[i] = 0;
  __syncthreads();
  const int64_t tile = s.tile;
  const int64_t base = tile * NV;
  const int tile_count = (int)min_<int64_t>(NV, count - base);

  // Each wave owns a contiguous chunk of 64*VT keys; slot i holds keys
  // [chunk + 64 i, chunk + 64 i + 64) in lane order, so wave-local rank order
  // (slot, lane) is the input order.
  const int chunk = warp * warp_size * VT;
  bits_t k[VT];
  uint32_t v[VT];
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    int j = chunk + i * warp_size + lane;
    k[i] = j < tile_count ? perm_t::to_radix_bits(keys_in[base + j]) : ~(bits_t)0;
  }
  if constexpr (has_values) {
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      int j = chunk + i * warp_size + lane;
      v[i] = j < tile_count ? vals_in[base + j] : 0u;
    }
  }

  // Wave-local stable rank of every key among equal digits.
  uint32_t r[VT];
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const int j = chunk + i * warp_size + lane;
    const bool valid = j < tile_count;
    const uint32_t d = (uint32_t)(k[i] >> shift) & 255u;
    const uint64_t active = __ballot(valid);
    const uint64_t m = match_digit8(d, active);
    const uint32_t below = mask_count_below(m);
    uint32_t c = 0;
    if (valid) c = s.hist[warp][d];
    r[i] = c + below;
    if (valid && below == 0) s.hist[warp][d] = c + (uint32_t)__popcll(m);
  }
  __syncthreads();

  // Per digit: exclusive prefix over waves, the tile total, the tile-local
  // digit offsets, the global digit base, then the look-back over tiles.
  uint32_t total = 0, gcount = 0;
  if (tid < radix_num_bins) {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      uint32_t c = s.hist[w][tid];
      s.hist[w][tid] = total;
      total += c;
    }
    atomic_store_agent(lookback + tile * radix_num_bins + tid,
                       (tile == 0 ? radix_flag_inclusive : radix_flag_aggregate) | total);
    gcount = hist[tid];
  }
  uint32_t incl_local = 0, incl_global = 0;
  if (tid < radix_num_bins) {
    incl_local = wave_scan_incl_u32(total);
    incl_global = wave_scan_incl_u32(gcount);
    if (lane == warp_size - 1) {
      s.scan_tot[0][warp] = incl_local;
      s.scan_tot[1][warp] = incl_global;
    }
  }
  __syncthreads();
  if (tid < radix_num_bins) {
    uint32_t off_local = incl_local - total, off_global = incl_global - gcount;
    for (int w = 0; w < warp; ++w) {
      off_local += s.scan_tot[0][w];
      off_global += s.scan_tot[1][w];
    }
    uint32_t excl_tiles = 0;
    if (tile > 0) {
      int64_t j = tile - 1;
      while (true) {
        uint32_t word = atomic_load_agent(lookback + j * radix_num_bins + tid);
        uint32_t flag = word & ~radix_value_mask;
        if (flag == 0) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl_tiles += word & radix_value_mask;
        if (flag == radix_flag_inclusive) break;
        --j;
      }
      atomic_store_agent(lookback + tile * radix_num_bins + tid,
                         radix_flag_inclusive | (excl_tiles + total));
    }
    s.local_off[tid] = off_local;
    s.global_off[tid] = off_global + excl_tiles - off_local;
  }
  __syncthreads();

  // Reorder the tile by digit in LDS.
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const int j = chunk + i * warp_size + lane;
    if (j < tile_count) {
      const uint32_t d = (uint32_t)(k[i] >> shift) & 255u;
      const uint32_t pos = s.local_off[d] + s.hist[warp][d] + r[i];
      s.keys[pos] = k[i];
      r[i] = pos;
    }
  }
  __syncthreads();

  uint32_t dst[VT];
#pragma unroll
  for (int i = 0; i < VT; ++i) {
    const int p = i * NT + tid;
    if (p < tile_count) {
      bits_t b = s.keys[p];
      const uint32_t d = (uint32_t)(b >> shift) & 255u;
      dst[i] = s.global_off[d] + (uint32_t)p;
      if (dst[i] < count) keys_out[dst[i]] = perm_t::from_radix_bits(b);
    }
  }
  if constexpr (has_values) {
    __syncthreads();
    uint32_t* s_vals = (uint32_t*)s.keys;
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int j = chunk + i * warp_size + lane;
      if (j < tile_count) s_vals[r[i]] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < VT; ++i) {
      const int p = i * NT + tid;
      if (p < tile_count && dst[i] < count) vals_out[dst[i]] = s_vals[p];
    }
  }
}

template <int NT, int VT, typename key_t, typename val_t>
inline size_t radix_aux_bytes(int64_t count) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  constexpr int D = (int)sizeof(key_t);
  const int64_t num_tiles = div_up(count, (int64_t)(NT * VT));
  size_t need = aux_bytes<key_t>(count);
  if (has_values) need += aux_bytes<uint32_t>(count);
  need += aux_bytes<uint32_t>((int64_t)D * radix_num_bins + (int64_t)D * num_tiles * radix_num_bins + D);
  return need;
}

// Stable LSD radix sort in place; val_t = uint32_t or empty_t.
template <int NT = 512, int VT = 16, typename key_t, typename val_t>
mgpu_status_t radix_sort(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                         val_t* vals, int64_t count) {
  constexpr bool has_values = !std::is_same_v<val_t, empty_t>;
  static_assert(!has_values || std::is_same_v<val_t, uint32_t>, "values are uint32");
  constexpr int D = (int)sizeof(key_t);
  constexpr int NV = NT * VT;
  if (count < 0) return MGPU_INVALID_ARG;
  if (count == 0) return MGPU_OK;
  if (!keys || (has_values && !vals)) return MGPU_INVALID_ARG;
  if (count > (int64_t)radix_value_mask) return MGPU_UNSUPPORTED;
  const int64_t num_tiles = div_up(count, (int64_t)NV);
  const size_t need = radix_aux_bytes<NT, VT, key_t, val_t>(count);
  if (!aux_data) {
    aux_size += need;
    return MGPU_OK;
  }
  if (aux_size < need) return MGPU_AUX_TOO_SMALL;
  void* p = aux_data;
  key_t* keys2 = advance_pointer<key_t>(p, count);
  uint32_t* vals2 = has_values ? advance_pointer<uint32_t>(p, count) : nullptr;
  const int64_t state_words = (int64_t)D * radix_num_bins + (int64_t)D * num_tiles * radix_num_bins + D;
  uint32_t* state = advance_pointer<uint32_t>(p, state_words);
  uint32_t* hist = state;
  uint32_t* lookback = hist + D * radix_num_bins;
  uint32_t* counters = lookback + (int64_t)D * num_tiles * radix_num_bins;

  hipError_t e = fill_u32(state, state_words, 0u, stream);
  if (e != hipSuccess) return hip_status(e);
  constexpr int HNT = 1024;
  int64_t hblocks = min_<int64_t>(div_up(count, (int64_t)HNT * 16), 512);
  e = launch_kernel("mgpu_radix_histogram", kernel_radix_histogram<HNT, key_t>,
                    dim3((unsigned)hblocks), dim3(HNT), 0, stream, (const key_t*)keys, count, hist);
  if (e != hipSuccess) return hip_status(e);

  key_t* k_in = keys;
  key_t* k_out = keys2;
  uint32_t* v_in = has_values ? (uint32_t*)vals : nullptr;
  uint32_t* v_out = vals2;
  for (int pass = 0; pass < D; ++pass) {
    e = launch_kernel("mgpu_radix_onesweep", kernel_radix_onesweep<NT, VT, key_t, has_values>,
                      dim3((unsigned)num_tiles), dim3(NT), 0, stream, (const key_t*)k_in, k_out,
                      (const uint32_t*)v_in, v_out, count, 8 * pass,
                      (const uint32_t*)(hist + pass * radix_num_bins),
                      lookback + (int64_t)pass * num_tiles * radix_num_bins, counters + pass);
    if (e != hipSuccess) return hip_status(e);
    std::swap(k_in, k_out);
    std::swap(v_in, v_out);
  }
  return MGPU_OK;
}

template <int NT = 512, int VT = 16, typename key_t>
mgpu_status_t radix_sort(void* aux_data, size_t& aux_size, hipStream_t stream, key_t* keys,
                         int64_t count) {
  return radix_sort<NT, VT>(aux_data, aux_size, stream, keys, (empty_t*)nullptr, count);
}

}  // namespace hip
}  // namespace mgpu
