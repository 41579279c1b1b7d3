// Launch shapes used by the C ABI instantiations.
#pragma once
#include <cstdint>
#include <type_traits>

// Merge / mergesort tiles, measured on MI355X with 32-bit keys: merge of
// 2 x 2^27 keys 0.404 ms at 512 x 23 (256 x 15: 0.489); mergesort of 2^28
// key-value pairs 13.6 ms at 512 x 19 (256 x 15: 16.2, 512 x 11: 14.3) with two-way
// passes; with four-way passes 9.47-9.51 ms at 512 x 19 (1024 x 9: 9.42-9.50, 512 x 13:
// 9.55, 512 x 17: 9.60, 256 x 19: 9.61, 512 x 11: 9.90, 512 x 15: 10.06, 1024 x 7:
// 10.12, 256 x 23: 10.27, 1024 x 13: 12.2, 1024 x 11: 12.4, 512 x 21: 14.2).
// 8-byte keys: merge of 2 x 2^26 u64 keys 0.395 ms at 512 x 15 (256 x 15:
// 0.435, 256 x 23: 0.421, 512 x 11: 0.427); mergesort of 2^27 u64 pairs
// 9.80 ms at 512 x 13 (512 x 11: 9.87, 512 x 9: 10.34, 256 x 15: 10.65,
// 512 x 15: 12.08).
template <typename key_t, int bytes = sizeof(key_t)> struct merge_tuning { static constexpr int nt = 512, vt = 23; };
template <typename key_t> struct merge_tuning<key_t, 8> { static constexpr int nt = 512, vt = 15; };
template <typename key_t, int bytes = sizeof(key_t)> struct mergesort_tuning { static constexpr int nt = 512, vt = 19; };
template <typename key_t> struct mergesort_tuning<key_t, 8> { static constexpr int nt = 512, vt = 13; };
// Scan: 256 x 80 (20480-element tiles) for 32-bit integers, 256 x 96 for
// float32.  Larger tiles mean fewer look-back steps per byte; 2^29 int32
// exclusive: 0.775 ms at 256 x 80 vs 1.19 ms at 256 x 16 on MI355X, 0.781 at
// 256 x 96, 0.823 at 256 x 128, 0.816 at 512 x 48, 0.830 at 1024 x 32;
// float32 inclusive: 0.788 ms at 256 x 96 vs 0.818 at 256 x 80, 0.846 at
// 256 x 112, 0.842 at 512 x 64.  Forcing four waves per SIMD (128 VGPRs)
// was slower for both (int32 0.796, float32 0.820 with spills).
template <typename T> struct scan_tuning {
  static constexpr int nt = 256, vt = std::is_same<T, float>::value ? 96 : 80;
};
// Segmented scan: integers 256 x 48, float32 256 x 32.  2^29 int32 over
// 65536 segments, exclusive: 0.843 ms at 256 x 48, 0.890 at 256 x 32, 0.854
// at 256 x 64 (inclusive 0.889 / 0.921 / 1.022).  The float32 kernels carry
// float64 partial sums and at 256 x 48 would need 244 VGPRs (2 waves/SIMD).
template <typename T> struct segscan_tuning {
  static constexpr int nt = 256, vt = std::is_floating_point<T>::value ? 32 : 48;
};
