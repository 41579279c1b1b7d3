// Launch shapes used by the C ABI instantiations.
#pragma once
#include <cstdint>

template <typename key_t> struct merge_tuning { static constexpr int nt = 256, vt = 15; };
template <typename key_t> struct mergesort_tuning { static constexpr int nt = 256, vt = 15; };
// Radix: scatter block nt x vt (tile = nt*vt keys), count block cnt threads.
// 512 x 24 (12288-key tiles) measured fastest for 2^28 u32 key-value on MI355X.
template <typename key_t> struct radix_tuning { static constexpr int nt = 512, vt = 24, cnt = 1024; };
template <> struct radix_tuning<uint64_t> { static constexpr int nt = 512, vt = 12, cnt = 512; };
