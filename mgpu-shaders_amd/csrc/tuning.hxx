// Launch shapes used by the C ABI instantiations.
#pragma once
#include <cstdint>

template <typename key_t> struct merge_tuning { static constexpr int nt = 256, vt = 15; };
template <typename key_t> struct mergesort_tuning { static constexpr int nt = 256, vt = 15; };
template <typename key_t> struct radix_tuning { static constexpr int nt = 512, vt = 16; };
template <> struct radix_tuning<uint64_t> { static constexpr int nt = 512, vt = 8; };
