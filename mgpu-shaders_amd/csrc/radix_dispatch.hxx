// Radix sort launch shape by key size, values and count, for the C ABI and
// the serving engine.  Small sorts are bound by launches and by how few tiles
// there are to spread over 256 CUs, so they take small tiles; large ones take
// the tile that moves HBM fastest.  Measured on MI355X (Gkeys/s, whole sort):
//   32-bit kv:    2^22: 512x8 33.2, 512x20 32.4; 6M: 35.9 vs 33.6; 8M: 40.1 vs 41.0;
//                 2^28: 512x20 61.2 (512x16 57.7, 512x24 4.47 ms vs 4.36)
//   32-bit keys:  2^22: 512x8 34.8, 512x12 33.2, 512x20 31.3; 16M: 512x12 67.7, 512x20 62.1;
//                 25M: 512x12 75.3, 512x24 73.3; 2^28: 512x24 91.5, 512x20 89.6, 512x12 86.2
//   64-bit kv:    1M: 512x8 5.5, 512x12 4.9; 16M: 512x12 18.4, 512x8 17.7;
//                 2^27: 256x24 21.3, 512x12 19.8, 256x32 20.4, 512x16 19.4
//   64-bit keys:  2^27: 512x12 29.8, 512x16 29.9, 256x24 29.9, 256x32 30.1
#pragma once
#include <type_traits>
#include "mgpu/hip/radix.hxx"

namespace mgpu {
namespace hip {

template <typename key_t, typename val_t>
inline mgpu_status_t radix_sort_tuned(void* aux, size_t& aux_size, hipStream_t stream,
                                      key_t* keys, val_t* vals, int64_t count) {
  constexpr bool kv = !std::is_same_v<val_t, empty_t>;
  if constexpr (sizeof(key_t) == 8) {
    if (count <= ((int64_t)1 << 22))
      return radix_sort<512, 8, 8, 512>(aux, aux_size, stream, keys, vals, count);
    if constexpr (kv) return radix_sort<256, 24, 8, 512>(aux, aux_size, stream, keys, vals, count);
    return radix_sort<512, 12, 8, 512>(aux, aux_size, stream, keys, vals, count);
  } else if constexpr (kv) {
    if (count <= ((int64_t)6 << 20))
      return radix_sort<512, 8, 8, 512>(aux, aux_size, stream, keys, vals, count);
    return radix_sort<512, 20, 8, 512>(aux, aux_size, stream, keys, vals, count);
  } else {
    if (count <= ((int64_t)1 << 22))
      return radix_sort<512, 8, 8, 512>(aux, aux_size, stream, keys, vals, count);
    if (count <= ((int64_t)1 << 25))
      return radix_sort<512, 12, 8, 512>(aux, aux_size, stream, keys, vals, count);
    return radix_sort<512, 24, 8, 512>(aux, aux_size, stream, keys, vals, count);
  }
}

}  // namespace hip
}  // namespace mgpu
