// Counter-based SplitMix64 input generators (Steele, Lea & Flood 2014), the
// same formulas on host and device so both produce identical inputs:
//   word(seed, stream, i) = mix64(seed * 2^40 + stream * 2^36 + i + 0x9E3779B97F4A7C15)
//   u32(stream, i)        = word(...) >> 32
// Kinds:
//   u32         u32(0, i)
//   lowentropy  u32(0,i) & u32(1,i) & u32(2,i) & u32(3,i)
//   i32_0_16    u32(0, i) >> 28
//   f32_01      (u32(0, i) >> 8) * 2^-24          (exact multiples of 2^-24)
//   u64         u32(0, i) << 32 | u32(1, i)
//   f32_pm1     (u32(0, i) >> 8) * 2^-23 - 1      (exact, in [-1, 1))
//   iota        i
//   sorted_u32  i * s + (u32(0, i) % s), s = max(1, 2^32 / n)   (ascending)
//   sorted_u64  i * s + (u64 % s),       s = max(1, 2^64 / n)   (ascending)
//   sorted_f32  (i + (u32(0, i) >> 8) * 2^-24) / n               (ascending)
// (sorted kinds use n = the generated count and ignore the offset)
#pragma once
#include "transform.hxx"

namespace mgpu {
namespace hip {

enum gen_kind_t : int {
  gen_u32 = 0,
  gen_lowentropy = 1,
  gen_i32_0_16 = 2,
  gen_f32_01 = 3,
  gen_u64 = 4,
  gen_f32_pm1 = 5,
  gen_iota = 6,
  gen_sorted_u32 = 7,
  gen_sorted_u64 = 8,
  gen_sorted_f32 = 9,
};

__host__ __device__ inline uint64_t splitmix64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint32_t gen_word32(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t c = (seed << 40) + (stream << 36) + i + 0x9E3779B97F4A7C15ull;
  return (uint32_t)(splitmix64_mix(c) >> 32);
}

// Element i of a generator kind, as raw bits (4 or 8 bytes); n is the
// length of the generated array (used by the sorted kinds).
__host__ __device__ inline uint64_t gen_element(int kind, uint64_t seed, uint64_t i, uint64_t n = 1) {
  switch (kind) {
    case gen_sorted_u32: {
      uint64_t s = n ? (((uint64_t)1 << 32) / n) : 1;
      s = s ? s : 1;
      return i * s + gen_word32(seed, 0, i) % s;
    }
    case gen_sorted_u64: {
      uint64_t s = n > 1 ? (~0ull / n) : ~0ull;
      uint64_t r = ((uint64_t)gen_word32(seed, 0, i) << 32) | gen_word32(seed, 1, i);
      return i * s + r % s;
    }
    case gen_sorted_f32: {
      float f = ((float)i + (float)(gen_word32(seed, 0, i) >> 8) * (1.0f / 16777216.0f)) / (float)n;
      uint32_t b; __builtin_memcpy(&b, &f, 4); return b;
    }
    case gen_u32: return gen_word32(seed, 0, i);
    case gen_lowentropy:
      return gen_word32(seed, 0, i) & gen_word32(seed, 1, i) & gen_word32(seed, 2, i) &
             gen_word32(seed, 3, i);
    case gen_i32_0_16: return gen_word32(seed, 0, i) >> 28;
    case gen_f32_01: {
      float f = (float)(gen_word32(seed, 0, i) >> 8) * (1.0f / 16777216.0f);
      uint32_t b; __builtin_memcpy(&b, &f, 4); return b;
    }
    case gen_u64: return ((uint64_t)gen_word32(seed, 0, i) << 32) | gen_word32(seed, 1, i);
    case gen_f32_pm1: {
      float f = (float)(gen_word32(seed, 0, i) >> 8) * (1.0f / 8388608.0f) - 1.0f;
      uint32_t b; __builtin_memcpy(&b, &f, 4); return b;
    }
    case gen_iota: return i;
    default: return 0;
  }
}

// Fill out[0, count) with a generator kind (4- or 8-byte elements), one
// element per thread through transform().
inline mgpu_status_t generate(void* out, int64_t count, int elem_bytes, int kind, uint64_t seed,
                              uint64_t offset, hipStream_t stream) {
  if (elem_bytes == 8) {
    uint64_t* o = (uint64_t*)out;
    return transform<256>(count, stream, [=] __device__(int64_t i) {
      o[i] = kind >= gen_sorted_u32 ? gen_element(kind, seed, (uint64_t)i, (uint64_t)count)
                                    : gen_element(kind, seed, offset + (uint64_t)i);
    }, "mgpu_generate");
  }
  uint32_t* o = (uint32_t*)out;
  return transform<256>(count, stream, [=] __device__(int64_t i) {
    o[i] = (uint32_t)(kind >= gen_sorted_u32 ? gen_element(kind, seed, (uint64_t)i, (uint64_t)count)
                                             : gen_element(kind, seed, offset + (uint64_t)i));
  }, "mgpu_generate");
}

}  // namespace hip
}  // namespace mgpu
