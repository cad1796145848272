// Host-side IEEE binary16 and bfloat16 conversions (round to nearest, ties to
// even; with subnormals, overflow to infinity, NaN kept quiet), matching the
// device's conversions in their default modes.
#pragma once

#include <cstdint>
#include <cstring>

namespace gd {

inline uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u)  // infinity or NaN
    return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // >= 65520 rounds to infinity
  if (ax >= 0x38800000u) {                                     // normal result (>= 2^-14)
    const uint32_t mant = ax & 0x7fffffu;
    const uint32_t e = (ax >> 23) - 127u + 15u;
    uint32_t h = (e << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // carries into the exponent
    return (uint16_t)(sign | h);
  }
  if (ax < 0x33000000u) return (uint16_t)sign;  // below 2^-25: rounds to zero
  // subnormal result, unit 2^-24: value = m24 * 2^(E-23), E in [-25, -15]
  const uint32_t m24 = (ax & 0x7fffffu) | 0x800000u;
  const int E = (int)(ax >> 23) - 127;
  const int shift = -(E + 1);  // 14..24
  uint32_t h = m24 >> shift;
  const uint32_t rem = m24 & ((1u << shift) - 1u);
  const uint32_t halfway = 1u << (shift - 1);
  if (rem > halfway || (rem == halfway && (h & 1u))) ++h;
  return (uint16_t)(sign | h);
}

inline float f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  uint32_t x;
  if (e == 0x1fu) {
    x = sign | 0x7f800000u | (m << 13);
  } else if (e != 0) {
    x = sign | ((e + 112u) << 23) | (m << 13);
  } else if (m == 0) {
    x = sign;
  } else {  // subnormal: normalise
    int s = 0;
    while (!(m & 0x400u)) {
      m <<= 1;
      ++s;
    }
    x = sign | ((uint32_t)(113 - s) << 23) | ((m & 0x3ffu) << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

// bfloat16 (the upper 16 bits of a float): round to nearest, ties to even, on
// the bit pattern (subnormals and overflow to infinity follow from the carry;
// NaN kept quiet), as the device's v_cvt_pk_bf16_f32.
inline uint16_t f32_to_bf16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}

inline float bf16_to_f32(uint16_t h) {
  const uint32_t x = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

}  // namespace gd
