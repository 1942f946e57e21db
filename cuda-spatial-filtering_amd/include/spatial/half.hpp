// IEEE binary16 storage for host code that must not depend on HIP headers.
// Stands in for the reference's __half / std::complex<__half> host storage
// (pipeline/common.hpp:52-55, packet_formats.hpp:346-348): same bits, same size.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace spatial {

struct half_t {
  uint16_t bits = 0;
};

// Round-to-nearest-even float -> binary16 (subnormals, inf and nan kept).
inline half_t float_to_half(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t mant = x & 0x7fffffu;
  const int exp = (int)((x >> 23) & 0xff);
  half_t h;
  if (exp == 0xff) {  // inf / nan
    h.bits = (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
    return h;
  }
  int e = exp - 127 + 15;
  if (e >= 0x1f) {  // overflow -> inf
    h.bits = (uint16_t)(sign | 0x7c00u);
    return h;
  }
  if (e <= 0) {  // subnormal or zero
    if (e < -10) {
      h.bits = (uint16_t)sign;
      return h;
    }
    mant |= 0x800000u;
    const int shift = 14 - e;
    uint32_t hm = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (hm & 1u))) ++hm;
    h.bits = (uint16_t)(sign | hm);
    return h;
  }
  uint32_t hm = mant >> 13;
  const uint32_t rem = mant & 0x1fffu;
  uint32_t he = (uint32_t)e;
  if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) {
    if (++hm == 0x400u) {
      hm = 0;
      ++he;
    }
  }
  h.bits = (uint16_t)(sign | (he << 10) | hm);
  if (he >= 0x1f) h.bits = (uint16_t)(sign | 0x7c00u);
  return h;
}

inline float half_to_float(half_t h) {
  const uint32_t sign = (uint32_t)(h.bits & 0x8000u) << 16;
  const uint32_t exp = (h.bits >> 10) & 0x1f, mant = h.bits & 0x3ffu;
  float f;
  if (exp == 0) {
    f = std::ldexp((float)mant, -24);
  } else if (exp == 0x1f) {
    f = mant ? NAN : INFINITY;
  } else {
    f = std::ldexp((float)(mant | 0x400u), (int)exp - 25);
  }
  return sign ? -f : f;
}

// Interleaved complex of two binary16 values (layout of std::complex<__half>).
struct complex_half {
  half_t re, im;
  complex_half() = default;
  complex_half(float r, float i) : re(float_to_half(r)), im(float_to_half(i)) {}
  float real() const { return half_to_float(re); }
  float imag() const { return half_to_float(im); }
};
static_assert(sizeof(complex_half) == 4, "complex_half must match std::complex<__half>");

}  // namespace spatial
