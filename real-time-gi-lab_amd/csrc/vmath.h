// Device-side f32 vector helpers with a fixed evaluation order. Every expression is written so that the rounding
// sequence is the one of the WGSL it restates (left-to-right sums, dot(a, b) = (a.x b.x + a.y b.y) + a.z b.z, no
// contraction: the sources are compiled with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct f3 {
  float x, y, z;
};
struct i3 {
  int x, y, z;
};

// f32 -> bf16 with round to nearest even (NaN stays a quiet NaN) and back: the precision form's storage.
__device__ __forceinline__ uint32_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float((h & 0xffffu) << 16); }

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 splat3(float s) { return f3{s, s, s}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 abs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ f3 min3(f3 a, f3 b) { return f3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ f3 max3(f3 a, f3 b) { return f3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ f3 floor3(f3 a) { return f3{floorf(a.x), floorf(a.y), floorf(a.z)}; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 normalize3(f3 a) {
  float l = sqrtf(dot3(a, a));
  return a / l;
}
__device__ __forceinline__ f3 tof3(i3 a) { return f3{(float)a.x, (float)a.y, (float)a.z}; }
// f32 -> i32 conversion that saturates like WGSL's instead of being undefined out of range.
__device__ __forceinline__ int f2i(float f) {
  f = fminf(fmaxf(f, -2147483648.0f), 2147483520.0f);
  return (int)f;
}
__device__ __forceinline__ i3 f2i3(f3 a) { return i3{f2i(a.x), f2i(a.y), f2i(a.z)}; }
__device__ __forceinline__ i3 operator+(i3 a, i3 b) { return i3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ i3 operator-(i3 a, i3 b) { return i3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ i3 operator*(i3 a, int s) { return i3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ i3 operator+(i3 a, int s) { return i3{a.x + s, a.y + s, a.z + s}; }
__device__ __forceinline__ i3 operator-(i3 a, int s) { return i3{a.x - s, a.y - s, a.z - s}; }
__device__ __forceinline__ i3 shr3(i3 a, int s) { return i3{a.x >> s, a.y >> s, a.z >> s}; }
__device__ __forceinline__ i3 clampi3(i3 a, i3 lo, i3 hi) {
  return i3{min(max(a.x, lo.x), hi.x), min(max(a.y, lo.y), hi.y), min(max(a.z, lo.z), hi.z)};
}
__device__ __forceinline__ i3 maxi3(i3 a, i3 b) { return i3{max(a.x, b.x), max(a.y, b.y), max(a.z, b.z)}; }
// 0 <= a < n per component (n >= 0): one unsigned compare each, a negative component wrapping above n.
__device__ __forceinline__ bool inside3(i3 a, int n) {
  return (uint32_t)a.x < (uint32_t)n && (uint32_t)a.y < (uint32_t)n && (uint32_t)a.z < (uint32_t)n;
}

// 2^-e for 0 <= e < 126 as an f32 built from its exponent bits: equal to 1.0f / (float)(1u << e), which is exact.
__device__ __forceinline__ float inv_pow2(int e) { return __int_as_float((127 - e) << 23); }

// x / 255 and x / 32767, correctly rounded, for the integer x of an 8-bit unorm (0..255) and of a 16-bit snorm
// (-32768..32767): a multiply by the rounded reciprocal and one FMA correction step. For these inputs the result
// equals IEEE division (checked exhaustively in exact rational arithmetic), at 3 instructions instead of about 10.
__device__ __forceinline__ float div255(float x) {
  const float r = 1.0f / 255.0f;
  float q = x * r;
  return __builtin_fmaf(__builtin_fmaf(-q, 255.0f, x), r, q);
}
// log2 of a positive normal float: x = m 2^e with m in [sqrt(1/2), sqrt(2)], ln m = 2 atanh(s), s = (m - 1) / (m + 1),
// by the series s + s^3/3 + ... + s^9/9 (|s| < 0.172: the next term is below 2^-30 of the sum). Plain IEEE float
// operations in a fixed order, so that every conforming implementation gives the same bits (the level blend of the
// dynamic layer's read depends on it; a library log2f differs between implementations in the last place).
__device__ __forceinline__ float log2_series(float x) {
  uint32_t u = __float_as_uint(x);
  int e = (int)(u >> 23) - 127;
  float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f), s2 = s * s;
  float p = 1.0f / 9.0f;
  p = p * s2 + 1.0f / 7.0f;
  p = p * s2 + 1.0f / 5.0f;
  p = p * s2 + 1.0f / 3.0f;
  p = p * s2 + 1.0f;
  return (float)e + ((2.0f * s) * p) * 1.44269504f;
}
__device__ __forceinline__ float div32767(float x) {
  const float r = 1.0f / 32767.0f;
  float q = x * r;
  return __builtin_fmaf(__builtin_fmaf(-q, 32767.0f, x), r, q);
}
