// pk_math.h — single-precision arithmetic contract shared by host and device.
//
// Every f32 result the renderer produces is defined by this file: vector
// helpers with a fixed evaluation order, and transcendental functions built
// only from IEEE add/sub/mul, correctly rounded divide and sqrt, fused
// multiply-add and integer bit operations.  Compiled with -ffp-contract=off
// on both sides (g++ for host code, hipcc for gfx950), the same inputs give
// the same bits on the CPU and on the GPU, which is what makes per-pixel and
// per-hit parity checkable bit for bit.
//
// Methods (all derived from the mathematics, not from any vendor library):
//   sin/cos : Cody–Waite reduction by pi/2 with a three-part constant,
//             Taylor polynomials on [-pi/4, pi/4] (truncation < 2e-9).
//   atan    : argument reduction x -> 1/x and x -> (x*sqrt3-1)/(x+sqrt3)
//             (Hart et al., "Computer Approximations", 1968), Taylor series
//             on |z| <= tan(pi/12) (truncation < 1e-9).
//   acos    : acos(x) = atan2(sqrt((1-x)(1+x)), x).
//   log2    : x = m*2^e, m in [sqrt(1/2), sqrt 2), log(m) = 2 atanh(s),
//             s = (m-1)/(m+1), odd Taylor series of atanh.
//   exp2    : 2^y = 2^n * e^(f ln 2), |f| <= 1/2, Taylor series of e^x.
// WGSL semantics that these stand in for: camera.wgsl:58-73 (tan, sin, cos,
// pow), shade_hit.wgsl:202-227 (atan2, acos, sin, cos), shade_miss.wgsl:37-43.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define PK_HD __host__ __device__ inline
#else
#define PK_HD inline
#endif

namespace pkm {

PK_HD uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
PK_HD float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
PK_HD float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
PK_HD float sqrt_(float x) { return __builtin_sqrtf(x); }
PK_HD float rint_(float x) { return __builtin_rintf(x); }
PK_HD float abs_(float x) { return u2f(f2u(x) & 0x7fffffffu); }
PK_HD float min_(float a, float b) { return b < a ? b : a; }
PK_HD float max_(float a, float b) { return a < b ? b : a; }
PK_HD float clamp_(float x, float lo, float hi) { return min_(max_(x, lo), hi); }
// accumulate.wgsl:29's max(color, 0) with the GPU max's NaN rule (the
// operand that is not a NaN): a sample whose color is NaN adds 0 instead of
// poisoning the running mean.  Equal to max_(c, 0) for every other c (up to
// the sign of a zero, which the running-mean update does not see).
PK_HD float max0_(float c) { return c > 0.0f ? c : 0.0f; }
PK_HD float copysign_(float m, float s) { return u2f((f2u(m) & 0x7fffffffu) | (f2u(s) & 0x80000000u)); }

constexpr float kPi = 0x1.921fb6p+1f;        // 3.14159274
constexpr float kTau = 0x1.921fb6p+2f;       // 6.28318548
constexpr float kHalfPi = 0x1.921fb6p+0f;    // 1.57079637
constexpr float kInvPi = 0x1.45f306p-2f;     // 0.318309873
constexpr float kInvTau = 0x1.45f306p-3f;    // 0.159154937
constexpr float kTwoOverPi = 0x1.45f306p-1f; // 0.636619747
// pi/2 = C1 + C2 + C3, C1 and C2 carry 12 significant bits so k*C1 and k*C2
// are exact for |k| < 2^12.
constexpr float kPio2C1 = 0x1.92p+0f;
constexpr float kPio2C2 = 0x1.fb4p-12f;
constexpr float kPio2C3 = 0x1.4442d2p-24f;

// sin(r), cos(r) for |r| <= pi/4 (Taylor to r^9 and r^10).
PK_HD float sin_poly(float r) {
  float r2 = r * r;
  float p = fma_(r2, 1.0f / 362880.0f, -1.0f / 5040.0f);
  p = fma_(p, r2, 1.0f / 120.0f);
  p = fma_(p, r2, -1.0f / 6.0f);
  return fma_(p * r2, r, r);
}
PK_HD float cos_poly(float r) {
  float r2 = r * r;
  float p = fma_(r2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
  p = fma_(p, r2, -1.0f / 720.0f);
  p = fma_(p, r2, 1.0f / 24.0f);
  p = fma_(p, r2, -0.5f);
  return fma_(p, r2, 1.0f);
}
// Reduce x to r in [-pi/4, pi/4] and the quadrant k mod 4.  Valid for
// |x| < 2^12 * pi/2, which covers every argument the renderer forms.
PK_HD float reduce_pio2(float x, int* q) {
  float k = rint_(x * kTwoOverPi);
  float r = fma_(-k, kPio2C1, x);
  r = fma_(-k, kPio2C2, r);
  r = fma_(-k, kPio2C3, r);
  *q = ((int)k) & 3;
  return r;
}
PK_HD float sinf_(float x) {
  int q; float r = reduce_pio2(x, &q);
  float s = sin_poly(r), c = cos_poly(r);
  float v = (q & 1) ? c : s;
  return (q & 2) ? -v : v;
}
PK_HD float cosf_(float x) {
  int q; float r = reduce_pio2(x, &q);
  float s = sin_poly(r), c = cos_poly(r);
  float v = (q & 1) ? s : c;
  return ((q + 1) & 2) ? -v : v;
}
PK_HD void sincosf_(float x, float* s_out, float* c_out) {
  int q; float r = reduce_pio2(x, &q);
  float s = sin_poly(r), c = cos_poly(r);
  float sv = (q & 1) ? c : s;
  float cv = (q & 1) ? s : c;
  *s_out = (q & 2) ? -sv : sv;
  *c_out = ((q + 1) & 2) ? -cv : cv;
}
PK_HD float tanf_(float x) {
  float s, c; sincosf_(x, &s, &c);
  return s / c;
}

// atan on |z| <= tan(pi/12): z - z^3/3 + ... - z^15/15 would be the next term.
PK_HD float atan_poly(float z) {
  float z2 = z * z;
  float p = fma_(z2, 1.0f / 13.0f, -1.0f / 11.0f);
  p = fma_(p, z2, 1.0f / 9.0f);
  p = fma_(p, z2, -1.0f / 7.0f);
  p = fma_(p, z2, 1.0f / 5.0f);
  p = fma_(p, z2, -1.0f / 3.0f);
  return fma_(p * z2, z, z);
}
// atan(x) for x in [0, 1]: arguments above tan(pi/12) are reduced by
// atan(x) = pi/6 + atan((x sqrt3 - 1) / (x + sqrt3)).  The reduced value is
// selected rather than branched to, so a wave forms it once whichever of its
// lanes need it.
PK_HD float atan_unit(float x) {
  constexpr float kTanPi12 = 0x1.126146p-2f;  // 0.267949194 = 2 - sqrt3
  constexpr float kSqrt3 = 0x1.bb67aep+0f;
  constexpr float kPi6 = 0x1.0c1524p-1f;
  const bool red = x > kTanPi12;
  const float xr = (x * kSqrt3 - 1.0f) / (x + kSqrt3);
  return (red ? kPi6 : 0.0f) + atan_poly(red ? xr : x);
}
// atan(a) for a >= 0.
PK_HD float atan_pos(float a) {
  bool inv = a > 1.0f;
  float r = atan_unit(inv ? 1.0f / a : a);
  return inv ? kHalfPi - r : r;
}
PK_HD float atanf_(float x) {
  float r = atan_pos(abs_(x));
  return copysign_(r, x);
}
// atan2(y, x) in [-pi, pi]; atan2(+-0, x>=0) = +-0, atan2(y, 0) = +-pi/2.
// The smaller magnitude over the larger (never above 1) goes through one
// division and one atan_unit, then the octant's reflections: the same
// operations as atan_pos(ay / ax) for ay <= ax and pi/2 - atan_pos(ax / ay)
// otherwise, without a wave evaluating both.
PK_HD float atan2f_(float y, float x) {
  float ay = abs_(y), ax = abs_(x);
  float r;
  if (ax == 0.0f && ay == 0.0f) {
    r = (f2u(x) >> 31) ? kPi : 0.0f;
  } else {
    const bool steep = !(ay <= ax);
    r = atan_unit((steep ? ax : ay) / (steep ? ay : ax));
    if (steep) r = kHalfPi - r;
    if (f2u(x) >> 31) r = kPi - r;
  }
  return copysign_(r, y);
}
PK_HD float acosf_(float x) {
  float s = sqrt_(max_(0.0f, (1.0f - x) * (1.0f + x)));
  return abs_(atan2f_(s, x));
}

// log2(x) for finite x > 0 (x <= 0 returns -inf).
PK_HD float log2f_(float x) {
  if (!(x > 0.0f)) return u2f(0xff800000u);
  uint32_t u = f2u(x);
  int e;
  if (u < 0x00800000u) {           // subnormal: scale by 2^23 first
    u = f2u(x * 8388608.0f);
    e = (int)(u >> 23) - 127 - 23;
  } else {
    e = (int)(u >> 23) - 127;
  }
  float m = u2f((u & 0x007fffffu) | 0x3f800000u);   // [1, 2)
  if (m > 0x1.6a09e6p+0f) { m *= 0.5f; e += 1; }    // [sqrt(1/2), sqrt 2)
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = fma_(s2, 1.0f / 11.0f, 1.0f / 9.0f);
  p = fma_(p, s2, 1.0f / 7.0f);
  p = fma_(p, s2, 1.0f / 5.0f);
  p = fma_(p, s2, 1.0f / 3.0f);
  float ln_m = 2.0f * fma_(p * s2, s, s);
  constexpr float kInvLn2 = 0x1.715476p+0f;
  return fma_(ln_m, kInvLn2, (float)e);
}
// 2^y; saturates to 0 / +inf outside the f32 range.
PK_HD float exp2f_(float y) {
  if (!(y == y)) return y;
  if (y >= 128.0f) return u2f(0x7f800000u);
  if (y < -149.0f) return 0.0f;
  float n = rint_(y);
  float f = y - n;                 // exact, |f| <= 1/2
  constexpr float kLn2 = 0x1.62e430p-1f;
  float x = f * kLn2;
  float p = fma_(x, 1.0f / 40320.0f, 1.0f / 5040.0f);
  p = fma_(p, x, 1.0f / 720.0f);
  p = fma_(p, x, 1.0f / 120.0f);
  p = fma_(p, x, 1.0f / 24.0f);
  p = fma_(p, x, 1.0f / 6.0f);
  p = fma_(p, x, 0.5f);
  p = fma_(p, x, 1.0f);
  p = fma_(p, x, 1.0f);
  int ni = (int)n;
  // scale by 2^ni in two steps so subnormal results round once per step
  int n1 = ni / 2, n2 = ni - n1;
  float s1 = u2f((uint32_t)(n1 + 127) << 23);
  float s2 = u2f((uint32_t)(n2 + 127) << 23);
  return (p * s1) * s2;
}
// pow(x, y) for x >= 0 as used by camera.wgsl:68 (pow(rand(), bokeh)).
PK_HD float powf_(float x, float y) {
  if (y == 0.0f) return 1.0f;
  if (x == 0.0f) return y > 0.0f ? 0.0f : u2f(0x7f800000u);
  if (x == 1.0f) return 1.0f;
  return exp2f_(y * log2f_(x));
}

// ---- vectors: fixed left-to-right evaluation, no contraction -------------
struct v3 { float x, y, z; };
PK_HD v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
PK_HD v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PK_HD v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PK_HD v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
PK_HD v3 scale(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
PK_HD v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
PK_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PK_HD v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PK_HD float length(v3 a) { return sqrt_(dot(a, a)); }
PK_HD v3 normalize(v3 a) { float l = length(a); return mk(a.x / l, a.y / l, a.z / l); }
// WGSL reflect(e1, e2) = e1 - 2 * dot(e2, e1) * e2
PK_HD v3 reflect(v3 e1, v3 e2) { float k = 2.0f * dot(e2, e1); return sub(e1, scale(e2, k)); }
// WGSL mix(a, b, t) = a * (1 - t) + b * t
PK_HD float mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
PK_HD v3 mix3(v3 a, v3 b, float t) { return mk(mix(a.x, b.x, t), mix(a.y, b.y, t), mix(a.z, b.z, t)); }
// Column-major 3x3 (WGSL mat3x3): m * v = c0*v.x + c1*v.y + c2*v.z
struct m3 { v3 c0, c1, c2; };
PK_HD v3 mat_mul(const m3& m, v3 v) {
  return mk(m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z,
            m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
            m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z);
}
// v * m = (dot(v, c0), dot(v, c1), dot(v, c2))
PK_HD v3 vec_mul(v3 v, const m3& m) { return mk(dot(v, m.c0), dot(v, m.c1), dot(v, m.c2)); }

// ---- packing (WGSL builtins) ----------------------------------------------
// pack2x16unorm: u32(0.5 + 65535 * clamp(e, 0, 1)) per component, x in low half
PK_HD uint32_t pack2x16unorm(float a, float b) {
  uint32_t lo = (uint32_t)(0.5f + 65535.0f * clamp_(a, 0.0f, 1.0f));
  uint32_t hi = (uint32_t)(0.5f + 65535.0f * clamp_(b, 0.0f, 1.0f));
  return lo | (hi << 16);
}
// q / den, correctly rounded, for an integer-valued q with |q| <= 65535 and
// den one of 127, 255, 32767, 65535: the product with the correctly rounded
// reciprocal, corrected once with its fma residual (Markstein's iteration),
// 3 operations instead of a full division.  Equal to the correctly rounded
// quotient for every such q (checked exhaustively against q / den).
PK_HD float div_by_den(float q, float den, float inv_den) {
  const float r = q * inv_den;
  const float e = fma_(-r, den, q);
  return fma_(e, inv_den, r);
}
PK_HD float unpack_unorm16(uint32_t q) { return div_by_den((float)q, 65535.0f, 1.0f / 65535.0f); }
PK_HD float unpack_snorm16(uint32_t q16) {
  int v = (int)(int16_t)(uint16_t)q16;
  return max_(div_by_den((float)v, 32767.0f, 1.0f / 32767.0f), -1.0f);
}
PK_HD float unpack_snorm8(uint32_t q8) {
  int v = (int)(int8_t)(uint8_t)q8;
  return max_(div_by_den((float)v, 127.0f, 1.0f / 127.0f), -1.0f);
}
PK_HD float unpack_unorm8(uint32_t q8) { return div_by_den((float)q8, 255.0f, 1.0f / 255.0f); }
PK_HD v3 unpack4x8snorm_xyz(uint32_t w) {
  return mk(unpack_snorm8(w & 0xffu), unpack_snorm8((w >> 8) & 0xffu), unpack_snorm8((w >> 16) & 0xffu));
}

// ---- PCG hash (camera.wgsl:46-56, shade_hit.wgsl:153-163) ------------------
// Jarzynski & Olano 2020, "Hash Functions for GPU Rendering", pcg_hash: the
// state advances by the LCG x*747796405+2891336453 and the output is the
// RXS-M-XS permutation of the old state.
PK_HD uint32_t pcg_next(uint32_t* seed) {
  uint32_t state = *seed;
  *seed = state * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
// f32(h) / 2^32 — f32(h) rounds to nearest, so values >= 2^32 - 128 give 1.0.
PK_HD float pcg_rand(uint32_t* seed) { return (float)pcg_next(seed) / 4294967296.0f; }

}  // namespace pkm
