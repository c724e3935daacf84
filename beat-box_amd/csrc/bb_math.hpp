// bb_math.hpp — device float helpers for the rigid-body step (gfx950).
//
// Compiled with -ffp-contract=off: every expression below is evaluated in the
// order written (no fused multiply-add), matching the operation order of the
// reference shaders (src/shaders/extensions.slang, include/math.hpp) as
// restated here.  Matrices are row-major m[row][col]; mulvm is the row vector
// times a matrix (Slang mul(v, M)), mulmv the matrix times a column vector
// (Slang mul(M, v)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BB_DI __device__ __forceinline__

#include <utility>

namespace bb {
// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1, so
// array indices inside f are constants from the start and small private
// arrays stay in registers.
template <class F, int... I>
BB_DI void static_for_impl(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
BB_DI void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}
}  // namespace bb

namespace bb {

struct v3 {
  float x, y, z;
};
struct qt {
  v3 v;
  float w;
};
struct mat3 {
  float m[3][3];
};

BB_DI v3 V(float x, float y, float z) { return v3{x, y, z}; }
BB_DI v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
BB_DI v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
BB_DI v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
BB_DI v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
BB_DI v3 vdiv(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
BB_DI float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
BB_DI v3 cross(v3 a, v3 b) {
  return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
BB_DI float vlen(v3 a) { return __builtin_sqrtf(dot(a, a)); }
BB_DI v3 vnorm(v3 a) { return vdiv(a, vlen(a)); }
// min/max as selects: (b < a) ? b : a keeps the operand order of the CPU
// restatement for signed zeros.
BB_DI float fmn(float a, float b) { return (b < a) ? b : a; }
BB_DI float fmx(float a, float b) { return (a < b) ? b : a; }
BB_DI float fcl(float x, float lo, float hi) { return fmn(fmx(x, lo), hi); }
BB_DI float fab(float a) { return __builtin_fabsf(a); }
BB_DI v3 vabs(v3 a) { return v3{fab(a.x), fab(a.y), fab(a.z)}; }
BB_DI v3 vmin(v3 a, v3 b) { return v3{fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
BB_DI v3 vmax(v3 a, v3 b) { return v3{fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)}; }
BB_DI float cmp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
BB_DI void setcmp(v3& a, int i, float v) {
  if (i == 0) a.x = v;
  else if (i == 1) a.y = v;
  else a.z = v;
}
BB_DI float fsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

BB_DI v3 ld3(const float4& f) { return v3{f.x, f.y, f.z}; }
BB_DI float4 st3(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

BB_DI mat3 mrows(v3 r0, v3 r1, v3 r2) {
  mat3 r;
  r.m[0][0] = r0.x; r.m[0][1] = r0.y; r.m[0][2] = r0.z;
  r.m[1][0] = r1.x; r.m[1][1] = r1.y; r.m[1][2] = r1.z;
  r.m[2][0] = r2.x; r.m[2][1] = r2.y; r.m[2][2] = r2.z;
  return r;
}
BB_DI v3 mrow(const mat3& a, int i) { return v3{a.m[i][0], a.m[i][1], a.m[i][2]}; }
BB_DI v3 mcol(const mat3& a, int j) { return v3{a.m[0][j], a.m[1][j], a.m[2][j]}; }
BB_DI mat3 mT(const mat3& a) {
  mat3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}
BB_DI mat3 mabs(const mat3& a) {
  mat3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.m[i][j] = fab(a.m[i][j]);
  return r;
}
BB_DI mat3 mmul(const mat3& a, const mat3& b) {
  mat3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
  return r;
}
BB_DI v3 mulmv(const mat3& a, v3 v) {
  return v3{(a.m[0][0] * v.x + a.m[0][1] * v.y) + a.m[0][2] * v.z,
            (a.m[1][0] * v.x + a.m[1][1] * v.y) + a.m[1][2] * v.z,
            (a.m[2][0] * v.x + a.m[2][1] * v.y) + a.m[2][2] * v.z};
}
BB_DI v3 mulvm(v3 v, const mat3& a) {
  return v3{(v.x * a.m[0][0] + v.y * a.m[1][0]) + v.z * a.m[2][0],
            (v.x * a.m[0][1] + v.y * a.m[1][1]) + v.z * a.m[2][1],
            (v.x * a.m[0][2] + v.y * a.m[1][2]) + v.z * a.m[2][2]};
}

// Quaternion algebra, include/math.hpp:262-363.
BB_DI qt qmul(qt a, qt b) {
  qt r;
  r.v.x = ((a.w * b.v.x + a.v.x * b.w) + a.v.y * b.v.z) - a.v.z * b.v.y;
  r.v.y = ((a.w * b.v.y + a.v.y * b.w) + a.v.z * b.v.x) - a.v.x * b.v.z;
  r.v.z = ((a.w * b.v.z + a.v.z * b.w) + a.v.x * b.v.y) - a.v.y * b.v.x;
  r.w = ((a.w * b.w - a.v.x * b.v.x) - a.v.y * b.v.y) - a.v.z * b.v.z;
  return r;
}
BB_DI qt qconj(qt a) { return qt{v3{-a.v.x, -a.v.y, -a.v.z}, a.w}; }
BB_DI qt qnorm(qt a) {
  float m = __builtin_sqrtf(((a.v.x * a.v.x + a.v.y * a.v.y) + a.v.z * a.v.z) + a.w * a.w);
  return qt{vdiv(a.v, m), a.w / m};
}
BB_DI mat3 qmat(qt q) {
  float x2 = q.v.x + q.v.x, y2 = q.v.y + q.v.y, z2 = q.v.z + q.v.z;
  float xx = q.v.x * x2, xy = q.v.x * y2, xz = q.v.x * z2;
  float yy = q.v.y * y2, yz = q.v.y * z2, zz = q.v.z * z2;
  float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  return mrows(v3{1 - (yy + zz), xy - wz, xz + wy}, v3{xy + wz, 1 - (xx + zz), yz - wx},
               v3{xz - wy, yz + wx, 1 - (xx + yy)});
}
BB_DI qt ldq(const float4& f) { return qt{v3{f.x, f.y, f.z}, f.w}; }
BB_DI float4 stq(qt q) { return make_float4(q.v.x, q.v.y, q.v.z, q.w); }
// q (v,0) q^-1 and q^-1 (v,0) q, include/shared.inl:455-463.
BB_DI v3 qrot(qt q, v3 v) { return qmul(qmul(q, qt{v, 0.0f}), qconj(q)).v; }
BB_DI v3 qrot_inv(qt q, v3 v) { return qmul(qmul(qconj(q), qt{v, 0.0f}), q).v; }

// sin/cos by the Cephes single-precision method (S. L. Moshier, sinf.c /
// cosf.c): three-part Cody-Waite reduction by pi/4, minimax polynomials.
BB_DI void sincos_c(float x, float* so, float* co) {
  const float FOPI = 1.27323954473516f;
  const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
  int ss = 1;
  float ax = x;
  if (x < 0.0f) {
    ax = -x;
    ss = -1;
  }
  int j = (int)(FOPI * ax);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  j = j & 7;
  int sc = 1;
  if (j > 3) {
    j -= 4;
    ss = -ss;
    sc = -sc;
  }
  if (j > 1) sc = -sc;
  float z = ((ax - y * DP1) - y * DP2) - y * DP3;
  float zz = z * z;
  float ps = ((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz * z + z;
  float pc = ((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz * zz -
             0.5f * zz + 1.0f;
  float s, c;
  if (j == 1 || j == 2) {
    s = pc;
    c = ps;
  } else {
    s = ps;
    c = pc;
  }
  *so = ss < 0 ? -s : s;
  *co = sc < 0 ? -c : c;
}

// Order-preserving float <-> uint map (include/math.hpp:228-256): integer
// order equals float order with -0 < +0, so min/max through it are exact.
BB_DI uint32_t omap(float v) {
  uint32_t u = __float_as_uint(v);
  uint32_t mask = (uint32_t)(-(int32_t)(u >> 31)) | 0x80000000u;
  return u ^ mask;
}
BB_DI float ounmap(uint32_t v) {
  uint32_t mask = ((v >> 31) - 1u) | 0x80000000u;
  return __uint_as_float(v ^ mask);
}
BB_DI float omin(float a, float b) { return omap(b) < omap(a) ? b : a; }
BB_DI float omax(float a, float b) { return omap(a) < omap(b) ? b : a; }

}  // namespace bb
