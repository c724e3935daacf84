// oracle_math.hpp — TEST INFRASTRUCTURE ONLY (the CPU oracle).
//
// Float helpers of the CPU restatement of the reference rigid-body step.
// Every helper evaluates in the operation order written here, built with
// -ffp-contract=off so no multiply-add is fused.  Nothing in the product
// includes this file.
//
// Conventions restated from the reference's shader language (Slang):
//  * float3x3 is indexed m[row][col]; float3x3(r0, r1, r2) takes ROWS
//    (src/shaders/extensions.slang:228-242 uses this).
//  * mul(v, M) is the row vector times M; mul(M, v) is M times the column v.
//  * dot(a, b) = (a.x*b.x + a.y*b.y) + a.z*b.z.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

struct f3 {
  float x, y, z;
};
struct quat {
  f3 v;
  float w;
};
struct m3 {
  float m[3][3];
};

static inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
static inline f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
static inline f3 muls(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
static inline f3 mulv(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
static inline float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline float length(f3 a) { return std::sqrt(dot(a, a)); }
static inline f3 normalize(f3 a) { return divs(a, length(a)); }
static inline float fminr(float a, float b) { return (b < a) ? b : a; }
static inline float fmaxr(float a, float b) { return (a < b) ? b : a; }
static inline float clampf(float x, float lo, float hi) { return fminr(fmaxr(x, lo), hi); }
static inline float fabsr(float a) { return std::fabs(a); }
static inline f3 absv(f3 a) { return f3{fabsr(a.x), fabsr(a.y), fabsr(a.z)}; }
static inline f3 minv(f3 a, f3 b) { return f3{fminr(a.x, b.x), fminr(a.y, b.y), fminr(a.z, b.z)}; }
static inline f3 maxv(f3 a, f3 b) { return f3{fmaxr(a.x, b.x), fmaxr(a.y, b.y), fmaxr(a.z, b.z)}; }
static inline float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline void setc(f3& a, int i, float v) {
  if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}
static inline float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// ---- matrices (row-major m[row][col]) ------------------------------------
static inline m3 rows(f3 r0, f3 r1, f3 r2) {
  m3 r;
  r.m[0][0] = r0.x; r.m[0][1] = r0.y; r.m[0][2] = r0.z;
  r.m[1][0] = r1.x; r.m[1][1] = r1.y; r.m[1][2] = r1.z;
  r.m[2][0] = r2.x; r.m[2][1] = r2.y; r.m[2][2] = r2.z;
  return r;
}
static inline f3 row(const m3& a, int i) { return f3{a.m[i][0], a.m[i][1], a.m[i][2]}; }
static inline f3 col(const m3& a, int j) { return f3{a.m[0][j], a.m[1][j], a.m[2][j]}; }
static inline m3 transpose(const m3& a) {
  m3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}
static inline m3 absm(const m3& a) {
  m3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = fabsr(a.m[i][j]);
  return r;
}
// mul(M, N)
static inline m3 mulmm(const m3& a, const m3& b) {
  m3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
  return r;
}
// mul(M, v): column vector
static inline f3 mulmv(const m3& a, f3 v) {
  return f3{(a.m[0][0] * v.x + a.m[0][1] * v.y) + a.m[0][2] * v.z,
            (a.m[1][0] * v.x + a.m[1][1] * v.y) + a.m[1][2] * v.z,
            (a.m[2][0] * v.x + a.m[2][1] * v.y) + a.m[2][2] * v.z};
}
// mul(v, M): row vector
static inline f3 mulvm(f3 v, const m3& a) {
  return f3{(v.x * a.m[0][0] + v.y * a.m[1][0]) + v.z * a.m[2][0],
            (v.x * a.m[0][1] + v.y * a.m[1][1]) + v.z * a.m[2][1],
            (v.x * a.m[0][2] + v.y * a.m[1][2]) + v.z * a.m[2][2]};
}

// ---- quaternions (include/math.hpp:262-363) -----------------------------
static inline quat qmul(quat a, quat b) {
  quat r;
  r.v.x = ((a.w * b.v.x + a.v.x * b.w) + a.v.y * b.v.z) - a.v.z * b.v.y;
  r.v.y = ((a.w * b.v.y + a.v.y * b.w) + a.v.z * b.v.x) - a.v.x * b.v.z;
  r.v.z = ((a.w * b.v.z + a.v.z * b.w) + a.v.x * b.v.y) - a.v.y * b.v.x;
  r.w = ((a.w * b.w - a.v.x * b.v.x) - a.v.y * b.v.y) - a.v.z * b.v.z;
  return r;
}
static inline quat qconj(quat a) { return quat{f3{-a.v.x, -a.v.y, -a.v.z}, a.w}; }
static inline float qmag(quat a) {
  return std::sqrt(((a.v.x * a.v.x + a.v.y * a.v.y) + a.v.z * a.v.z) + a.w * a.w);
}
static inline quat qnormalize(quat a) {
  float m = qmag(a);
  return quat{divs(a.v, m), a.w / m};
}
static inline m3 qto_matrix(quat q) {
  float x2 = q.v.x + q.v.x, y2 = q.v.y + q.v.y, z2 = q.v.z + q.v.z;
  float xx = q.v.x * x2, xy = q.v.x * y2, xz = q.v.x * z2;
  float yy = q.v.y * y2, yz = q.v.y * z2, zz = q.v.z * z2;
  float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  return rows(f3{1 - (yy + zz), xy - wz, xz + wy}, f3{xy + wz, 1 - (xx + zz), yz - wx},
              f3{xz - wy, yz + wx, 1 - (xx + yy)});
}
// RigidBody::rotate_vector / rotate_vector_inverse (include/shared.inl:455-463)
static inline f3 qrotate(quat q, f3 v) { return qmul(qmul(q, quat{v, 0.0f}), qconj(q)).v; }
static inline f3 qrotate_inv(quat q, f3 v) { return qmul(qmul(qconj(q), quat{v, 0.0f}), q).v; }

// ---- sin / cos ------------------------------------------------------------
// Single-precision sine/cosine after Cephes sinf.c/cosf.c (S. L. Moshier):
// three-part Cody-Waite reduction by pi/4 and the Cephes minimax polynomials.
// Written out so that CPU and GPU evaluate the same operations.
static inline void sincos_cephes(float x, float* s_out, float* c_out) {
  const float FOPI = 1.27323954473516f;
  const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f,
              DP3 = 3.77489497744594108e-8f;
  int sgn_s = 1;
  float ax = x;
  if (x < 0.0f) { ax = -x; sgn_s = -1; }
  int j = (int)(FOPI * ax);
  float y = (float)j;
  if (j & 1) { j += 1; y += 1.0f; }
  j = j & 7;
  int sgn_c = 1;
  if (j > 3) { j -= 4; sgn_s = -sgn_s; sgn_c = -sgn_c; }
  if (j > 1) sgn_c = -sgn_c;
  float z = ((ax - y * DP1) - y * DP2) - y * DP3;
  float zz = z * z;
  float ps = ((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz * z + z;
  float pc = ((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz * zz -
             0.5f * zz + 1.0f;
  float s, c;
  if (j == 1 || j == 2) { s = pc; c = ps; } else { s = ps; c = pc; }
  *s_out = sgn_s < 0 ? -s : s;
  *c_out = sgn_c < 0 ? -c : c;
}

// ---- order-preserving float <-> uint (include/math.hpp:228-256) -----------
static inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static inline uint32_t order_map(float v) {
  uint32_t u = f2u(v);
  uint32_t mask = (uint32_t)(-(int32_t)(u >> 31)) | 0x80000000u;
  return u ^ mask;
}
static inline float order_unmap(uint32_t v) {
  uint32_t mask = ((v >> 31) - 1u) | 0x80000000u;
  return u2f(v ^ mask);
}

}  // namespace oracle
