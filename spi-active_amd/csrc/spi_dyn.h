// Per-lane Go2 dynamics for gfx950: PD + motor model, floating-base articulated-body
// algorithm specialised to the folded 13-body tree, penalty foot contact, semi-implicit
// Euler.  One lane integrates one rollout; every model constant is a compile-time value
// (spi_go2_model.h) so the compiler folds the zeros of the joint axes and offsets.
//
// Conventions: spatial motion [w; v], force [n; f], both in body coordinates, angular
// first.  Joint j of leg l rotates about coordinate axis SPI_JOINT_AXIS[3l+j] of its child
// frame; the child frame is the parent frame translated by the joint origin and rotated
// by q.  Root state rows use IsaacGym's layout (isaacgym.py:567-577): pos, quat xyzw,
// linear and angular velocity in world axes; the linear velocity is that of the base frame
// origin.  Internally the base velocity is kept in base coordinates.
#pragma once
#include <hip/hip_runtime.h>

#include "spi_go2_model.h"
#include "spi_types.h"

#define SPI_HD __host__ __device__ __forceinline__

// Arithmetic type of every runtime quantity: float on the device.  A host-only build may
// define SPI_REAL_T to an operation-counting type to take the FLOP census of the algorithm.
#ifndef SPI_REAL_T
#define SPI_REAL_T float
#endif
// A second instantiation in another translation unit (e.g. tangent arithmetic) puts the
// code in its own namespace so the two never share a symbol.
#ifndef SPI_DYN_NS
#define SPI_DYN_NS spi
#endif

namespace SPI_DYN_NS {

using Real = SPI_REAL_T;

// Compile-time constant vector (model data); never counted, folded by the compiler.
struct CV3 { float x, y, z; };

// ------------------------------------------------------------------ small algebra
struct V3 { Real x, y, z; };
SPI_HD V3 v3(Real x, Real y, Real z) { return V3{x, y, z}; }
SPI_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SPI_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SPI_HD V3 operator*(Real s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
SPI_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
SPI_HD Real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SPI_HD Real comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
SPI_HD void add_comp(V3& a, int k, Real s) {
  if (k == 0) a.x += s; else if (k == 1) a.y += s; else a.z += s;
}

// multiply by a compile-time constant that may be zero: folds to 0 without fast-math
SPI_HD Real mc(float c, Real x) { return c == 0.0f ? Real(0.0f) : c * x; }
SPI_HD V3 cvec(CV3 c) { return V3{c.x, c.y, c.z}; }
// cross product with a constant vector r
SPI_HD V3 crossc(CV3 r, V3 b) {
  return V3{mc(r.y, b.z) - mc(r.z, b.y), mc(r.z, b.x) - mc(r.x, b.z), mc(r.x, b.y) - mc(r.y, b.x)};
}

struct S3 { Real xx, yy, zz, xy, xz, yz; };            // symmetric 3x3
struct G3 { Real m00, m01, m02, m10, m11, m12, m20, m21, m22; };  // general 3x3

template <int A, int B> SPI_HD Real& sref(S3& m) {
  static_assert(A >= 0 && A < 3 && B >= 0 && B < 3, "index");
  if constexpr (A == B) return A == 0 ? m.xx : (A == 1 ? m.yy : m.zz);
  else if constexpr ((A == 0 && B == 1) || (A == 1 && B == 0)) return m.xy;
  else if constexpr ((A == 0 && B == 2) || (A == 2 && B == 0)) return m.xz;
  else return m.yz;
}
template <int A, int B> SPI_HD Real sget(const S3& m) { return sref<A, B>(const_cast<S3&>(m)); }
template <int A, int B> SPI_HD Real& gref(G3& m) {
  static_assert(A >= 0 && A < 3 && B >= 0 && B < 3, "index");
  if constexpr (A == 0) return B == 0 ? m.m00 : (B == 1 ? m.m01 : m.m02);
  else if constexpr (A == 1) return B == 0 ? m.m10 : (B == 1 ? m.m11 : m.m12);
  else return B == 0 ? m.m20 : (B == 1 ? m.m21 : m.m22);
}

SPI_HD V3 mul(const S3& s, V3 v) {
  return V3{s.xx * v.x + s.xy * v.y + s.xz * v.z, s.xy * v.x + s.yy * v.y + s.yz * v.z,
            s.xz * v.x + s.yz * v.y + s.zz * v.z};
}
SPI_HD V3 mul(const G3& g, V3 v) {
  return V3{g.m00 * v.x + g.m01 * v.y + g.m02 * v.z, g.m10 * v.x + g.m11 * v.y + g.m12 * v.z,
            g.m20 * v.x + g.m21 * v.y + g.m22 * v.z};
}
SPI_HD V3 mulT(const G3& g, V3 v) {
  return V3{g.m00 * v.x + g.m10 * v.y + g.m20 * v.z, g.m01 * v.x + g.m11 * v.y + g.m21 * v.z,
            g.m02 * v.x + g.m12 * v.y + g.m22 * v.z};
}
SPI_HD V3 col(const S3& s, int k) { return k == 0 ? V3{s.xx, s.xy, s.xz} : (k == 1 ? V3{s.xy, s.yy, s.yz} : V3{s.xz, s.yz, s.zz}); }
SPI_HD V3 row(const G3& g, int k) { return k == 0 ? V3{g.m00, g.m01, g.m02} : (k == 1 ? V3{g.m10, g.m11, g.m12} : V3{g.m20, g.m21, g.m22}); }
SPI_HD Real diag(const S3& s, int k) { return k == 0 ? s.xx : (k == 1 ? s.yy : s.zz); }

// ------------------------------------------------------------------ rotations about a coordinate axis
// Child->parent rotation R_k(q) applied to a vector (c = cos q, s = sin q).
template <int K> SPI_HD V3 rot(V3 v, Real c, Real s) {
  if (K == 0) return V3{v.x, c * v.y - s * v.z, s * v.y + c * v.z};
  if (K == 1) return V3{c * v.x + s * v.z, v.y, -s * v.x + c * v.z};
  return V3{c * v.x - s * v.y, s * v.x + c * v.y, v.z};
}
// Parent->child coordinates: R_k(q)^T v
template <int K> SPI_HD V3 rotT(V3 v, Real c, Real s) { return rot<K>(v, c, -s); }

// R S R^T for symmetric S (child coords -> parent coords), rotation about axis K.
// With (i, j) = ((K+1)%3, (K+2)%3) the rotation mixes i and j only:
//   S'_KK = S_KK,  S'_Ki = c S_Ki - s S_Kj,  S'_Kj = s S_Ki + c S_Kj,
//   2x2 block by the double angle (c2 = c^2 - s^2, s2 = 2 c s):
//   S'_ii = p + q c2 - S_ij s2,  S'_jj = p - q c2 + S_ij s2,  S'_ij = q s2 + S_ij c2,
//   p = (S_ii + S_jj) / 2, q = (S_ii - S_jj) / 2.
template <int K> SPI_HD S3 rot_sym(const S3& S, Real c, Real s, Real c2, Real s2) {
  constexpr int I = (K + 1) % 3, J = (K + 2) % 3;
  S3 o = S;
  const Real ski = sget<K, I>(S), skj = sget<K, J>(S), sij = sget<I, J>(S);
  const Real p = 0.5f * (sget<I, I>(S) + sget<J, J>(S)), q = 0.5f * (sget<I, I>(S) - sget<J, J>(S));
  const Real t = q * c2 - sij * s2;
  sref<K, I>(o) = c * ski - s * skj;
  sref<K, J>(o) = s * ski + c * skj;
  sref<I, I>(o) = p + t;
  sref<J, J>(o) = p - t;
  sref<I, J>(o) = q * s2 + sij * c2;
  return o;
}
// R G R^T for general G: row K and column K rotate as vectors, the (i, j) block by the
// double angle with its symmetric (r) and antisymmetric (t) parts:
//   G'_ii = p + q c2 - r s2, G'_jj = p - q c2 + r s2, G'_ij = q s2 + r c2 + t, G'_ji = q s2 + r c2 - t.
template <int K> SPI_HD G3 rot_gen(const G3& G, Real c, Real s, Real c2, Real s2) {
  constexpr int I = (K + 1) % 3, J = (K + 2) % 3;
  G3 o = G;
  G3& g = const_cast<G3&>(G);
  const Real gki = gref<K, I>(g), gkj = gref<K, J>(g), gik = gref<I, K>(g), gjk = gref<J, K>(g);
  const Real gii = gref<I, I>(g), gjj = gref<J, J>(g), gij = gref<I, J>(g), gji = gref<J, I>(g);
  gref<K, I>(o) = c * gki - s * gkj;
  gref<K, J>(o) = s * gki + c * gkj;
  gref<I, K>(o) = c * gik - s * gjk;
  gref<J, K>(o) = s * gik + c * gjk;
  const Real p = 0.5f * (gii + gjj), q = 0.5f * (gii - gjj), r = 0.5f * (gij + gji), t = 0.5f * (gij - gji);
  const Real u = q * c2 - r * s2, w = q * s2 + r * c2;
  gref<I, I>(o) = p + u;
  gref<J, J>(o) = p - u;
  gref<I, J>(o) = w + t;
  gref<J, I>(o) = w - t;
  return o;
}

// ------------------------------------------------------------------ spatial inertia blocks
// Articulated (or rigid) inertia [A B; B^T C]; rigid bodies have B = h^ (h = m c), C = m 1.
struct ABI { S3 A; G3 B; S3 C; };
struct F6 { V3 n, f; };
struct M6 { V3 w, v; };

SPI_HD F6 abi_mul(const ABI& I, const M6& m) {
  return F6{mul(I.A, m.w) + mul(I.B, m.v), mulT(I.B, m.w) + mul(I.C, m.v)};
}

// Compile-time rigid-body inertia of folded leg body b (1..12) about its frame origin.
template <class T> struct RBT { T m, hx, hy, hz, ixx, iyy, izz, ixy, ixz, iyz; };
using RB = RBT<float>;
constexpr RB leg_rb(int b) {
  // Ibar_O = Ic + m (|c|^2 1 - c c^T)
  double m = SPI_LINK_MASS[b - 1];
  double cx = SPI_LINK_COM[3 * (b - 1)], cy = SPI_LINK_COM[3 * (b - 1) + 1], cz = SPI_LINK_COM[3 * (b - 1) + 2];
  const double* I = &SPI_LINK_INERTIA[6 * (b - 1)];
  return RB{float(m), float(m * cx), float(m * cy), float(m * cz),
            float(I[0] + m * (cy * cy + cz * cz)), float(I[1] + m * (cx * cx + cz * cz)),
            float(I[2] + m * (cx * cx + cy * cy)), float(I[3] - m * cx * cy), float(I[4] - m * cx * cz),
            float(I[5] - m * cy * cz)};
}

// Rigid inertia times motion, using the (m, h, Ibar) form:  I v = [Ibar w + h x v; m v - h x w]
template <class T> SPI_HD F6 rb_mul(const RBT<T>& r, const M6& m) {
  V3 h{r.hx, r.hy, r.hz};
  S3 Ib{r.ixx, r.iyy, r.izz, r.ixy, r.ixz, r.iyz};
  return F6{mul(Ib, m.w) + cross(h, m.v), r.m * m.v - cross(h, m.w)};
}
template <class T> SPI_HD ABI rb_abi(const RBT<T>& r) {
  // B = h^ : [[0,-hz,hy],[hz,0,-hx],[-hy,hx,0]]
  return ABI{S3{r.ixx, r.iyy, r.izz, r.ixy, r.ixz, r.iyz},
             G3{Real(0.f), Real(-r.hz), r.hy, r.hz, Real(0.f), Real(-r.hx), Real(-r.hy), r.hx, Real(0.f)},
             S3{r.m, r.m, r.m, Real(0.f), Real(0.f), Real(0.f)}};
}
// Force cross product: v x* f = [w x n + v x f; w x f]
SPI_HD F6 crf(const M6& v, const F6& f) { return F6{cross(v.w, f.n) + cross(v.v, f.f), cross(v.w, f.f)}; }

// ------------------------------------------------------------------ joint transforms
// X = [E 0; -E r^ E] with E = R_k(q)^T and r the constant joint origin (parent coords).
// motion parent -> child
// The origin r is a compile-time CV3 (per-leg code) or a runtime V3 whose structural zeros are
// literal (leg-parallel code, spi_dyn_leg.h); both fold the zero terms.
SPI_HD V3 crossc(V3 r, V3 b) { return cross(r, b); }
SPI_HD V3 cvec(V3 c) { return c; }
template <int K, class O = CV3> SPI_HD M6 xm(const M6& m, O r, Real c, Real s) {
  return M6{rotT<K>(m.w, c, s), rotT<K>(m.v - crossc(r, m.w), c, s)};
}
// force child -> parent: X^T f = [E^T n + r x E^T f; E^T f]
template <int K, class O = CV3> SPI_HD F6 xtf(const F6& f, O r, Real c, Real s) {
  V3 ff = rot<K>(f.f, c, s);
  return F6{rot<K>(f.n, c, s) + crossc(r, ff), ff};
}
// inertia child -> parent: X^T I X = T^T (R6^T I R6) T, with T the pure translation by r:
//   A' = A - B r^ + r^ B^T - r^ C r^,  B' = B + r^ C,  C' = C    (rotated blocks)
template <int K, class O = CV3> SPI_HD ABI xtix(const ABI& I, O r, Real c, Real s) {
  const Real c2 = c * c - s * s, s2 = 2.0f * c * s;
  S3 A = rot_sym<K>(I.A, c, s, c2, s2);
  G3 B = rot_gen<K>(I.B, c, s, c2, s2);
  S3 C = rot_sym<K>(I.C, c, s, c2, s2);
  // r^ C  (columns: r x C[:,j])
  V3 rc0 = crossc(r, V3{C.xx, C.xy, C.xz});
  V3 rc1 = crossc(r, V3{C.xy, C.yy, C.yz});
  V3 rc2 = crossc(r, V3{C.xz, C.yz, C.zz});
  G3 Bn{B.m00 + rc0.x, B.m01 + rc1.x, B.m02 + rc2.x, B.m10 + rc0.y, B.m11 + rc1.y, B.m12 + rc2.y,
        B.m20 + rc0.z, B.m21 + rc1.z, B.m22 + rc2.z};
  // A' = A + r^ B^T - Bn r^   (since -B r^ - r^ C r^ = -(B + r^ C) r^ = -Bn r^)
  // (r^ B^T)_{ij} = (r x B[j,:])_i ;  (Bn r^)_{ij} = -(r x Bn[i,:])_j
  V3 rb0 = crossc(r, V3{B.m00, B.m01, B.m02});  // column 0 of r^ B^T
  V3 rb1 = crossc(r, V3{B.m10, B.m11, B.m12});
  V3 rb2 = crossc(r, V3{B.m20, B.m21, B.m22});
  V3 bn0 = crossc(r, V3{Bn.m00, Bn.m01, Bn.m02});  // -(row 0 of Bn r^)
  V3 bn1 = crossc(r, V3{Bn.m10, Bn.m11, Bn.m12});
  V3 bn2 = crossc(r, V3{Bn.m20, Bn.m21, Bn.m22});
  S3 An{A.xx + rb0.x + bn0.x, A.yy + rb1.y + bn1.y, A.zz + rb2.z + bn2.z,
        A.xy + rb1.x + bn0.y, A.xz + rb2.x + bn0.z, A.yz + rb2.y + bn1.z};
  return ABI{An, Bn, C};
}
SPI_HD void abi_add(ABI& a, const ABI& b) {
  a.A = S3{a.A.xx + b.A.xx, a.A.yy + b.A.yy, a.A.zz + b.A.zz, a.A.xy + b.A.xy, a.A.xz + b.A.xz, a.A.yz + b.A.yz};
  a.B = G3{a.B.m00 + b.B.m00, a.B.m01 + b.B.m01, a.B.m02 + b.B.m02, a.B.m10 + b.B.m10, a.B.m11 + b.B.m11,
           a.B.m12 + b.B.m12, a.B.m20 + b.B.m20, a.B.m21 + b.B.m21, a.B.m22 + b.B.m22};
  a.C = S3{a.C.xx + b.C.xx, a.C.yy + b.C.yy, a.C.zz + b.C.zz, a.C.xy + b.C.xy, a.C.xz + b.C.xz, a.C.yz + b.C.yz};
}

// ------------------------------------------------------------------ scalar math
// (a scalar type that brings its own versions defines SPI_REAL_HAS_MATH)
#ifndef SPI_REAL_HAS_MATH
// tanh.  Device: 1 - 2/(1+e^{2|x|}) for every x, absolute error within about 2 ulp of 1
// (1.2e-7); near 0 that is a relative error of 1.2e-7 / |x| - the friction and contact laws
// that use it multiply tanh by a force scale, so only its absolute error reaches the torque.
// Host builds (double checks, the census) keep an odd Taylor series near 0.
SPI_HD Real spi_tanh(Real x) {
  Real ax = fabsf(x);
#ifdef __HIP_DEVICE_COMPILE__
  Real e = __expf(2.0f * ax);
  return copysignf(1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e), x);
#else
  Real e = expf(2.0f * ax);
  Real big = 1.0f - 2.0f / (1.0f + e);
  Real x2 = ax * ax;
  Real small = ax * (1.0f + x2 * (-1.0f / 3.0f + x2 * (2.0f / 15.0f + x2 * (-17.0f / 315.0f))));
  Real t = ax < 0.0625f ? small : big;
  return copysignf(t, x);
#endif
}
SPI_HD void spi_sincos(Real x, Real* s, Real* c) {
#ifdef __HIP_DEVICE_COMPILE__
  *s = __sinf(x);
  *c = __cosf(x);
#else
  *s = sinf(x);
  *c = cosf(x);
#endif
}
SPI_HD Real spi_rcp(Real x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);   // v_rcp_f32 (1 ulp), not the IEEE division sequence
#else
  return 1.0f / x;
#endif
}
SPI_HD Real spi_rsqrt(Real x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rsqf(x);   // v_rsq_f32
#else
  return 1.0f / sqrtf(x);
#endif
}

#endif  // SPI_REAL_HAS_MATH

// ------------------------------------------------------------------ state and parameters
struct State {
  V3 p;               // base position, world
  Real qx, qy, qz, qw;  // base orientation (body -> world), xyzw
  V3 w, v;            // base spatial velocity, base coords
  Real q[12], qd[12];
};

// rotation matrix (body -> world) of the base quaternion
SPI_HD G3 quat_to_rot(Real x, Real y, Real z, Real w) {
  return G3{1.f - 2.f * (y * y + z * z), 2.f * (x * y - z * w), 2.f * (x * z + y * w),
            2.f * (x * y + z * w), 1.f - 2.f * (x * x + z * z), 2.f * (y * z - x * w),
            2.f * (x * z - y * w), 2.f * (y * z + x * w), 1.f - 2.f * (x * x + y * y)};
}

// load from IsaacGym-layout root (13) + q (12) + qd (12)
SPI_HD void state_from_root(State& st, const Real* root, const Real* q, const Real* qd, int dof_stride) {
  st.p = V3{root[0], root[1], root[2]};
  st.qx = root[3]; st.qy = root[4]; st.qz = root[5]; st.qw = root[6];
  G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  st.v = mulT(R, V3{root[7], root[8], root[9]});
  st.w = mulT(R, V3{root[10], root[11], root[12]});
#pragma unroll
  for (int j = 0; j < 12; ++j) { st.q[j] = q[j * dof_stride]; st.qd[j] = qd[j * dof_stride]; }
}
SPI_HD void state_to_root(const State& st, Real* root, Real* q, Real* qd, int dof_stride) {
  G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  V3 vw = mul(R, st.v), ww = mul(R, st.w);
  root[0] = st.p.x; root[1] = st.p.y; root[2] = st.p.z;
  root[3] = st.qx; root[4] = st.qy; root[5] = st.qz; root[6] = st.qw;
  root[7] = vw.x; root[8] = vw.y; root[9] = vw.z;
  root[10] = ww.x; root[11] = ww.y; root[12] = ww.z;
#pragma unroll
  for (int j = 0; j < 12; ++j) { q[j * dof_stride] = st.q[j]; qd[j * dof_stride] = st.qd[j]; }
}

// Per-lane physical parameters that enter the dynamics.
struct BaseInertia { Real m, hx, hy, hz, ixx, iyy, izz, ixy, ixz, iyz; };

// Base link parameters (mass, CoM, inertia about CoM) + the constant rigidly attached links.
SPI_HD BaseInertia make_base_inertia(Real m, Real cx, Real cy, Real cz, Real ixx, Real iyy, Real izz,
                                     Real ixy, Real ixz, Real iyz) {
  constexpr float fm = float(SPI_BASE_FIXED_MASS[0]);
  constexpr double fcx = SPI_BASE_FIXED_COM[0], fcy = SPI_BASE_FIXED_COM[1], fcz = SPI_BASE_FIXED_COM[2];
  constexpr double fmd = SPI_BASE_FIXED_MASS[0];
  constexpr float fhx = float(fmd * fcx), fhy = float(fmd * fcy), fhz = float(fmd * fcz);
  constexpr float fIxx = float(SPI_BASE_FIXED_INERTIA[0] + fmd * (fcy * fcy + fcz * fcz));
  constexpr float fIyy = float(SPI_BASE_FIXED_INERTIA[1] + fmd * (fcx * fcx + fcz * fcz));
  constexpr float fIzz = float(SPI_BASE_FIXED_INERTIA[2] + fmd * (fcx * fcx + fcy * fcy));
  constexpr float fIxy = float(SPI_BASE_FIXED_INERTIA[3] - fmd * fcx * fcy);
  constexpr float fIxz = float(SPI_BASE_FIXED_INERTIA[4] - fmd * fcx * fcz);
  constexpr float fIyz = float(SPI_BASE_FIXED_INERTIA[5] - fmd * fcy * fcz);
  BaseInertia b;
  b.m = m + fm;
  b.hx = m * cx + fhx; b.hy = m * cy + fhy; b.hz = m * cz + fhz;
  b.ixx = ixx + m * (cy * cy + cz * cz) + fIxx;
  b.iyy = iyy + m * (cx * cx + cz * cz) + fIyy;
  b.izz = izz + m * (cx * cx + cy * cy) + fIzz;
  b.ixy = ixy - m * cx * cy + fIxy;
  b.ixz = ixz - m * cx * cz + fIxz;
  b.iyz = iyz - m * cy * cz + fIyz;
  return b;
}

struct LaneParams {
  BaseInertia base;
  Real kp_s[4], kd_s[4], fric[4], motor[3];
};

// ------------------------------------------------------------------ control law
// Follows legged_robot_base.py:185-198 (action clip) and :542-560 (PD, clip) for
// control_mode legged_base; go2_omni.py:436-462 (hip targets x0.5) for go2_omni;
// then the motor model of active_sysid_openloop.py:356-400 on the clipped torque.
SPI_HD void control_torques(const SpiSimArgs& a, const LaneParams& lp, const Real* act, const State& st, Real* tau) {
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    const int leg = j / 3, grp = j % 3;
    Real x = fminf(fmaxf(act[j], -a.action_clip), a.action_clip);
    Real as = x * a.action_scale;
    if (grp == 0) as *= a.hip_scale;
    Real t = lp.kp_s[leg] * a.kp[j] * (as + a.q0[j] - st.q[j]) - lp.kd_s[leg] * a.kd[j] * st.qd[j];
    t = fminf(fmaxf(t, -a.effort[j]), a.effort[j]);
    if (a.motor_model == SPI_MOTOR_SCALAR) {
      t = t * lp.motor[0];
    } else if (a.motor_model == SPI_MOTOR_VEC3) {
      t = t * lp.motor[grp];
    } else if (a.motor_model == SPI_MOTOR_VEC3_TANH) {
      Real A = lp.motor[grp];
      Real B = spi_rcp(A);
      t = A * spi_tanh(B * t);
    }
    tau[j] = t;
  }
}

// ------------------------------------------------------------------ articulated-body step
// Per-joint quantities kept from the backward pass for the acceleration pass.
struct JointAcc {
  M6 U;      // I^A S  (a force, stored as two 3-vectors)
  Real dinv, u;
  M6 c;      // velocity-product acceleration
  Real cq, sq;
};

template <int L, int J> constexpr CV3 joint_origin() {
  return CV3{float(SPI_JOINT_ORIGIN[3 * (3 * L + J)]), float(SPI_JOINT_ORIGIN[3 * (3 * L + J) + 1]),
             float(SPI_JOINT_ORIGIN[3 * (3 * L + J) + 2])};
}
template <int L, int J> constexpr int joint_axis() { return SPI_JOINT_AXIS[3 * L + J]; }

// unit-axis helpers
template <int K> SPI_HD M6 axis_motion(Real s) {
  const Real z(0.f);
  return M6{V3{K == 0 ? s : z, K == 1 ? s : z, K == 2 ? s : z}, V3{z, z, z}};
}
template <int K> SPI_HD M6 vel_product(const M6& v, Real qd) {
  // v x (e_K qd, 0) = [w x e_K qd ; v x e_K qd]
  const Real z(0.f);
  V3 e{K == 0 ? qd : z, K == 1 ? qd : z, K == 2 ? qd : z};
  return M6{cross(v.w, e), cross(v.v, e)};
}
// Articulated joint step for axis K: returns Ia (downdated) and pa, fills JointAcc
template <int K> SPI_HD void joint_downdate(ABI& IA, F6& pA, Real tau, JointAcc& ja) {
  V3 Ua = col(IA.A, K);
  V3 Ub = row(IA.B, K);
  Real D = diag(IA.A, K);
  Real dinv = spi_rcp(D);
  Real u = tau - comp(pA.n, K);
  ja.U = M6{Ua, Ub};
  ja.dinv = dinv;
  ja.u = u;
  // Ia = IA - U U^T / D
  V3 Wa = dinv * Ua, Wb = dinv * Ub;
  IA.A = S3{IA.A.xx - Wa.x * Ua.x, IA.A.yy - Wa.y * Ua.y, IA.A.zz - Wa.z * Ua.z,
            IA.A.xy - Wa.x * Ua.y, IA.A.xz - Wa.x * Ua.z, IA.A.yz - Wa.y * Ua.z};
  IA.B = G3{IA.B.m00 - Wa.x * Ub.x, IA.B.m01 - Wa.x * Ub.y, IA.B.m02 - Wa.x * Ub.z,
            IA.B.m10 - Wa.y * Ub.x, IA.B.m11 - Wa.y * Ub.y, IA.B.m12 - Wa.y * Ub.z,
            IA.B.m20 - Wa.z * Ub.x, IA.B.m21 - Wa.z * Ub.y, IA.B.m22 - Wa.z * Ub.z};
  IA.C = S3{IA.C.xx - Wb.x * Ub.x, IA.C.yy - Wb.y * Ub.y, IA.C.zz - Wb.z * Ub.z,
            IA.C.xy - Wb.x * Ub.y, IA.C.xz - Wb.x * Ub.z, IA.C.yz - Wb.y * Ub.z};
  // Ia S = 0 exactly (Featherstone): the K row/column of A and row K of B vanish
  sref<K, 0>(IA.A) = Real(0.0f);
  sref<K, 1>(IA.A) = Real(0.0f);
  sref<K, 2>(IA.A) = Real(0.0f);
  gref<K, 0>(IA.B) = Real(0.0f);
  gref<K, 1>(IA.B) = Real(0.0f);
  gref<K, 2>(IA.B) = Real(0.0f);
  // pa = pA + Ia c + U u / D
  F6 Iac = abi_mul(IA, ja.c);
  Real ud = u * dinv;
  pA = F6{pA.n + Iac.n + ud * Ua, pA.f + Iac.f + ud * Ub};
}

// Contact of one foot sphere with the ground plane z = 0.  Inputs: calf pose (world) and
// calf spatial velocity (calf coords).  Output: contact wrench on the calf in calf coords.
// Normal force: spring k and damper c, the damping force saturating at the spring force;
// tangential: viscous with slope c_tangent, capped by the Coulomb cone mu Fn.
template <int L> SPI_HD F6 foot_contact(const SpiSimArgs& a, const G3& Rc, V3 pc_world, const M6& vc, Real* fout) {
  constexpr CV3 pfc{float(SPI_FOOT_POS[3 * L]), float(SPI_FOOT_POS[3 * L + 1]), float(SPI_FOOT_POS[3 * L + 2])};
  constexpr float rad = float(SPI_FOOT_RADIUS[L]);
  const V3 pf = cvec(pfc);
  V3 centre = pc_world + mul(Rc, pf);
  Real pen = rad - centre.z;
  F6 out{V3{0.f, 0.f, 0.f}, V3{0.f, 0.f, 0.f}};
  if (fout) { fout[0] = 0.f; fout[1] = 0.f; fout[2] = 0.f; }
  if (pen > 0.0f) {
    // contact point = lowest point of the sphere, as a calf-fixed material point
    V3 down{-rad * Rc.m20, -rad * Rc.m21, -rad * Rc.m22};  // R^T (0,0,-rad)
    V3 pcl = pf + down;
    V3 vl = vc.v + cross(vc.w, pcl);
    V3 vw = mul(Rc, vl);
    // Spring-damper normal force whose damping force saturates at the spring force:
    // Fn = k pen (1 - tanh(c vz / (k pen))).  Where |c vz| << k pen it is k pen - c vz, the
    // spring-damper with stiffness k and damping c; it stays in [0, 2 k pen], so it vanishes at
    // the contact onset whatever the approach speed and never pulls, and inside the contact it
    // is smooth in the state and the parameters.  The unsaturated law max(0, k pen - c vz)
    // jumps by c |vz| at touchdown, which makes the identification objective jump under
    // arbitrarily small parameter changes.
    const Real kpen = a.k_contact * pen;
    const Real fn = kpen - kpen * spi_tanh(a.c_contact * vw.z * spi_rcp(kpen));
    Real vt2 = vw.x * vw.x + vw.y * vw.y;
    Real sc = 0.0f;
    if (vt2 > 0.0f) {
      Real inv = spi_rsqrt(vt2);
      Real vt = vt2 * inv;
      Real ft = fminf(a.c_tangent * vt, a.mu * fn);
      sc = ft * inv;
    }
    V3 fw{-sc * vw.x, -sc * vw.y, fn};
    V3 fl = mulT(Rc, fw);
    out = F6{cross(pcl, fl), fl};
    if (fout) { fout[0] = fw.x; fout[1] = fw.y; fout[2] = fw.z; }
  }
  return out;
}

// Backward pass of one leg; accumulates its articulated inertia and bias into the base.
template <int L> SPI_HD void leg_backward(const SpiSimArgs& a, const State& st, const G3& R, const M6& v0,
                                          const Real* tau, ABI& IA0, F6& p0, JointAcc* ja, Real* fout) {
  constexpr int K0 = joint_axis<L, 0>(), K1 = joint_axis<L, 1>(), K2 = joint_axis<L, 2>();
  constexpr CV3 r0 = joint_origin<L, 0>(), r1 = joint_origin<L, 1>(), r2 = joint_origin<L, 2>();
  constexpr RB b0 = leg_rb(1 + 3 * L), b1 = leg_rb(2 + 3 * L), b2 = leg_rb(3 + 3 * L);
  Real c0, s0, c1, s1, c2, s2;
  spi_sincos(st.q[3 * L], &s0, &c0);
  spi_sincos(st.q[3 * L + 1], &s1, &c1);
  spi_sincos(st.q[3 * L + 2], &s2, &c2);
  ja[0].cq = c0; ja[0].sq = s0;
  ja[1].cq = c1; ja[1].sq = s1;
  ja[2].cq = c2; ja[2].sq = s2;
  // velocities
  M6 v1 = xm<K0>(v0, r0, c0, s0);
  v1.w = v1.w + axis_motion<K0>(st.qd[3 * L]).w;
  M6 v2 = xm<K1>(v1, r1, c1, s1);
  v2.w = v2.w + axis_motion<K1>(st.qd[3 * L + 1]).w;
  M6 v3 = xm<K2>(v2, r2, c2, s2);
  v3.w = v3.w + axis_motion<K2>(st.qd[3 * L + 2]).w;
  ja[0].c = vel_product<K0>(v1, st.qd[3 * L]);
  ja[1].c = vel_product<K1>(v2, st.qd[3 * L + 1]);
  ja[2].c = vel_product<K2>(v3, st.qd[3 * L + 2]);
  // calf pose in world: base R, then hip, thigh, calf rotations
  V3 pb = cvec(r0) + rot<K0>(cvec(r1) + rot<K1>(cvec(r2), c1, s1), c0, s0);  // calf origin in base coords
  V3 pcw = st.p + mul(R, pb);
  // R_calf = R * R_K0(q0) * R_K1(q1) * R_K2(q2): apply to the base rotation's rows
  G3 Rc;
  {
    // rows of Rc: row_i(R) * Rk0 * Rk1 * Rk2 == (Rk2^T Rk1^T Rk0^T row_i(R)^T)^T
    V3 x0 = rotT<K2>(rotT<K1>(rotT<K0>(V3{R.m00, R.m01, R.m02}, c0, s0), c1, s1), c2, s2);
    V3 x1 = rotT<K2>(rotT<K1>(rotT<K0>(V3{R.m10, R.m11, R.m12}, c0, s0), c1, s1), c2, s2);
    V3 x2 = rotT<K2>(rotT<K1>(rotT<K0>(V3{R.m20, R.m21, R.m22}, c0, s0), c1, s1), c2, s2);
    Rc = G3{x0.x, x0.y, x0.z, x1.x, x1.y, x1.z, x2.x, x2.y, x2.z};
  }
  F6 fext = foot_contact<L>(a, Rc, pcw, v3, fout);
  // calf (leaf)
  ABI IA = rb_abi(b2);
  F6 pA = crf(v3, rb_mul(b2, v3));
  pA = F6{pA.n - fext.n, pA.f - fext.f};
  joint_downdate<K2>(IA, pA, tau[3 * L + 2], ja[2]);
  // to thigh
  {
    ABI t = xtix<K2>(IA, r2, c2, s2);
    F6 pt = xtf<K2>(pA, r2, c2, s2);
    IA = rb_abi(b1);
    abi_add(IA, t);
    F6 pb1 = crf(v2, rb_mul(b1, v2));
    pA = F6{pb1.n + pt.n, pb1.f + pt.f};
  }
  joint_downdate<K1>(IA, pA, tau[3 * L + 1], ja[1]);
  {
    ABI t = xtix<K1>(IA, r1, c1, s1);
    F6 pt = xtf<K1>(pA, r1, c1, s1);
    IA = rb_abi(b0);
    abi_add(IA, t);
    F6 pb0 = crf(v1, rb_mul(b0, v1));
    pA = F6{pb0.n + pt.n, pb0.f + pt.f};
  }
  joint_downdate<K0>(IA, pA, tau[3 * L], ja[0]);
  {
    ABI t = xtix<K0>(IA, r0, c0, s0);
    F6 pt = xtf<K0>(pA, r0, c0, s0);
    abi_add(IA0, t);
    p0 = F6{p0.n + pt.n, p0.f + pt.f};
  }
}

template <int L> SPI_HD void leg_forward(const M6& a0, const JointAcc* ja, Real* qdd) {
  constexpr int K0 = joint_axis<L, 0>(), K1 = joint_axis<L, 1>(), K2 = joint_axis<L, 2>();
  constexpr CV3 r0 = joint_origin<L, 0>(), r1 = joint_origin<L, 1>(), r2 = joint_origin<L, 2>();
  M6 a = xm<K0>(a0, r0, ja[0].cq, ja[0].sq);
  a = M6{a.w + ja[0].c.w, a.v + ja[0].c.v};
  Real x = ja[0].dinv * (ja[0].u - dot(ja[0].U.w, a.w) - dot(ja[0].U.v, a.v));
  qdd[3 * L] = x;
  a.w = a.w + axis_motion<K0>(x).w;
  a = xm<K1>(a, r1, ja[1].cq, ja[1].sq);
  a = M6{a.w + ja[1].c.w, a.v + ja[1].c.v};
  x = ja[1].dinv * (ja[1].u - dot(ja[1].U.w, a.w) - dot(ja[1].U.v, a.v));
  qdd[3 * L + 1] = x;
  a.w = a.w + axis_motion<K1>(x).w;
  a = xm<K2>(a, r2, ja[2].cq, ja[2].sq);
  a = M6{a.w + ja[2].c.w, a.v + ja[2].c.v};
  x = ja[2].dinv * (ja[2].u - dot(ja[2].U.w, a.w) - dot(ja[2].U.v, a.v));
  qdd[3 * L + 2] = x;
}

// Solve [A B; B^T C] x = b for the symmetric positive-definite base inertia by LDL^T,
// eliminating the linear block first (it carries the total mass, the largest pivots).
SPI_HD M6 solve_base(const ABI& I, const F6& rhs) {
  // order: v (C block) then w (A block)
  Real m[6][6];
  // indices 0-2: linear (C), 3-5: angular (A)
  m[0][0] = I.C.xx; m[0][1] = I.C.xy; m[0][2] = I.C.xz;
  m[1][1] = I.C.yy; m[1][2] = I.C.yz; m[2][2] = I.C.zz;
  // coupling: rows linear i, cols angular j: (B^T)_{ij} = B_{ji}
  m[0][3] = I.B.m00; m[0][4] = I.B.m10; m[0][5] = I.B.m20;
  m[1][3] = I.B.m01; m[1][4] = I.B.m11; m[1][5] = I.B.m21;
  m[2][3] = I.B.m02; m[2][4] = I.B.m12; m[2][5] = I.B.m22;
  m[3][3] = I.A.xx; m[3][4] = I.A.xy; m[3][5] = I.A.xz;
  m[4][4] = I.A.yy; m[4][5] = I.A.yz; m[5][5] = I.A.zz;
  Real b[6] = {rhs.f.x, rhs.f.y, rhs.f.z, rhs.n.x, rhs.n.y, rhs.n.z};
  Real dinv[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    dinv[k] = spi_rcp(m[k][k]);
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      Real l = m[k][i] * dinv[k];
#pragma unroll
      for (int j = i; j < 6; ++j) m[i][j] -= l * m[k][j];
      b[i] -= l * b[k];
    }
  }
  Real x[6];
#pragma unroll
  for (int k = 5; k >= 0; --k) {
    Real s = b[k];
#pragma unroll
    for (int j = k + 1; j < 6; ++j) s -= m[k][j] * x[j];
    x[k] = s * dinv[k];
  }
  return M6{V3{x[3], x[4], x[5]}, V3{x[0], x[1], x[2]}};
}

// Forward dynamics: base acceleration (base coords, incl. gravity) and joint accelerations.
SPI_HD void forward_dynamics(const SpiSimArgs& a, const LaneParams& lp, const State& st, const Real* tau_motor,
                             M6& acc0, Real* qdd, Real* foot_force) {
  G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  M6 v0{st.w, st.v};
  Real tau[12];
#pragma unroll
  for (int j = 0; j < 12; ++j) tau[j] = tau_motor[j] - lp.fric[j / 3] * spi_tanh(st.qd[j] * a.fric_vel_inv);
  const BaseInertia& B = lp.base;
  RBT<Real> rb{B.m, B.hx, B.hy, B.hz, B.ixx, B.iyy, B.izz, B.ixy, B.ixz, B.iyz};
  ABI IA0 = rb_abi(rb);
  F6 p0 = crf(v0, rb_mul(rb, v0));
  JointAcc ja[12];
  leg_backward<0>(a, st, R, v0, tau, IA0, p0, ja + 0, foot_force ? foot_force + 0 : nullptr);
  leg_backward<1>(a, st, R, v0, tau, IA0, p0, ja + 3, foot_force ? foot_force + 3 : nullptr);
  leg_backward<2>(a, st, R, v0, tau, IA0, p0, ja + 6, foot_force ? foot_force + 6 : nullptr);
  leg_backward<3>(a, st, R, v0, tau, IA0, p0, ja + 9, foot_force ? foot_force + 9 : nullptr);
  M6 a0 = solve_base(IA0, F6{-1.0f * p0.n, -1.0f * p0.f});  // gravity-free (relative) acceleration
  leg_forward<0>(a0, ja + 0, qdd);
  leg_forward<1>(a0, ja + 3, qdd);
  leg_forward<2>(a0, ja + 6, qdd);
  leg_forward<3>(a0, ja + 9, qdd);
  // true acceleration: add the gravity field [0; R^T g]
  Real g = a.gravity;
  acc0 = M6{a0.w, a0.v + V3{-g * R.m20, -g * R.m21, -g * R.m22}};
}

// Semi-implicit Euler over one sub-interval h.
SPI_HD void integrate(State& st, const M6& acc0, const Real* qdd, Real h) {
  st.w = st.w + h * acc0.w;
  st.v = st.v + h * acc0.v;
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    st.qd[j] += h * qdd[j];
    st.q[j] += h * st.qd[j];
  }
  G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  st.p = st.p + h * mul(R, st.v);
  // q <- q (x) [sin(th/2) w/|w|, cos(th/2)], th = h|w|: series in (h|w|/2)^2, then renormalise
  Real hh = 0.5f * h;
  Real t2 = hh * hh * dot(st.w, st.w);
  Real sfac = hh * (1.0f - t2 * (1.0f / 6.0f));
  Real cw = 1.0f - 0.5f * t2 + t2 * t2 * (1.0f / 24.0f);
  Real dx = sfac * st.w.x, dy = sfac * st.w.y, dz = sfac * st.w.z;
  Real x = st.qw * dx + st.qx * cw + st.qy * dz - st.qz * dy;
  Real y = st.qw * dy - st.qx * dz + st.qy * cw + st.qz * dx;
  Real z = st.qw * dz + st.qx * dy - st.qy * dx + st.qz * cw;
  Real w = st.qw * cw - st.qx * dx - st.qy * dy - st.qz * dz;
  Real inv = spi_rsqrt(x * x + y * y + z * z + w * w);
  st.qx = x * inv; st.qy = y * inv; st.qz = z * inv; st.qw = w * inv;
}

// One physics step: motor torque held over `substeps` integration sub-intervals.
SPI_HD void physics_step(const SpiSimArgs& a, const LaneParams& lp, State& st, const Real* tau_motor,
                         Real* foot_force) {
  Real h = a.dt / Real(a.substeps);
  for (int k = 0; k < a.substeps; ++k) {
    M6 acc0;
    Real qdd[12];
    forward_dynamics(a, lp, st, tau_motor, acc0, qdd, foot_force);
    integrate(st, acc0, qdd, h);
  }
}

// Divergence guard: sum of |base velocity| and |joint velocity| components.
// "x < limit", false for NaN and infinities also in files built with -ffinite-math-only, where
// the compiler may assume no value is NaN and rewrite or drop the test: on the device the bits
// pass through an empty asm, whose output the compiler knows nothing about (host builds of
// this header are IEEE).
SPI_HD bool below_limit(float x, float limit) {
#ifdef __HIP_DEVICE_COMPILE__
  unsigned int u = __float_as_uint(x);
  asm volatile("" : "+v"(u));
  return (u & 0x7f800000u) != 0x7f800000u && x < limit;
#else
  return x < limit;
#endif
}
// other scalar types (host builds in double, the operation-counting type)
template <class T> SPI_HD bool below_limit(const T& x, float limit) { return x < T(limit); }

SPI_HD Real state_norm(const State& st) {
  Real s = fabsf(st.w.x) + fabsf(st.w.y) + fabsf(st.w.z) + fabsf(st.v.x) + fabsf(st.v.y) + fabsf(st.v.z);
#pragma unroll
  for (int j = 0; j < 12; ++j) s += fabsf(st.qd[j]);
  return s;
}

// Weighted error of the state after a control step against a 37-float target record:
// tracking mode sums w . squared errors of [pos, quat (raw xyzw), linvel, angvel, q, qd];
// reference mode is scripts/eval.py:290-292 (L2 norms of pos, quat, q).
template <class TgtT> SPI_HD Real step_cost(const SpiCostArgs& c, const State& st, const TgtT* tg) {
  G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  V3 vw = mul(R, st.v), ww = mul(R, st.w);
  Real e0(0.f), e1(0.f), e2(0.f), e3(0.f), e4(0.f), e5(0.f), d;
  d = st.p.x - tg[0]; e0 += d * d;
  d = st.p.y - tg[1]; e0 += d * d;
  d = st.p.z - tg[2]; e0 += d * d;
  d = st.qx - tg[3]; e1 += d * d;
  d = st.qy - tg[4]; e1 += d * d;
  d = st.qz - tg[5]; e1 += d * d;
  d = st.qw - tg[6]; e1 += d * d;
  d = vw.x - tg[7]; e2 += d * d;
  d = vw.y - tg[8]; e2 += d * d;
  d = vw.z - tg[9]; e2 += d * d;
  d = ww.x - tg[10]; e3 += d * d;
  d = ww.y - tg[11]; e3 += d * d;
  d = ww.z - tg[12]; e3 += d * d;
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    d = st.q[j] - tg[13 + j]; e4 += d * d;
    d = st.qd[j] - tg[25 + j]; e5 += d * d;
  }
  if (c.mode == SPI_COST_TRACKING)
    return c.w[0] * e0 + c.w[1] * e1 + c.w[2] * e2 + c.w[3] * e3 + c.w[4] * e4 + c.w[5] * e5;
  return c.w[0] * sqrtf(e0) + c.w[1] * sqrtf(e1) + c.w[4] * sqrtf(e4);
}

}  // namespace SPI_DYN_NS
