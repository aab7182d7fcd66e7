// Reverse-mode (adjoint) of the per-lane Go2 rollout of spi_dyn.h.
//
// One integration sub-interval maps (x, tau_m, theta) -> x+ through the forward dynamics
// a = FD(x, tau_m, theta) and semi-implicit Euler.  Its adjoint is built from two identities:
//   * FD solves ID(x, a, theta) = (0, tau_joint(x, tau_m, theta)) with M = dID/da symmetric, so
//     for an incoming acceleration adjoint abar, mu = M^-1 abar (an articulated-body solve with
//     no bias terms) and
//       xbar += (d tau_joint / dx)^T mu_joint - (d ID / dx)^T mu,  thetabar likewise,
//       taubar_m += mu_joint;
//   * (d ID / dx)^T mu is the reverse of the recursive Newton-Euler pass (velocities and
//     accelerations out to the feet, forces back to the base) at the solved accelerations,
//     with the foot contact wrenches (which depend on the base pose, joint angles and the calf
//     velocity) differentiated through the contact law and the forward kinematics.
// Joint transforms use d(X m)/dq = -crm(S, X m) and d(X^T f)/dq = X^T crf(S, f) (S the joint
// axis in child coordinates).  Control law, joint friction, quaternion update and tracking cost
// have closed-form adjoints.  The derivative is that of the discrete rollout actually computed
// (same contact / cone / clip branches as the value).
#pragma once
#include "spi_dyn.h"

namespace SPI_DYN_NS {

// ------------------------------------------------------------------ adjoint state
struct StateBar {
  V3 p;
  Real qx, qy, qz, qw;
  V3 w, v;
  Real q[12], qd[12];
};
SPI_HD V3 v3zero() { return V3{Real(0.f), Real(0.f), Real(0.f)}; }
SPI_HD void zero_bar(StateBar& b) {
  b.p = v3zero(); b.w = v3zero(); b.v = v3zero();
  b.qx = b.qy = b.qz = b.qw = Real(0.f);
#pragma unroll
  for (int j = 0; j < 12; ++j) { b.q[j] = Real(0.f); b.qd[j] = Real(0.f); }
}

// parameter adjoint in the SoA row order (mass, com xyz, Ixx Iyy Izz Ixy Ixz Iyz, kp_s x4,
// kd_s x4, friction x4, motor x3)
struct ParBar { Real g[SPI_P_COUNT]; };
SPI_HD void zero_par(ParBar& b) {
#pragma unroll
  for (int k = 0; k < SPI_P_COUNT; ++k) b.g[k] = Real(0.f);
}

// ------------------------------------------------------------------ small algebra
SPI_HD M6 crm(const M6& v, const M6& m) { return M6{cross(v.w, m.w), cross(v.w, m.v) + cross(v.v, m.w)}; }
SPI_HD Real dot6(const M6& m, const F6& f) { return dot(m.w, f.n) + dot(m.v, f.f); }
SPI_HD M6 madd(const M6& a, const M6& b) { return M6{a.w + b.w, a.v + b.v}; }
SPI_HD F6 fadd(const F6& a, const F6& b) { return F6{a.n + b.n, a.f + b.f}; }
SPI_HD F6 fneg(const F6& a) { return F6{Real(-1.f) * a.n, Real(-1.f) * a.f}; }
SPI_HD F6 fzero() { return F6{v3zero(), v3zero()}; }
SPI_HD M6 mzero() { return M6{v3zero(), v3zero()}; }
template <int K> SPI_HD V3 unit() {
  return V3{Real(K == 0 ? 1.f : 0.f), Real(K == 1 ? 1.f : 0.f), Real(K == 2 ? 1.f : 0.f)};
}
template <int K> SPI_HD M6 axis6() { return M6{unit<K>(), v3zero()}; }
// crm(S, y) and crf(S, f) for S = [e_K; 0]
template <int K> SPI_HD M6 crm_axis(const M6& y) { return M6{cross(unit<K>(), y.w), cross(unit<K>(), y.v)}; }
template <int K> SPI_HD F6 crf_axis(const F6& f) { return F6{cross(unit<K>(), f.n), cross(unit<K>(), f.f)}; }
SPI_HD Real comp6n(const F6& f, int k) { return comp(f.n, k); }

SPI_HD G3 gmul(const G3& a, const G3& b) {
  return G3{a.m00 * b.m00 + a.m01 * b.m10 + a.m02 * b.m20, a.m00 * b.m01 + a.m01 * b.m11 + a.m02 * b.m21,
            a.m00 * b.m02 + a.m01 * b.m12 + a.m02 * b.m22, a.m10 * b.m00 + a.m11 * b.m10 + a.m12 * b.m20,
            a.m10 * b.m01 + a.m11 * b.m11 + a.m12 * b.m21, a.m10 * b.m02 + a.m11 * b.m12 + a.m12 * b.m22,
            a.m20 * b.m00 + a.m21 * b.m10 + a.m22 * b.m20, a.m20 * b.m01 + a.m21 * b.m11 + a.m22 * b.m21,
            a.m20 * b.m02 + a.m21 * b.m12 + a.m22 * b.m22};
}
SPI_HD G3 gtrans(const G3& a) { return G3{a.m00, a.m10, a.m20, a.m01, a.m11, a.m21, a.m02, a.m12, a.m22}; }
SPI_HD G3 gadd(const G3& a, const G3& b) {
  return G3{a.m00 + b.m00, a.m01 + b.m01, a.m02 + b.m02, a.m10 + b.m10, a.m11 + b.m11,
            a.m12 + b.m12, a.m20 + b.m20, a.m21 + b.m21, a.m22 + b.m22};
}
SPI_HD G3 gzero() { return G3{Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f)}; }
// outer product a b^T
SPI_HD G3 outer(V3 a, V3 b) {
  return G3{a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z};
}
SPI_HD Real gdot(const G3& a, const G3& b) {
  return a.m00 * b.m00 + a.m01 * b.m01 + a.m02 * b.m02 + a.m10 * b.m10 + a.m11 * b.m11 + a.m12 * b.m12 +
         a.m20 * b.m20 + a.m21 * b.m21 + a.m22 * b.m22;
}
// child -> parent rotation about coordinate axis K (the matrix of rot<K>)
template <int K> SPI_HD G3 axis_rot(Real c, Real s) {
  const Real o(1.f), z(0.f);
  if (K == 0) return G3{o, z, z, z, c, Real(-1.f) * s, z, s, c};
  if (K == 1) return G3{c, z, s, z, o, z, Real(-1.f) * s, z, c};
  return G3{c, Real(-1.f) * s, z, s, c, z, z, z, o};
}
// d/dq of axis_rot<K>: R_K [e_K]x, contracted with Gbar
template <int K> SPI_HD Real axis_rot_vjp(const G3& Gbar, Real c, Real s) {
  const Real z(0.f);
  G3 D;
  if (K == 0) D = G3{z, z, z, z, Real(-1.f) * s, Real(-1.f) * c, z, c, Real(-1.f) * s};
  else if (K == 1) D = G3{Real(-1.f) * s, z, c, z, z, z, Real(-1.f) * c, z, Real(-1.f) * s};
  else D = G3{Real(-1.f) * s, Real(-1.f) * c, z, c, Real(-1.f) * s, z, z, z, z};
  return gdot(Gbar, D);
}
// quaternion (xyzw) adjoint of R = quat_to_rot(q) for an incoming Rbar (SB: any adjoint state
// with members qx, qy, qz, qw)
template <class SB> SPI_HD void quat_rot_vjp(Real x, Real y, Real z, Real w, const G3& B, SB& sb) {
  const Real two(2.f), four(4.f);
  sb.qx += two * (y * B.m01 + z * B.m02 + y * B.m10 - w * B.m12 + z * B.m20 + w * B.m21) -
           four * x * (B.m11 + B.m22);
  sb.qy += two * (x * B.m01 + w * B.m02 + x * B.m10 + z * B.m12 - w * B.m20 + z * B.m21) -
           four * y * (B.m00 + B.m22);
  sb.qz += two * (Real(-1.f) * w * B.m01 + x * B.m02 + w * B.m10 + y * B.m12 + x * B.m20 + y * B.m21) -
           four * z * (B.m00 + B.m11);
  sb.qw += two * (Real(-1.f) * z * B.m01 + y * B.m02 + z * B.m10 - x * B.m12 - y * B.m20 + x * B.m21);
}
SPI_HD V3 mulG(const G3& g, V3 v) { return mul(g, v); }

// gradient of y^T I x with respect to (m, h, Ibar) of a rigid inertia, accumulated
struct RBBar { Real m, hx, hy, hz, ixx, iyy, izz, ixy, ixz, iyz; };
SPI_HD void rb_bilinear_vjp(const M6& y, const M6& x, Real scale, RBBar& b) {
  // y^T I x = yw.(Ibar xw + h x xv) + yv.(m xv - h x xw)
  b.m += scale * dot(y.v, x.v);
  V3 gh = cross(x.v, y.w) - cross(x.w, y.v);
  b.hx += scale * gh.x; b.hy += scale * gh.y; b.hz += scale * gh.z;
  b.ixx += scale * y.w.x * x.w.x;
  b.iyy += scale * y.w.y * x.w.y;
  b.izz += scale * y.w.z * x.w.z;
  b.ixy += scale * (y.w.x * x.w.y + y.w.y * x.w.x);
  b.ixz += scale * (y.w.x * x.w.z + y.w.z * x.w.x);
  b.iyz += scale * (y.w.y * x.w.z + y.w.z * x.w.y);
}
// (m, h, Ibar) of the base (make_base_inertia) -> mass, com, inertia-about-CoM rows 0..9 of g
SPI_HD void base_inertia_vjp(const RBBar& b, Real m, Real cx, Real cy, Real cz, Real* g) {
  // h = m c (+ const): dh/dm = c, dh/dc = m
  // Ibar_xx = ixx + m (cy^2 + cz^2), Ibar_xy = ixy - m cx cy, ...
  g[0] += b.m + b.hx * cx + b.hy * cy + b.hz * cz + b.ixx * (cy * cy + cz * cz) + b.iyy * (cx * cx + cz * cz) +
          b.izz * (cx * cx + cy * cy) - b.ixy * cx * cy - b.ixz * cx * cz - b.iyz * cy * cz;
  const Real two(2.f);
  g[1] += b.hx * m + two * m * cx * (b.iyy + b.izz) - m * (b.ixy * cy + b.ixz * cz);
  g[2] += b.hy * m + two * m * cy * (b.ixx + b.izz) - m * (b.ixy * cx + b.iyz * cz);
  g[3] += b.hz * m + two * m * cz * (b.ixx + b.iyy) - m * (b.ixz * cx + b.iyz * cy);
  g[4] += b.ixx; g[5] += b.iyy; g[6] += b.izz;
  g[7] += b.ixy; g[8] += b.ixz; g[9] += b.iyz;
}

// ------------------------------------------------------------------ foot contact adjoint
// Reverse of foot_contact<L>: out = (pcl x fl, fl).  fbar (motion-like) is the adjoint of
// out; accumulates Rcbar, pcwbar and vcbar.
template <int L> SPI_HD void foot_contact_vjp(const SpiSimArgs& a, const G3& Rc, V3 pcw, const M6& vc,
                                              const M6& fbar, G3& Rcb, V3& pcwb, F6& vcb) {
  constexpr CV3 pfc{float(SPI_FOOT_POS[3 * L]), float(SPI_FOOT_POS[3 * L + 1]), float(SPI_FOOT_POS[3 * L + 2])};
  constexpr float rad = float(SPI_FOOT_RADIUS[L]);
  const V3 pf = cvec(pfc);
  const V3 centre = pcw + mul(Rc, pf);
  const Real pen = rad - centre.z;
  if (!(pen > 0.0f)) return;
  const V3 down{-rad * Rc.m20, -rad * Rc.m21, -rad * Rc.m22};
  const V3 pcl = pf + down;
  const V3 vl = vc.v + cross(vc.w, pcl);
  const V3 vw = mul(Rc, vl);
  // foot_contact's normal force fn = k pen (1 - tanh(u)), u = c vz / (k pen)
  const Real kpen = a.k_contact * pen;
  const Real u = a.c_contact * vw.z * spi_rcp(kpen);
  const Real th = spi_tanh(u);
  const Real fn = kpen - kpen * th;
  const Real vt2 = vw.x * vw.x + vw.y * vw.y;
  Real sc(0.f), vt(0.f);
  bool cone = false;
  if (vt2 > 0.0f) {
    vt = sqrtf(vt2);
    const Real visc = a.c_tangent * vt, cap = a.mu * fn;
    cone = cap < visc;
    sc = cone ? cap / vt : Real(a.c_tangent);
  }
  const V3 fw{Real(-1.f) * sc * vw.x, Real(-1.f) * sc * vw.y, fn};
  const V3 fl = mulT(Rc, fw);
  // out = (pcl x fl, fl)
  V3 pclb = cross(fl, fbar.w);
  V3 flb = cross(fbar.w, pcl) + fbar.v;
  // fl = Rc^T fw
  V3 fwb = mul(Rc, flb);
  Rcb = gadd(Rcb, outer(fw, flb));
  // fw
  Real scb = Real(-1.f) * (fwb.x * vw.x + fwb.y * vw.y);
  V3 vwb{Real(-1.f) * sc * fwb.x, Real(-1.f) * sc * fwb.y, Real(0.f)};
  Real fnb = fwb.z;
  if (vt2 > 0.0f && cone) {
    // sc = mu fn / vt
    fnb += scb * a.mu / vt;
    const Real vtb = Real(-1.f) * scb * a.mu * fn / (vt * vt);
    vwb.x += vtb * vw.x / vt;
    vwb.y += vtb * vw.y / vt;
  }
  // d fn / d pen = k ((1 - tanh u) + u sech^2 u),  d fn / d vz = -c sech^2 u
  const Real sech2 = Real(1.f) - th * th;
  const Real penb = a.k_contact * (Real(1.f) - th + u * sech2) * fnb;
  vwb.z += Real(-1.f) * a.c_contact * sech2 * fnb;
  // vw = Rc vl
  V3 vlb = mulT(Rc, vwb);
  Rcb = gadd(Rcb, outer(vwb, vl));
  // vl = vc.v + vc.w x pcl
  vcb.f = vcb.f + vlb;
  vcb.n = vcb.n + cross(pcl, vlb);
  pclb = pclb + cross(vlb, vc.w);
  // pcl = pf + down, down_i = -rad Rc[2][i]
  Rcb.m20 += Real(-rad) * pclb.x;
  Rcb.m21 += Real(-rad) * pclb.y;
  Rcb.m22 += Real(-rad) * pclb.z;
  // pen = rad - centre.z, centre = pcw + Rc pf
  const V3 centreb{Real(0.f), Real(0.f), Real(-1.f) * penb};
  pcwb = pcwb + centreb;
  Rcb = gadd(Rcb, outer(centreb, pf));
}

// ------------------------------------------------------------------ RNEA adjoint of one leg
// s = -(mu_base . F_base + sum_j mu_j tau_j) differentiated with respect to the state, given
// the solved accelerations (base a0 relative to the free-fall frame, qdd).  Fb_bar is the
// (negated) base covector; mu_neg holds the negated joint covectors.  v0b collects the base
// velocity adjoint, Rb the base rotation adjoint, sb the rest.
template <int L> SPI_HD void leg_id_vjp(const SpiSimArgs& a, const State& st, const G3& R, const M6& v0,
                                        const M6& a0, const Real* qdd, const M6& Fb_bar, const Real* mu_neg,
                                        F6& v0b, G3& Rb, StateBar& sb) {
  constexpr int K0 = joint_axis<L, 0>(), K1 = joint_axis<L, 1>(), K2 = joint_axis<L, 2>();
  constexpr CV3 r0 = joint_origin<L, 0>(), r1 = joint_origin<L, 1>(), r2 = joint_origin<L, 2>();
  constexpr RB b0 = leg_rb(1 + 3 * L), b1 = leg_rb(2 + 3 * L), b2 = leg_rb(3 + 3 * L);
  const int j0 = 3 * L, j1 = 3 * L + 1, j2 = 3 * L + 2;
  Real c0, s0, c1, s1, c2, s2;
  spi_sincos(st.q[j0], &s0, &c0);
  spi_sincos(st.q[j1], &s1, &c1);
  spi_sincos(st.q[j2], &s2, &c2);
  // ---- forward: velocities, accelerations, forces
  const M6 xv1 = xm<K0>(v0, r0, c0, s0);
  M6 v1 = xv1; v1.w = v1.w + axis_motion<K0>(st.qd[j0]).w;
  const M6 xv2 = xm<K1>(v1, r1, c1, s1);
  M6 v2 = xv2; v2.w = v2.w + axis_motion<K1>(st.qd[j1]).w;
  const M6 xv3 = xm<K2>(v2, r2, c2, s2);
  M6 v3 = xv3; v3.w = v3.w + axis_motion<K2>(st.qd[j2]).w;
  const M6 xa1 = xm<K0>(a0, r0, c0, s0);
  M6 a1 = madd(xa1, vel_product<K0>(v1, st.qd[j0])); a1.w = a1.w + axis_motion<K0>(qdd[j0]).w;
  const M6 xa2 = xm<K1>(a1, r1, c1, s1);
  M6 a2 = madd(xa2, vel_product<K1>(v2, st.qd[j1])); a2.w = a2.w + axis_motion<K1>(qdd[j1]).w;
  const M6 xa3 = xm<K2>(a2, r2, c2, s2);
  M6 a3 = madd(xa3, vel_product<K2>(v3, st.qd[j2])); a3.w = a3.w + axis_motion<K2>(qdd[j2]).w;
  // calf pose in world (as leg_backward)
  const V3 pb_in = cvec(r1) + rot<K1>(cvec(r2), c1, s1);
  const V3 pb = cvec(r0) + rot<K0>(pb_in, c0, s0);
  const V3 pcw = st.p + mul(R, pb);
  const G3 R0 = axis_rot<K0>(c0, s0), R1 = axis_rot<K1>(c1, s1), R2 = axis_rot<K2>(c2, s2);
  const G3 R12 = gmul(R1, R2), A = gmul(R0, R12);
  const G3 Rc = gmul(R, A);
  const F6 fext = foot_contact<L>(a, Rc, pcw, v3, nullptr);
  const F6 h3 = rb_mul(b2, v3), h2 = rb_mul(b1, v2), h1 = rb_mul(b0, v1);
  const F6 f3 = fadd(fadd(rb_mul(b2, a3), crf(v3, h3)), fneg(fext));
  const F6 f2 = fadd(rb_mul(b1, a2), crf(v2, h2));
  const F6 f1 = fadd(rb_mul(b0, a1), crf(v1, h1));
  const F6 F3 = f3;
  const F6 F2 = fadd(f2, xtf<K2>(F3, r2, c2, s2));
  const F6 F1 = fadd(f1, xtf<K1>(F2, r1, c1, s1));
  // ---- reverse phase 1: force adjoints from the base down to the calf
  M6 t = xm<K0>(Fb_bar, r0, c0, s0);
  sb.q[j0] += dot6(t, crf_axis<K0>(F1));
  M6 F1b = t; F1b.w = F1b.w + Real(mu_neg[j0]) * unit<K0>();
  t = xm<K1>(F1b, r1, c1, s1);
  sb.q[j1] += dot6(t, crf_axis<K1>(F2));
  M6 F2b = t; F2b.w = F2b.w + Real(mu_neg[j1]) * unit<K1>();
  t = xm<K2>(F2b, r2, c2, s2);
  sb.q[j2] += dot6(t, crf_axis<K2>(F3));
  M6 F3b = t; F3b.w = F3b.w + Real(mu_neg[j2]) * unit<K2>();
  // ---- reverse phase 2: calf, thigh, hip
  // calf: f3 = I a3 + v3 x* I v3 - fext
  F6 a3b = rb_mul(b2, F3b);
  F6 v3b = fneg(fadd(crf(F3b, h3), rb_mul(b2, crm(v3, F3b))));
  G3 Rcb = gzero();
  V3 pcwb = v3zero();
  foot_contact_vjp<L>(a, Rc, pcw, v3, M6{Real(-1.f) * F3b.w, Real(-1.f) * F3b.v}, Rcb, pcwb, v3b);
  // a3 = X2 a2 + e qdd + v3 x (e qd)
  F6 a2b = fadd(rb_mul(b1, F2b), xtf<K2>(a3b, r2, c2, s2));
  sb.q[j2] -= dot6(crm_axis<K2>(xa3), a3b);
  v3b = fadd(v3b, crf(axis_motion<K2>(st.qd[j2]), a3b));
  sb.qd[j2] += dot6(crm(v3, axis6<K2>()), a3b);
  // v3 = X2 v2 + e qd
  F6 v2b = fadd(fneg(fadd(crf(F2b, h2), rb_mul(b1, crm(v2, F2b)))), xtf<K2>(v3b, r2, c2, s2));
  sb.qd[j2] += comp(v3b.n, K2);
  sb.q[j2] -= dot6(crm_axis<K2>(xv3), v3b);
  // thigh: a2 = X1 a1 + ..., v2 = X1 v1 + ...
  F6 a1b = fadd(rb_mul(b0, F1b), xtf<K1>(a2b, r1, c1, s1));
  sb.q[j1] -= dot6(crm_axis<K1>(xa2), a2b);
  v2b = fadd(v2b, crf(axis_motion<K1>(st.qd[j1]), a2b));
  sb.qd[j1] += dot6(crm(v2, axis6<K1>()), a2b);
  F6 v1b = fadd(fneg(fadd(crf(F1b, h1), rb_mul(b0, crm(v1, F1b)))), xtf<K1>(v2b, r1, c1, s1));
  sb.qd[j1] += comp(v2b.n, K1);
  sb.q[j1] -= dot6(crm_axis<K1>(xv2), v2b);
  // hip: a1 = X0 a0 + ... (a0 is fixed), v1 = X0 v0 + ...
  sb.q[j0] -= dot6(crm_axis<K0>(xa1), a1b);
  v1b = fadd(v1b, crf(axis_motion<K0>(st.qd[j0]), a1b));
  sb.qd[j0] += dot6(crm(v1, axis6<K0>()), a1b);
  v0b = fadd(v0b, xtf<K0>(v1b, r0, c0, s0));
  sb.qd[j0] += comp(v1b.n, K0);
  sb.q[j0] -= dot6(crm_axis<K0>(xv1), v1b);
  // ---- calf pose: pcw = p + R pb(q0, q1), Rc = R R0 R1 R2
  sb.p = sb.p + pcwb;
  Rb = gadd(Rb, outer(pcwb, pb));
  const V3 pbb = mulT(R, pcwb);
  // pb = r0 + R0 pb_in, pb_in = r1 + R1 r2
  G3 R0b = outer(pbb, pb_in);
  const V3 pinb = mulT(R0, pbb);
  G3 R1b = outer(pinb, cvec(r2));
  Rb = gadd(Rb, gmul(Rcb, gtrans(A)));
  const G3 Ab = gmul(gtrans(R), Rcb);
  R0b = gadd(R0b, gmul(Ab, gtrans(R12)));
  const G3 Bb = gmul(gtrans(R0), Ab);
  R1b = gadd(R1b, gmul(Bb, gtrans(R2)));
  const G3 R2b = gmul(gtrans(R1), Bb);
  sb.q[j0] += axis_rot_vjp<K0>(R0b, c0, s0);
  sb.q[j1] += axis_rot_vjp<K1>(R1b, c1, s1);
  sb.q[j2] += axis_rot_vjp<K2>(R2b, c2, s2);
}

// ------------------------------------------------------------------ M^-1 (no bias terms)
template <int L> SPI_HD void leg_minv_backward(const State& st, const Real* rhs_joint, ABI& IA0, F6& p0,
                                               JointAcc* ja) {
  constexpr int K0 = joint_axis<L, 0>(), K1 = joint_axis<L, 1>(), K2 = joint_axis<L, 2>();
  constexpr CV3 r0 = joint_origin<L, 0>(), r1 = joint_origin<L, 1>(), r2 = joint_origin<L, 2>();
  constexpr RB b0 = leg_rb(1 + 3 * L), b1 = leg_rb(2 + 3 * L), b2 = leg_rb(3 + 3 * L);
  Real c0, s0, c1, s1, c2, s2;
  spi_sincos(st.q[3 * L], &s0, &c0);
  spi_sincos(st.q[3 * L + 1], &s1, &c1);
  spi_sincos(st.q[3 * L + 2], &s2, &c2);
  ja[0].cq = c0; ja[0].sq = s0; ja[1].cq = c1; ja[1].sq = s1; ja[2].cq = c2; ja[2].sq = s2;
  ja[0].c = mzero(); ja[1].c = mzero(); ja[2].c = mzero();
  ABI IA = rb_abi(b2);
  F6 pA = fzero();
  joint_downdate<K2>(IA, pA, rhs_joint[3 * L + 2], ja[2]);
  {
    ABI t = xtix<K2>(IA, r2, c2, s2);
    F6 pt = xtf<K2>(pA, r2, c2, s2);
    IA = rb_abi(b1);
    abi_add(IA, t);
    pA = pt;
  }
  joint_downdate<K1>(IA, pA, rhs_joint[3 * L + 1], ja[1]);
  {
    ABI t = xtix<K1>(IA, r1, c1, s1);
    F6 pt = xtf<K1>(pA, r1, c1, s1);
    IA = rb_abi(b0);
    abi_add(IA, t);
    pA = pt;
  }
  joint_downdate<K0>(IA, pA, rhs_joint[3 * L], ja[0]);
  {
    ABI t = xtix<K0>(IA, r0, c0, s0);
    F6 pt = xtf<K0>(pA, r0, c0, s0);
    abi_add(IA0, t);
    p0 = fadd(p0, pt);
  }
}
// out = M^-1 rhs; rhs/out layout [base (w-part 3, v-part 3) | joints 12]
SPI_HD void minv_apply(const LaneParams& lp, const State& st, const Real* rhs, Real* out) {
  const BaseInertia& B = lp.base;
  RBT<Real> rb{B.m, B.hx, B.hy, B.hz, B.ixx, B.iyy, B.izz, B.ixy, B.ixz, B.iyz};
  ABI IA0 = rb_abi(rb);
  F6 p0 = fzero();
  JointAcc ja[12];
  leg_minv_backward<0>(st, rhs + 6, IA0, p0, ja + 0);
  leg_minv_backward<1>(st, rhs + 6, IA0, p0, ja + 3);
  leg_minv_backward<2>(st, rhs + 6, IA0, p0, ja + 6);
  leg_minv_backward<3>(st, rhs + 6, IA0, p0, ja + 9);
  const F6 f{V3{rhs[0], rhs[1], rhs[2]}, V3{rhs[3], rhs[4], rhs[5]}};
  const M6 a0 = solve_base(IA0, F6{f.n - p0.n, f.f - p0.f});
  leg_forward<0>(a0, ja + 0, out + 6);
  leg_forward<1>(a0, ja + 3, out + 6);
  leg_forward<2>(a0, ja + 6, out + 6);
  leg_forward<3>(a0, ja + 9, out + 6);
  out[0] = a0.w.x; out[1] = a0.w.y; out[2] = a0.w.z;
  out[3] = a0.v.x; out[4] = a0.v.y; out[5] = a0.v.z;
}

// ------------------------------------------------------------------ one sub-interval
// Reverse of {FD; integrate(h)} at the sub-interval's start state st.  On entry xb holds the
// adjoint of the end state; on exit that of st.  Accumulates taub (motor torque, held over the
// physics step) and pb (parameters).
// raw = the base link's own (mass, com x, y, z), the parameters the base inertia is built from.
SPI_HD void substep_vjp(const SpiSimArgs& a, const LaneParams& lp, const Real* raw, const State& st,
                        const Real* tau_m, Real h, StateBar& xb, Real* taub, ParBar& pb) {
  // forward recompute
  M6 acc;
  Real qdd[12];
  forward_dynamics(a, lp, st, tau_m, acc, qdd, nullptr);
  const G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  const V3 gb{Real(-1.f) * a.gravity * R.m20, Real(-1.f) * a.gravity * R.m21, Real(-1.f) * a.gravity * R.m22};
  const M6 a_rel{acc.w, acc.v - gb};
  const V3 w1 = st.w + h * acc.w, v1 = st.v + h * acc.v;
  const Real hh = 0.5f * h;
  const Real t2 = hh * hh * dot(w1, w1);
  const Real sfac = hh * (1.0f - t2 * (1.0f / 6.0f));
  const Real cw = 1.0f - 0.5f * t2 + t2 * t2 * (1.0f / 24.0f);
  const Real dx = sfac * w1.x, dy = sfac * w1.y, dz = sfac * w1.z;
  const Real qx = st.qx, qy = st.qy, qz = st.qz, qw = st.qw;
  const Real ux = qw * dx + qx * cw + qy * dz - qz * dy;
  const Real uy = qw * dy - qx * dz + qy * cw + qz * dx;
  const Real uz = qw * dz + qx * dy - qy * dx + qz * cw;
  const Real uw = qw * cw - qx * dx - qy * dy - qz * dz;
  const Real inv = spi_rsqrt(ux * ux + uy * uy + uz * uz + uw * uw);
  const Real nx = ux * inv, ny = uy * inv, nz = uz * inv, nw = uw * inv;
  // ---- integrate adjoint
  // quat1 = u / |u|
  const Real pr = nx * xb.qx + ny * xb.qy + nz * xb.qz + nw * xb.qw;
  const Real ubx = (xb.qx - nx * pr) * inv, uby = (xb.qy - ny * pr) * inv;
  const Real ubz = (xb.qz - nz * pr) * inv, ubw = (xb.qw - nw * pr) * inv;
  StateBar ob;                        // adjoint of the start state
  zero_bar(ob);
  ob.qx = ubx * cw - uby * dz + ubz * dy - ubw * dx;
  ob.qy = ubx * dz + uby * cw - ubz * dx - ubw * dy;
  ob.qz = Real(-1.f) * ubx * dy + uby * dx + ubz * cw - ubw * dz;
  ob.qw = ubx * dx + uby * dy + ubz * dz + ubw * cw;
  const Real dxb = ubx * qw + uby * qz - ubz * qy - ubw * qx;
  const Real dyb = Real(-1.f) * ubx * qz + uby * qw + ubz * qx - ubw * qy;
  const Real dzb = ubx * qy - uby * qx + ubz * qw - ubw * qz;
  const Real cwb = ubx * qx + uby * qy + ubz * qz + ubw * qw;
  V3 w1b = xb.w + sfac * V3{dxb, dyb, dzb};
  const Real sfacb = dxb * w1.x + dyb * w1.y + dzb * w1.z;
  const Real t2b = Real(-1.f / 6.0f) * hh * sfacb + cwb * (Real(-0.5f) + t2 * (1.0f / 12.0f));
  w1b = w1b + (2.0f * hh * hh * t2b) * w1;
  // p1 = p + h R v1
  ob.p = xb.p;
  V3 v1b = xb.v + h * mulT(R, xb.p);
  G3 Rbar = outer(h * xb.p, v1);
  // q1 = q + h qd1, qd1 = qd + h qdd
  Real qddb[12];
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    ob.q[j] = xb.q[j];
    const Real qd1b = xb.qd[j] + h * xb.q[j];
    ob.qd[j] = qd1b;
    qddb[j] = h * qd1b;
  }
  // v1 = v + h acc.v, w1 = w + h acc.w; acc.v = a_rel.v + R^T g
  ob.v = v1b;
  ob.w = w1b;
  const V3 avb = h * v1b, awb = h * w1b;
  Rbar = gadd(Rbar, outer(V3{Real(0.f), Real(0.f), Real(-1.f) * a.gravity}, avb));
  // ---- forward-dynamics adjoint: mu = M^-1 [awb, avb, qddb]
  Real rhs[18], mu[18];
  rhs[0] = awb.x; rhs[1] = awb.y; rhs[2] = awb.z; rhs[3] = avb.x; rhs[4] = avb.y; rhs[5] = avb.z;
#pragma unroll
  for (int j = 0; j < 12; ++j) rhs[6 + j] = qddb[j];
  minv_apply(lp, st, rhs, mu);
  // tau_joint = tau_m - fric tanh(qd / v0): + mu_joint . d tau_joint
  Real mu_neg[12];
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    const Real m = mu[6 + j];
    taub[j] += m;
    const Real th = spi_tanh(st.qd[j] * a.fric_vel_inv);
    pb.g[18 + j / 3] += Real(-1.f) * m * th;
    ob.qd[j] += Real(-1.f) * m * lp.fric[j / 3] * (1.0f - th * th) * a.fric_vel_inv;
    mu_neg[j] = Real(-1.f) * m;
  }
  // - (d ID / dx)^T mu: reverse RNEA seeded with -mu
  const M6 Fb_bar{V3{Real(-1.f) * mu[0], Real(-1.f) * mu[1], Real(-1.f) * mu[2]},
                  V3{Real(-1.f) * mu[3], Real(-1.f) * mu[4], Real(-1.f) * mu[5]}};
  const M6 v0{st.w, st.v};
  F6 v0b = fzero();
  leg_id_vjp<0>(a, st, R, v0, a_rel, qdd, Fb_bar, mu_neg, v0b, Rbar, ob);
  leg_id_vjp<1>(a, st, R, v0, a_rel, qdd, Fb_bar, mu_neg, v0b, Rbar, ob);
  leg_id_vjp<2>(a, st, R, v0, a_rel, qdd, Fb_bar, mu_neg, v0b, Rbar, ob);
  leg_id_vjp<3>(a, st, R, v0, a_rel, qdd, Fb_bar, mu_neg, v0b, Rbar, ob);
  // base body: f_b = I0 a0 + v0 x* I0 v0
  const BaseInertia& B = lp.base;
  RBT<Real> rb{B.m, B.hx, B.hy, B.hz, B.ixx, B.iyy, B.izz, B.ixy, B.ixz, B.iyz};
  const F6 h0 = rb_mul(rb, v0);
  v0b = fadd(v0b, fneg(fadd(crf(Fb_bar, h0), rb_mul(rb, crm(v0, Fb_bar)))));
  RBBar ib{Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f)};
  rb_bilinear_vjp(Fb_bar, a_rel, Real(1.f), ib);
  rb_bilinear_vjp(crm(v0, Fb_bar), v0, Real(-1.f), ib);
  base_inertia_vjp(ib, raw[0], raw[1], raw[2], raw[3], pb.g);
  ob.w = ob.w + v0b.n;
  ob.v = ob.v + v0b.f;
  // rotation adjoint -> quaternion
  quat_rot_vjp(qx, qy, qz, qw, Rbar, ob);
  xb = ob;
}

// ------------------------------------------------------------------ control law adjoint
SPI_HD void control_vjp(const SpiSimArgs& a, const LaneParams& lp, const Real* act, const State& st,
                        const Real* taub, StateBar& xb, ParBar& pb) {
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    const int leg = j / 3, grp = j % 3;
    Real x = fminf(fmaxf(act[j], -a.action_clip), a.action_clip);
    Real as = x * a.action_scale;
    if (grp == 0) as *= a.hip_scale;
    const Real e = as + a.q0[j] - st.q[j];
    const Real traw = lp.kp_s[leg] * a.kp[j] * e - lp.kd_s[leg] * a.kd[j] * st.qd[j];
    const Real t = fminf(fmaxf(traw, -a.effort[j]), a.effort[j]);
    Real tb = taub[j];
    if (a.motor_model == SPI_MOTOR_SCALAR) {
      pb.g[22] += tb * t;
      tb = tb * lp.motor[0];
    } else if (a.motor_model == SPI_MOTOR_VEC3) {
      pb.g[22 + grp] += tb * t;
      tb = tb * lp.motor[grp];
    } else if (a.motor_model == SPI_MOTOR_VEC3_TANH) {
      const Real A = lp.motor[grp];
      const Real B = spi_rcp(A);
      const Real th = spi_tanh(B * t);
      const Real d = 1.0f - th * th;
      pb.g[22 + grp] += tb * (th - t * B * d);
      tb = tb * d;
    }
    if (!(traw > -a.effort[j] && traw < a.effort[j])) tb = Real(0.f);
    pb.g[10 + leg] += tb * a.kp[j] * e;
    pb.g[14 + leg] += Real(-1.f) * tb * a.kd[j] * st.qd[j];
    xb.q[j] += Real(-1.f) * tb * lp.kp_s[leg] * a.kp[j];
    xb.qd[j] += Real(-1.f) * tb * lp.kd_s[leg] * a.kd[j];
  }
}

// ------------------------------------------------------------------ cost adjoint
// adds d step_cost / d st (scaled by cbar) to xb
template <class TgtT> SPI_HD void step_cost_vjp(const SpiCostArgs& c, const State& st, const TgtT* tg, Real cbar,
                                                StateBar& xb) {
  const G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  const V3 vw = mul(R, st.v), ww = mul(R, st.w);
  const V3 dp{st.p.x - tg[0], st.p.y - tg[1], st.p.z - tg[2]};
  const Real dq[4] = {st.qx - tg[3], st.qy - tg[4], st.qz - tg[5], st.qw - tg[6]};
  const V3 dv{vw.x - tg[7], vw.y - tg[8], vw.z - tg[9]};
  const V3 dw{ww.x - tg[10], ww.y - tg[11], ww.z - tg[12]};
  Real s[6];
  if (c.mode == SPI_COST_TRACKING) {
#pragma unroll
    for (int g = 0; g < 6; ++g) s[g] = 2.0f * c.w[g] * cbar;
  } else {
    Real e4(0.f);
#pragma unroll
    for (int j = 0; j < 12; ++j) { const Real d = st.q[j] - tg[13 + j]; e4 += d * d; }
    const Real e0 = dot(dp, dp), e1 = dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2] + dq[3] * dq[3];
    s[0] = e0 > 0.0f ? c.w[0] * cbar / sqrtf(e0) : Real(0.f);
    s[1] = e1 > 0.0f ? c.w[1] * cbar / sqrtf(e1) : Real(0.f);
    s[2] = Real(0.f); s[3] = Real(0.f);
    s[4] = e4 > 0.0f ? c.w[4] * cbar / sqrtf(e4) : Real(0.f);
    s[5] = Real(0.f);
  }
  xb.p = xb.p + s[0] * dp;
  xb.qx += s[1] * dq[0]; xb.qy += s[1] * dq[1]; xb.qz += s[1] * dq[2]; xb.qw += s[1] * dq[3];
  const V3 vwb = s[2] * dv, wwb = s[3] * dw;
  xb.v = xb.v + mulT(R, vwb);
  xb.w = xb.w + mulT(R, wwb);
  quat_rot_vjp(st.qx, st.qy, st.qz, st.qw, gadd(outer(vwb, st.v), outer(wwb, st.w)), xb);
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    xb.q[j] += s[4] * (st.q[j] - tg[13 + j]);
    xb.qd[j] += s[5] * (st.qd[j] - tg[25 + j]);
  }
}

// ------------------------------------------------------------------ one physics step
// Reverse of {control_torques; substeps x (FD, integrate)} from the physics step's start state
// st0; the sub-interval start states are recomputed into ss.
SPI_HD void physics_step_vjp(const SpiSimArgs& a, const LaneParams& lp, const Real* raw, const Real* act,
                             const State& st0, State* ss, StateBar& xb, ParBar& pb) {
  Real tau[12], taub[12];
  control_torques(a, lp, act, st0, tau);
  const Real h = a.dt / Real(a.substeps);
  ss[0] = st0;
  for (int k = 1; k < a.substeps; ++k) {
    ss[k] = ss[k - 1];
    M6 acc;
    Real qdd[12];
    forward_dynamics(a, lp, ss[k], tau, acc, qdd, nullptr);
    integrate(ss[k], acc, qdd, h);
  }
#pragma unroll
  for (int j = 0; j < 12; ++j) taub[j] = Real(0.f);
  for (int k = a.substeps - 1; k >= 0; --k) substep_vjp(a, lp, raw, ss[k], tau, h, xb, taub, pb);
  control_vjp(a, lp, act, st0, taub, xb, pb);
}

// ------------------------------------------------------------------ whole rollout
// Cost of one (candidate, segment) rollout and its gradient with respect to the 25 parameter
// rows, by reverse mode with checkpoints: the forward pass stores the state every `every`
// control steps (Store::put(slot, state)); the reverse pass recomputes each block of control
// steps from its checkpoint (slots nb .. nb + every - 1), then inside each control step the
// physics-step and sub-interval start states.  ps / ss: scratch for decimation / substeps
// states.  Returns false if the rollout left the finite / state_limit range.
template <class Store, class TgtT>
SPI_HD bool rollout_grad(const SpiSimArgs& a, const SpiCostArgs& cost, const LaneParams& lp, const Real* raw,
                         const TgtT* rec0, int H, int every, Store& store, State* ps, State* ss, Real& cost_out,
                         ParBar& pb) {
  const int nb = (H + every - 1) / every;
  State st;
  {
    Real r[13], q[12], qd[12];
    for (int i = 0; i < 13; ++i) r[i] = Real(rec0[i]);
    for (int j = 0; j < 12; ++j) { q[j] = Real(rec0[13 + j]); qd[j] = Real(rec0[25 + j]); }
    state_from_root(st, r, q, qd, 1);
  }
  Real acc(0.f);
  for (int t = 0; t < H; ++t) {
    if (t % every == 0) store.put(t / every, st);
    const TgtT* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
    Real act[12];
    for (int j = 0; j < 12; ++j) act[j] = Real(rec[j]);
    for (int d = 0; d < a.decimation; ++d) {
      Real tau[12];
      control_torques(a, lp, act, st, tau);
      physics_step(a, lp, st, tau, nullptr);
    }
    if (cost.mode == SPI_COST_TRACKING || t == H - 1) acc += step_cost(cost, st, rec + 12);
    const Real nrm = state_norm(st);
    if (!below_limit(nrm, a.state_limit)) return false;   // also for NaN and infinities
  }
  cost_out = acc;
  const State final_state = st;
  StateBar xb;
  zero_bar(xb);
  zero_par(pb);
  for (int b = nb - 1; b >= 0; --b) {
    const int t0 = b * every, t1 = (t0 + every < H) ? t0 + every : H;
    store.get(b, st);
    for (int t = t0; t < t1; ++t) {
      store.put(nb + (t - t0), st);
      if (t + 1 < t1) {
        const TgtT* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
        Real act[12];
        for (int j = 0; j < 12; ++j) act[j] = Real(rec[j]);
        for (int d = 0; d < a.decimation; ++d) {
          Real tau[12];
          control_torques(a, lp, act, st, tau);
          physics_step(a, lp, st, tau, nullptr);
        }
      }
    }
    for (int t = t1 - 1; t >= t0; --t) {
      const TgtT* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
      if (cost.mode == SPI_COST_TRACKING || t == H - 1) {
        State xe;
        if (t + 1 < t1) store.get(nb + (t + 1 - t0), xe);
        else if (b + 1 < nb) store.get(b + 1, xe);
        else xe = final_state;
        step_cost_vjp(cost, xe, rec + 12, Real(1.f), xb);
      }
      Real act[12];
      for (int j = 0; j < 12; ++j) act[j] = Real(rec[j]);
      store.get(nb + (t - t0), ps[0]);
      for (int d = 1; d < a.decimation; ++d) {
        ps[d] = ps[d - 1];
        Real tau[12];
        control_torques(a, lp, act, ps[d], tau);
        physics_step(a, lp, ps[d], tau, nullptr);
      }
      for (int d = a.decimation - 1; d >= 0; --d) physics_step_vjp(a, lp, raw, act, ps[d], ss, xb, pb);
    }
  }
  return true;
}

}  // namespace SPI_DYN_NS
