// Reverse mode of the leg-parallel Go2 step (spi_dyn_leg.h): four consecutive lanes hold one
// rollout, lane % 4 = leg.  The adjoint is the one of spi_adj.h, split the way the forward
// step is split:
//   * base adjoints (position, orientation, twist, base-inertia parameters) are computed
//     identically in the four lanes; every contribution a leg makes to them (the leg's share
//     of the base velocity adjoint, of the base rotation adjoint and of the calf position
//     adjoint) is summed over the four lanes with the same butterfly as the forward sums;
//   * each lane carries the adjoints of its own three joints and of its own leg's parameters
//     (Kp scale, Kd scale, Coulomb friction); its share of the three motor constants, which
//     the legs have in common, is summed over the lanes once at the end of the rollout.
// The articulated-body solve for mu = M^-1 abar runs as the forward dynamics do: each lane
// downdates its leg, the base articulated inertia is the 4-lane sum.  Same derivative as the
// one-rollout-per-lane adjoint up to the order of the legs' sums (float rounding).
#pragma once
#include "spi_adj.h"
#include "spi_dyn_leg.h"

namespace SPI_DYN_NS {

struct LegStateBar {
  V3 p;
  Real qx, qy, qz, qw;
  V3 w, v;
  Real q[3], qd[3];
};
__device__ __forceinline__ void zero_bar(LegStateBar& b) {
  b.p = v3zero(); b.w = v3zero(); b.v = v3zero();
  b.qx = b.qy = b.qz = b.qw = Real(0.f);
#pragma unroll
  for (int k = 0; k < 3; ++k) { b.q[k] = Real(0.f); b.qd[k] = Real(0.f); }
}

// One lane's parameter adjoint: base rows 0..9 (the same in the four lanes), this leg's
// Kp scale, Kd scale and friction, and this leg's share of the motor constants.
struct LegParBar {
  Real base[10];
  Real kp, kd, fric;
  Real motor[3];
};
__device__ __forceinline__ void zero_par(LegParBar& b) {
#pragma unroll
  for (int k = 0; k < 10; ++k) b.base[k] = Real(0.f);
  b.kp = b.kd = b.fric = Real(0.f);
#pragma unroll
  for (int g = 0; g < 3; ++g) b.motor[g] = Real(0.f);
}

__device__ __forceinline__ void sum4(G3& g) {
  g = G3{sum4(g.m00), sum4(g.m01), sum4(g.m02), sum4(g.m10), sum4(g.m11), sum4(g.m12),
         sum4(g.m20), sum4(g.m21), sum4(g.m22)};
}
__device__ __forceinline__ void sum4(F6& f) {
  sum4(f.n);
  sum4(f.f);
}

// ------------------------------------------------------------------ RNEA adjoint of this lane's leg
// leg_id_vjp<L> of spi_adj.h with the lane's constants (axes x, y, y for every leg).  The
// leg's contributions to the base velocity adjoint (v0b), the base rotation adjoint (Rb) and
// the base position adjoint (pb_sum) are returned for the caller to sum over the four lanes;
// the joint adjoints go to sb.q / sb.qd.
__device__ __forceinline__ void leg_id_vjp_lane(const SpiSimArgs& a, const LegK& K, const LegState& st, const G3& R,
                                                const M6& v0, const M6& a0, const Real* qdd, const M6& Fb_bar,
                                                const Real* mu_neg, F6& v0b, G3& Rb, V3& pb_sum, LegStateBar& sb) {
  Real c0, s0, c1, s1, c2, s2;
  spi_sincos(st.q[0], &s0, &c0);
  spi_sincos(st.q[1], &s1, &c1);
  spi_sincos(st.q[2], &s2, &c2);
  // ---- forward: velocities, accelerations, forces
  const M6 xv1 = xm<0>(v0, K.r0, c0, s0);
  M6 v1 = xv1; v1.w = v1.w + axis_motion<0>(st.qd[0]).w;
  const M6 xv2 = xm<1>(v1, K.r1, c1, s1);
  M6 v2 = xv2; v2.w = v2.w + axis_motion<1>(st.qd[1]).w;
  const M6 xv3 = xm<1>(v2, K.r2, c2, s2);
  M6 v3 = xv3; v3.w = v3.w + axis_motion<1>(st.qd[2]).w;
  const M6 xa1 = xm<0>(a0, K.r0, c0, s0);
  M6 a1 = madd(xa1, vel_product<0>(v1, st.qd[0])); a1.w = a1.w + axis_motion<0>(qdd[0]).w;
  const M6 xa2 = xm<1>(a1, K.r1, c1, s1);
  M6 a2 = madd(xa2, vel_product<1>(v2, st.qd[1])); a2.w = a2.w + axis_motion<1>(qdd[1]).w;
  const M6 xa3 = xm<1>(a2, K.r2, c2, s2);
  M6 a3 = madd(xa3, vel_product<1>(v3, st.qd[2])); a3.w = a3.w + axis_motion<1>(qdd[2]).w;
  const V3 pb_in = K.r1 + rot<1>(K.r2, c1, s1);
  const V3 pb = K.r0 + rot<0>(pb_in, c0, s0);
  const V3 pcw = st.p + mul(R, pb);
  const G3 R0 = axis_rot<0>(c0, s0), R1 = axis_rot<1>(c1, s1), R2 = axis_rot<1>(c2, s2);
  const G3 R12 = gmul(R1, R2), A = gmul(R0, R12);
  const G3 Rc = gmul(R, A);
  const F6 fext = foot_contact<0>(a, Rc, pcw, v3, nullptr);   // every leg has leg 0's foot sphere
  const F6 h3 = rb_mul(K.b2, v3), h2 = rb_mul(K.b1, v2), h1 = rb_mul(K.b0, v1);
  const F6 f3 = fadd(fadd(rb_mul(K.b2, a3), crf(v3, h3)), fneg(fext));
  const F6 f2 = fadd(rb_mul(K.b1, a2), crf(v2, h2));
  const F6 f1 = fadd(rb_mul(K.b0, a1), crf(v1, h1));
  const F6 F3 = f3;
  const F6 F2 = fadd(f2, xtf<1>(F3, K.r2, c2, s2));
  const F6 F1 = fadd(f1, xtf<1>(F2, K.r1, c1, s1));
  // ---- reverse phase 1: force adjoints from the base down to the calf
  M6 t = xm<0>(Fb_bar, K.r0, c0, s0);
  sb.q[0] += dot6(t, crf_axis<0>(F1));
  M6 F1b = t; F1b.w = F1b.w + Real(mu_neg[0]) * unit<0>();
  t = xm<1>(F1b, K.r1, c1, s1);
  sb.q[1] += dot6(t, crf_axis<1>(F2));
  M6 F2b = t; F2b.w = F2b.w + Real(mu_neg[1]) * unit<1>();
  t = xm<1>(F2b, K.r2, c2, s2);
  sb.q[2] += dot6(t, crf_axis<1>(F3));
  M6 F3b = t; F3b.w = F3b.w + Real(mu_neg[2]) * unit<1>();
  // ---- reverse phase 2: calf, thigh, hip
  F6 a3b = rb_mul(K.b2, F3b);
  F6 v3b = fneg(fadd(crf(F3b, h3), rb_mul(K.b2, crm(v3, F3b))));
  G3 Rcb = gzero();
  V3 pcwb = v3zero();
  foot_contact_vjp<0>(a, Rc, pcw, v3, M6{Real(-1.f) * F3b.w, Real(-1.f) * F3b.v}, Rcb, pcwb, v3b);
  F6 a2b = fadd(rb_mul(K.b1, F2b), xtf<1>(a3b, K.r2, c2, s2));
  sb.q[2] -= dot6(crm_axis<1>(xa3), a3b);
  v3b = fadd(v3b, crf(axis_motion<1>(st.qd[2]), a3b));
  sb.qd[2] += dot6(crm(v3, axis6<1>()), a3b);
  F6 v2b = fadd(fneg(fadd(crf(F2b, h2), rb_mul(K.b1, crm(v2, F2b)))), xtf<1>(v3b, K.r2, c2, s2));
  sb.qd[2] += comp(v3b.n, 1);
  sb.q[2] -= dot6(crm_axis<1>(xv3), v3b);
  F6 a1b = fadd(rb_mul(K.b0, F1b), xtf<1>(a2b, K.r1, c1, s1));
  sb.q[1] -= dot6(crm_axis<1>(xa2), a2b);
  v2b = fadd(v2b, crf(axis_motion<1>(st.qd[1]), a2b));
  sb.qd[1] += dot6(crm(v2, axis6<1>()), a2b);
  F6 v1b = fadd(fneg(fadd(crf(F1b, h1), rb_mul(K.b0, crm(v1, F1b)))), xtf<1>(v2b, K.r1, c1, s1));
  sb.qd[1] += comp(v2b.n, 1);
  sb.q[1] -= dot6(crm_axis<1>(xv2), v2b);
  sb.q[0] -= dot6(crm_axis<0>(xa1), a1b);
  v1b = fadd(v1b, crf(axis_motion<0>(st.qd[0]), a1b));
  sb.qd[0] += dot6(crm(v1, axis6<0>()), a1b);
  v0b = fadd(v0b, xtf<0>(v1b, K.r0, c0, s0));
  sb.qd[0] += comp(v1b.n, 0);
  sb.q[0] -= dot6(crm_axis<0>(xv1), v1b);
  // ---- calf pose: pcw = p + R pb(q0, q1), Rc = R R0 R1 R2
  pb_sum = pb_sum + pcwb;
  Rb = gadd(Rb, outer(pcwb, pb));
  const V3 pbb = mulT(R, pcwb);
  G3 R0b = outer(pbb, pb_in);
  const V3 pinb = mulT(R0, pbb);
  G3 R1b = outer(pinb, K.r2);
  Rb = gadd(Rb, gmul(Rcb, gtrans(A)));
  const G3 Ab = gmul(gtrans(R), Rcb);
  R0b = gadd(R0b, gmul(Ab, gtrans(R12)));
  const G3 Bb = gmul(gtrans(R0), Ab);
  R1b = gadd(R1b, gmul(Bb, gtrans(R2)));
  const G3 R2b = gmul(gtrans(R1), Bb);
  sb.q[0] += axis_rot_vjp<0>(R0b, c0, s0);
  sb.q[1] += axis_rot_vjp<1>(R1b, c1, s1);
  sb.q[2] += axis_rot_vjp<1>(R2b, c2, s2);
}

// ------------------------------------------------------------------ M^-1 (no bias terms)
// This lane's leg downdated with the right-hand side rhs (its three joints); returns the
// leg's articulated inertia and force seen by the base.
__device__ __forceinline__ void leg_minv_backward_lane(const LegK& K, const LegState& st, const Real* rhs, ABI& t0,
                                                       F6& pt0, JointAcc* ja) {
  Real c0, s0, c1, s1, c2, s2;
  spi_sincos(st.q[0], &s0, &c0);
  spi_sincos(st.q[1], &s1, &c1);
  spi_sincos(st.q[2], &s2, &c2);
  ja[0].cq = c0; ja[0].sq = s0; ja[1].cq = c1; ja[1].sq = s1; ja[2].cq = c2; ja[2].sq = s2;
  ja[0].c = mzero(); ja[1].c = mzero(); ja[2].c = mzero();
  ABI IA = rb_abi(K.b2);
  F6 pA = fzero();
  joint_downdate<1>(IA, pA, rhs[2], ja[2]);
  {
    ABI t = xtix<1>(IA, K.r2, c2, s2);
    F6 pt = xtf<1>(pA, K.r2, c2, s2);
    IA = rb_abi(K.b1);
    abi_add(IA, t);
    pA = pt;
  }
  joint_downdate<1>(IA, pA, rhs[1], ja[1]);
  {
    ABI t = xtix<1>(IA, K.r1, c1, s1);
    F6 pt = xtf<1>(pA, K.r1, c1, s1);
    IA = rb_abi(K.b0);
    abi_add(IA, t);
    pA = pt;
  }
  joint_downdate<0>(IA, pA, rhs[0], ja[0]);
  t0 = xtix<0>(IA, K.r0, c0, s0);
  pt0 = xtf<0>(pA, K.r0, c0, s0);
}
// (mu_base, mu_joint) = M^-1 (rhs_base = (w-part, v-part), rhs_joint = this lane's joints)
__device__ __forceinline__ void minv_apply_leg(const LegParams& lp, const LegK& K, const LegState& st, V3 rw, V3 rv,
                                               const Real* rhs_joint, M6& mu0, Real* mu_joint) {
  const BaseInertia& B = lp.base;
  RBT<Real> rb{B.m, B.hx, B.hy, B.hz, B.ixx, B.iyy, B.izz, B.ixy, B.ixz, B.iyz};
  JointAcc ja[3];
  ABI t;
  F6 pt;
  leg_minv_backward_lane(K, st, rhs_joint, t, pt, ja);
  sum4(t);
  sum4(pt);
  ABI IA0 = rb_abi(rb);
  abi_add(IA0, t);
  mu0 = solve_base(IA0, F6{rw - pt.n, rv - pt.f});
  leg_forward_lane(mu0, K, ja, mu_joint);
}

// ------------------------------------------------------------------ one sub-interval
// substep_vjp of spi_adj.h for this lane's leg: reverse of {FD; integrate(h)} at the
// sub-interval's start state st, whose forward dynamics (acc, qdd) the caller kept from the
// recompute of the sub-interval states.  xb: adjoint of the end state on entry, of st on exit.
__device__ __forceinline__ void substep_vjp_leg(const SpiSimArgs& a, const LegParams& lp, const LegK& K,
                                                const Real* raw, const LegState& st, const M6& acc, const Real* qdd,
                                                Real h, LegStateBar& xb, Real* taub, LegParBar& pb) {
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
  // ---- integrate adjoint (base: the same in the four lanes)
  const Real pr = nx * xb.qx + ny * xb.qy + nz * xb.qz + nw * xb.qw;
  const Real ubx = (xb.qx - nx * pr) * inv, uby = (xb.qy - ny * pr) * inv;
  const Real ubz = (xb.qz - nz * pr) * inv, ubw = (xb.qw - nw * pr) * inv;
  LegStateBar ob;
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
  ob.p = xb.p;
  V3 v1b = xb.v + h * mulT(R, xb.p);
  G3 Rbar = outer(h * xb.p, v1);
  // joints of this lane: q1 = q + h qd1, qd1 = qd + h qdd
  Real qddb[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ob.q[k] = xb.q[k];
    const Real qd1b = xb.qd[k] + h * xb.q[k];
    ob.qd[k] = qd1b;
    qddb[k] = h * qd1b;
  }
  ob.v = v1b;
  ob.w = w1b;
  const V3 avb = h * v1b, awb = h * w1b;
  Rbar = gadd(Rbar, outer(V3{Real(0.f), Real(0.f), Real(-1.f) * a.gravity}, avb));
  // ---- forward-dynamics adjoint: mu = M^-1 [awb, avb, qddb]
  M6 mu0;
  Real mu_j[3];
  minv_apply_leg(lp, K, st, awb, avb, qddb, mu0, mu_j);
  // tau_joint = tau_m - fric tanh(qd / v0)
  Real mu_neg[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const Real m = mu_j[k];
    taub[k] += m;
    const Real th = spi_tanh(st.qd[k] * a.fric_vel_inv);
    pb.fric += Real(-1.f) * m * th;
    ob.qd[k] += Real(-1.f) * m * lp.fric * (1.0f - th * th) * a.fric_vel_inv;
    mu_neg[k] = Real(-1.f) * m;
  }
  // - (d ID / dx)^T mu: reverse RNEA seeded with -mu; the legs' base contributions summed
  const M6 Fb_bar{Real(-1.f) * mu0.w, Real(-1.f) * mu0.v};
  const M6 v0{st.w, st.v};
  F6 v0b = fzero();
  G3 Rl = gzero();
  V3 pl = v3zero();
  leg_id_vjp_lane(a, K, st, R, v0, a_rel, qdd, Fb_bar, mu_neg, v0b, Rl, pl, ob);
  sum4(v0b);
  sum4(Rl);
  sum4(pl);
  ob.p = ob.p + pl;
  Rbar = gadd(Rbar, Rl);
  // base body: f_b = I0 a0 + v0 x* I0 v0
  const BaseInertia& B = lp.base;
  RBT<Real> rb{B.m, B.hx, B.hy, B.hz, B.ixx, B.iyy, B.izz, B.ixy, B.ixz, B.iyz};
  const F6 h0 = rb_mul(rb, v0);
  v0b = fadd(v0b, fneg(fadd(crf(Fb_bar, h0), rb_mul(rb, crm(v0, Fb_bar)))));
  RBBar ib{Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f), Real(0.f)};
  rb_bilinear_vjp(Fb_bar, a_rel, Real(1.f), ib);
  rb_bilinear_vjp(crm(v0, Fb_bar), v0, Real(-1.f), ib);
  base_inertia_vjp(ib, raw[0], raw[1], raw[2], raw[3], pb.base);
  ob.w = ob.w + v0b.n;
  ob.v = ob.v + v0b.f;
  quat_rot_vjp(qx, qy, qz, qw, Rbar, ob);
  xb = ob;
}

// ------------------------------------------------------------------ control law adjoint
// control_vjp of spi_adj.h for this lane's three joints; kp, kd, q0, effort: this leg's
// entries of the SpiSimArgs arrays.
__device__ __forceinline__ void control_vjp_leg(const SpiSimArgs& a, const LegParams& lp, const float* kp,
                                                const float* kd, const float* q0, const float* effort,
                                                const Real* act, const LegState& st, const Real* taub,
                                                LegStateBar& xb, LegParBar& pb) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Real x = fminf(fmaxf(act[k], -a.action_clip), a.action_clip);
    Real as = x * a.action_scale;
    if (k == 0) as *= a.hip_scale;
    const Real e = as + q0[k] - st.q[k];
    const Real traw = lp.kp_s * kp[k] * e - lp.kd_s * kd[k] * st.qd[k];
    const Real t = fminf(fmaxf(traw, -effort[k]), effort[k]);
    Real tb = taub[k];
    if (a.motor_model == SPI_MOTOR_SCALAR) {
      pb.motor[0] += tb * t;
      tb = tb * lp.motor[0];
    } else if (a.motor_model == SPI_MOTOR_VEC3) {
      pb.motor[k] += tb * t;
      tb = tb * lp.motor[k];
    } else if (a.motor_model == SPI_MOTOR_VEC3_TANH) {
      const Real A = lp.motor[k];
      const Real Bi = spi_rcp(A);
      const Real th = spi_tanh(Bi * t);
      const Real d = 1.0f - th * th;
      pb.motor[k] += tb * (th - t * Bi * d);
      tb = tb * d;
    }
    if (!(traw > -effort[k] && traw < effort[k])) tb = Real(0.f);
    pb.kp += tb * kp[k] * e;
    pb.kd += Real(-1.f) * tb * kd[k] * st.qd[k];
    xb.q[k] += Real(-1.f) * tb * lp.kp_s * kp[k];
    xb.qd[k] += Real(-1.f) * tb * lp.kd_s * kd[k];
  }
}

// ------------------------------------------------------------------ cost adjoint
// step_cost_vjp of spi_adj.h: base terms in every lane, this lane's joint terms.
template <class TgtT>
__device__ __forceinline__ void step_cost_vjp_leg(const SpiCostArgs& c, const LegState& st, const TgtT* tg, int l,
                                                  Real cbar, LegStateBar& xb) {
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
    for (int k = 0; k < 3; ++k) { const Real d = st.q[k] - tg[13 + 3 * l + k]; e4 += d * d; }
    e4 = sum4(e4);
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
  for (int k = 0; k < 3; ++k) {
    xb.q[k] += s[4] * (st.q[k] - tg[13 + 3 * l + k]);
    xb.qd[k] += s[5] * (st.qd[k] - tg[25 + 3 * l + k]);
  }
}

// This leg's constants of the SpiSimArgs arrays and parameters (kp, kd, q0, effort).
struct LegGains {
  float kp[3], kd[3], q0[3], effort[3];
};

// ------------------------------------------------------------------ one physics step
// Sub-interval start states and their forward dynamics, recomputed once into ss / acc / qdd
// (private memory, indexed by sub-interval), then the sub-interval adjoints backwards.
struct SubAcc {
  M6 acc;
  Real qdd[3];
};
__device__ __forceinline__ void physics_step_vjp_leg(const SpiSimArgs& a, const LegParams& lp, const LegK& K,
                                                     const LegGains& G, const Real* raw, const Real* act,
                                                     const LegState& st0, LegState* ss, SubAcc* sa, LegStateBar& xb,
                                                     LegParBar& pb) {
  Real tau[3], taub[3];
  control_torques_leg(a, lp, G.kp, G.kd, G.q0, G.effort, act, st0.q, st0.qd, tau);
  const Real h = a.dt / Real(a.substeps);
  LegState st = st0;
  for (int k = 0; k < a.substeps; ++k) {
    ss[k] = st;
    M6 acc;
    Real qdd[3];
    forward_dynamics_leg(a, lp, K, st, tau, acc, qdd);
    sa[k].acc = acc;
#pragma unroll
    for (int j = 0; j < 3; ++j) sa[k].qdd[j] = qdd[j];
    if (k + 1 < a.substeps) integrate_leg(st, acc, qdd, h);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) taub[k] = Real(0.f);
  for (int k = a.substeps - 1; k >= 0; --k)
    substep_vjp_leg(a, lp, K, raw, ss[k], sa[k].acc, sa[k].qdd, h, xb, taub, pb);
  control_vjp_leg(a, lp, G.kp, G.kd, G.q0, G.effort, act, st0, taub, xb, pb);
}

// ------------------------------------------------------------------ whole rollout
// rollout_grad of spi_adj.h, leg-parallel: Store::put / get hold this lane's LegState.  The
// forward pass stores the state every `every` control steps (slots 0 .. nb - 1); the reverse
// pass recomputes each block from its checkpoint and stores the start state of every physics
// step of the block (slot nb + (t - t0) * decimation + d), so no physics step is recomputed
// twice.  The four lanes of a rollout leave together (the divergence test is the same in all
// four).
template <class Store, class TgtT>
__device__ bool rollout_grad_leg(const SpiSimArgs& a, const SpiCostArgs& cost, const LegParams& lp, const LegK& K,
                                 const LegGains& G, const Real* raw, const TgtT* rec0, int H, int every, int l,
                                 Store& store, LegState* ss, SubAcc* sa, Real& cost_out, LegParBar& pb) {
  const int nb = (H + every - 1) / every;
  const int D = a.decimation;
  LegState st;
  leg_state_from_root(st, rec0, l);
  Real acc(0.f);
  for (int t = 0; t < H; ++t) {
    if (t % every == 0) store.put(t / every, st);
    const TgtT* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
    Real act[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) act[k] = Real(rec[3 * l + k]);
    for (int d = 0; d < D; ++d) {
      Real tau[3];
      control_torques_leg(a, lp, G.kp, G.kd, G.q0, G.effort, act, st.q, st.qd, tau);
      physics_step_leg(a, lp, K, st, tau);
    }
    if (cost.mode == SPI_COST_TRACKING || t == H - 1) acc += step_cost_leg(cost, st, rec + 12, l);
    const Real nrm = state_norm_leg(st);
    if (!below_limit(nrm, a.state_limit)) return false;   // also for NaN and infinities
  }
  cost_out = acc;
  const LegState final_state = st;
  LegStateBar xb;
  zero_bar(xb);
  zero_par(pb);
  for (int b = nb - 1; b >= 0; --b) {
    const int t0 = b * every, t1 = (t0 + every < H) ? t0 + every : H;
    store.get(b, st);
    for (int t = t0; t < t1; ++t) {
      const TgtT* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
      Real act[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) act[k] = Real(rec[3 * l + k]);
      for (int d = 0; d < D; ++d) {
        store.put(nb + (t - t0) * D + d, st);
        if (t + 1 == t1 && d + 1 == D) break;       // the block's end state is not needed here
        Real tau[3];
        control_torques_leg(a, lp, G.kp, G.kd, G.q0, G.effort, act, st.q, st.qd, tau);
        physics_step_leg(a, lp, K, st, tau);
      }
    }
    for (int t = t1 - 1; t >= t0; --t) {
      const TgtT* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
      if (cost.mode == SPI_COST_TRACKING || t == H - 1) {
        LegState xe;
        if (t + 1 < t1) store.get(nb + (t + 1 - t0) * D, xe);
        else if (b + 1 < nb) store.get(b + 1, xe);
        else xe = final_state;
        step_cost_vjp_leg(cost, xe, rec + 12, l, Real(1.f), xb);
      }
      Real act[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) act[k] = Real(rec[3 * l + k]);
      for (int d = D - 1; d >= 0; --d) {
        LegState ps;
        store.get(nb + (t - t0) * D + d, ps);
        physics_step_vjp_leg(a, lp, K, G, raw, act, ps, ss, sa, xb, pb);
      }
    }
  }
  return true;
}

}  // namespace SPI_DYN_NS
