// Leg-parallel form of the Go2 step of spi_dyn.h for small batches (gfx950): four consecutive
// lanes hold one env, lane % 4 = leg.  Each lane runs its own leg's articulated-body passes
// (the bulk of the work) with the leg's constants in registers; the base quantities are
// computed identically in all four lanes and the legs' articulated inertias are summed with a
// 4-lane butterfly.  A wavefront then carries 16 envs instead of 64, with about half the
// instructions per lane on the critical path: for batches that fill only a fraction of the
// SIMDs, the control step's latency drops; for large batches the one-env-per-lane form is
// cheaper in total.  Same model, contact law and integrator as spi_dyn.h; only the order of
// the four legs' sums into the base differs (float rounding).
#pragma once
#include "spi_dyn.h"
#include "spi_quad.h"

namespace SPI_DYN_NS {

// structural zeros of the folded tree, shared by all four legs: hip origin (x, y, 0), thigh
// origin (0, y, 0), calf origin (0, 0, z); axes x, y, y; the same foot sphere on every calf
constexpr bool legs_share_structure() {
  for (int l = 0; l < 4; ++l) {
    if (SPI_JOINT_AXIS[3 * l] != 0 || SPI_JOINT_AXIS[3 * l + 1] != 1 || SPI_JOINT_AXIS[3 * l + 2] != 1) return false;
    const double* o = &SPI_JOINT_ORIGIN[9 * l];
    if (o[2] != 0.0 || o[3] != 0.0 || o[5] != 0.0 || o[6] != 0.0 || o[7] != 0.0) return false;
    for (int k = 0; k < 3; ++k)
      if (SPI_FOOT_POS[3 * l + k] != SPI_FOOT_POS[k]) return false;
    if (SPI_FOOT_RADIUS[l] != SPI_FOOT_RADIUS[0]) return false;
  }
  return true;
}
static_assert(legs_share_structure(), "leg-parallel step assumes the Go2 leg structure");

// Per-leg constants in constant memory: link inertias and the non-zero origin components.
struct LegTable {
  RB b[12];
  float r0x[4], r0y[4], r1y[4], r2z[4];
};
constexpr LegTable leg_table() {
  LegTable t{};
  for (int b = 0; b < 12; ++b) t.b[b] = leg_rb(b + 1);
  for (int l = 0; l < 4; ++l) {
    t.r0x[l] = float(SPI_JOINT_ORIGIN[9 * l]);
    t.r0y[l] = float(SPI_JOINT_ORIGIN[9 * l + 1]);
    t.r1y[l] = float(SPI_JOINT_ORIGIN[9 * l + 4]);
    t.r2z[l] = float(SPI_JOINT_ORIGIN[9 * l + 8]);
  }
  return t;
}
static __constant__ LegTable kLegTab = leg_table();

// This lane's leg: joint origins (runtime values, literal structural zeros), link inertias.
struct LegK {
  V3 r0, r1, r2;
  RBT<Real> b0, b1, b2;
};
SPI_HD RBT<Real> rb_real(const RB& r) {
  return RBT<Real>{r.m, r.hx, r.hy, r.hz, r.ixx, r.iyy, r.izz, r.ixy, r.ixz, r.iyz};
}
__device__ __forceinline__ LegK leg_constants(int l) {
  const Real z(0.f);
  LegK k;
  k.r0 = V3{kLegTab.r0x[l], kLegTab.r0y[l], z};
  k.r1 = V3{z, kLegTab.r1y[l], z};
  k.r2 = V3{z, z, kLegTab.r2z[l]};
  k.b0 = rb_real(kLegTab.b[3 * l]);
  k.b1 = rb_real(kLegTab.b[3 * l + 1]);
  k.b2 = rb_real(kLegTab.b[3 * l + 2]);
  return k;
}

// Base state (identical in the four lanes) and this lane's three joints.
struct LegState {
  V3 p;
  Real qx, qy, qz, qw;
  V3 w, v;
  Real q[3], qd[3];
};

// Parameters seen by one lane: the base inertia and this leg's scales / friction.
struct LegParams {
  BaseInertia base;
  Real kp_s, kd_s, fric, motor[3];
};

// Sum over the four lanes of an env (lanes 4e .. 4e + 3 of one wavefront, one DPP quad); a
// scalar type with its own lanes (tangent arithmetic) provides shfl_xor_lane<M> for itself.
// The butterfly leaves the same bits in the four lanes (float addition commutes); its operands
// and partial sums are pinned so that reassociation cannot regroup them with the code around.
template <int M> __device__ __forceinline__ float shfl_xor_lane(float x) { return spi_quad_xor<M>(x); }
__device__ __forceinline__ float quad_pin(float x) { return spi_quad_pin(x); }
template <class T> __device__ __forceinline__ T sum4(T x) {
  x = quad_pin(x);
  x = quad_pin(x + shfl_xor_lane<1>(x));
  return quad_pin(x + shfl_xor_lane<2>(x));
}
__device__ __forceinline__ void sum4(V3& v) { v = V3{sum4(v.x), sum4(v.y), sum4(v.z)}; }
__device__ __forceinline__ void sum4(ABI& I) {
  I.A = S3{sum4(I.A.xx), sum4(I.A.yy), sum4(I.A.zz), sum4(I.A.xy), sum4(I.A.xz), sum4(I.A.yz)};
  I.B = G3{sum4(I.B.m00), sum4(I.B.m01), sum4(I.B.m02), sum4(I.B.m10), sum4(I.B.m11), sum4(I.B.m12),
           sum4(I.B.m20), sum4(I.B.m21), sum4(I.B.m22)};
  I.C = S3{sum4(I.C.xx), sum4(I.C.yy), sum4(I.C.zz), sum4(I.C.xy), sum4(I.C.xz), sum4(I.C.yz)};
}

// PD + motor model for this lane's three joints (control_torques of spi_dyn.h); kp, kd, q0,
// effort: this leg's entries of the SpiSimArgs arrays.
__device__ __forceinline__ void control_torques_leg(const SpiSimArgs& a, const LegParams& lp, const float* kp,
                                                    const float* kd, const float* q0, const float* effort,
                                                    const Real* act, const Real* q, const Real* qd, Real* tau) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Real x = fminf(fmaxf(act[k], -a.action_clip), a.action_clip);
    Real as = x * a.action_scale;
    if (k == 0) as *= a.hip_scale;
    Real t = lp.kp_s * kp[k] * (as + q0[k] - q[k]) - lp.kd_s * kd[k] * qd[k];
    t = fminf(fmaxf(t, -effort[k]), effort[k]);
    if (a.motor_model == SPI_MOTOR_SCALAR) {
      t = t * lp.motor[0];
    } else if (a.motor_model == SPI_MOTOR_VEC3) {
      t = t * lp.motor[k];
    } else if (a.motor_model == SPI_MOTOR_VEC3_TANH) {
      Real A = lp.motor[k];
      Real B = spi_rcp(A);
      t = A * spi_tanh(B * t);
    }
    tau[k] = t;
  }
}

// leg_backward<L> of spi_dyn.h with this lane's constants; returns the leg's contribution to
// the base articulated inertia (t) and bias (pt) instead of adding it.
__device__ __forceinline__ void leg_backward_lane(const SpiSimArgs& a, const LegK& K, const LegState& st, const G3& R,
                                                  const M6& v0, const Real* tau, ABI& t0, F6& pt0, JointAcc* ja) {
  Real c0, s0, c1, s1, c2, s2;
  spi_sincos(st.q[0], &s0, &c0);
  spi_sincos(st.q[1], &s1, &c1);
  spi_sincos(st.q[2], &s2, &c2);
  ja[0].cq = c0; ja[0].sq = s0;
  ja[1].cq = c1; ja[1].sq = s1;
  ja[2].cq = c2; ja[2].sq = s2;
  M6 v1 = xm<0>(v0, K.r0, c0, s0);
  v1.w = v1.w + axis_motion<0>(st.qd[0]).w;
  M6 v2 = xm<1>(v1, K.r1, c1, s1);
  v2.w = v2.w + axis_motion<1>(st.qd[1]).w;
  M6 v3 = xm<1>(v2, K.r2, c2, s2);
  v3.w = v3.w + axis_motion<1>(st.qd[2]).w;
  ja[0].c = vel_product<0>(v1, st.qd[0]);
  ja[1].c = vel_product<1>(v2, st.qd[1]);
  ja[2].c = vel_product<1>(v3, st.qd[2]);
  V3 pb = K.r0 + rot<0>(K.r1 + rot<1>(K.r2, c1, s1), c0, s0);
  V3 pcw = st.p + mul(R, pb);
  G3 Rc;
  {
    V3 x0 = rotT<1>(rotT<1>(rotT<0>(V3{R.m00, R.m01, R.m02}, c0, s0), c1, s1), c2, s2);
    V3 x1 = rotT<1>(rotT<1>(rotT<0>(V3{R.m10, R.m11, R.m12}, c0, s0), c1, s1), c2, s2);
    V3 x2 = rotT<1>(rotT<1>(rotT<0>(V3{R.m20, R.m21, R.m22}, c0, s0), c1, s1), c2, s2);
    Rc = G3{x0.x, x0.y, x0.z, x1.x, x1.y, x1.z, x2.x, x2.y, x2.z};
  }
  F6 fext = foot_contact<0>(a, Rc, pcw, v3, nullptr);  // every leg has leg 0's foot sphere
  ABI IA = rb_abi(K.b2);
  F6 pA = crf(v3, rb_mul(K.b2, v3));
  pA = F6{pA.n - fext.n, pA.f - fext.f};
  joint_downdate<1>(IA, pA, tau[2], ja[2]);
  {
    ABI t = xtix<1>(IA, K.r2, c2, s2);
    F6 pt = xtf<1>(pA, K.r2, c2, s2);
    IA = rb_abi(K.b1);
    abi_add(IA, t);
    F6 pb1 = crf(v2, rb_mul(K.b1, v2));
    pA = F6{pb1.n + pt.n, pb1.f + pt.f};
  }
  joint_downdate<1>(IA, pA, tau[1], ja[1]);
  {
    ABI t = xtix<1>(IA, K.r1, c1, s1);
    F6 pt = xtf<1>(pA, K.r1, c1, s1);
    IA = rb_abi(K.b0);
    abi_add(IA, t);
    F6 pb0 = crf(v1, rb_mul(K.b0, v1));
    pA = F6{pb0.n + pt.n, pb0.f + pt.f};
  }
  joint_downdate<0>(IA, pA, tau[0], ja[0]);
  t0 = xtix<0>(IA, K.r0, c0, s0);
  pt0 = xtf<0>(pA, K.r0, c0, s0);
}

__device__ __forceinline__ void leg_forward_lane(const M6& a0, const LegK& K, const JointAcc* ja, Real* qdd) {
  M6 a = xm<0>(a0, K.r0, ja[0].cq, ja[0].sq);
  a = M6{a.w + ja[0].c.w, a.v + ja[0].c.v};
  Real x = ja[0].dinv * (ja[0].u - dot(ja[0].U.w, a.w) - dot(ja[0].U.v, a.v));
  qdd[0] = x;
  a.w = a.w + axis_motion<0>(x).w;
  a = xm<1>(a, K.r1, ja[1].cq, ja[1].sq);
  a = M6{a.w + ja[1].c.w, a.v + ja[1].c.v};
  x = ja[1].dinv * (ja[1].u - dot(ja[1].U.w, a.w) - dot(ja[1].U.v, a.v));
  qdd[1] = x;
  a.w = a.w + axis_motion<1>(x).w;
  a = xm<1>(a, K.r2, ja[2].cq, ja[2].sq);
  a = M6{a.w + ja[2].c.w, a.v + ja[2].c.v};
  x = ja[2].dinv * (ja[2].u - dot(ja[2].U.w, a.w) - dot(ja[2].U.v, a.v));
  qdd[2] = x;
}

// forward_dynamics of spi_dyn.h, leg-parallel.  All four lanes of an env must be active.
__device__ __forceinline__ void forward_dynamics_leg(const SpiSimArgs& a, const LegParams& lp, const LegK& K,
                                                     const LegState& st, const Real* tau_motor, M6& acc0, Real* qdd) {
  G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  M6 v0{st.w, st.v};
  Real tau[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) tau[k] = tau_motor[k] - lp.fric * spi_tanh(st.qd[k] * a.fric_vel_inv);
  const BaseInertia& B = lp.base;
  RBT<Real> rb{B.m, B.hx, B.hy, B.hz, B.ixx, B.iyy, B.izz, B.ixy, B.ixz, B.iyz};
  JointAcc ja[3];
  ABI t;
  F6 pt;
  leg_backward_lane(a, K, st, R, v0, tau, t, pt, ja);
  sum4(t);
  sum4(pt.n);
  sum4(pt.f);
  ABI IA0 = rb_abi(rb);
  abi_add(IA0, t);
  const F6 pv = crf(v0, rb_mul(rb, v0));
  const F6 p0{pv.n + pt.n, pv.f + pt.f};
  M6 a0 = solve_base(IA0, F6{-1.0f * p0.n, -1.0f * p0.f});
  leg_forward_lane(a0, K, ja, qdd);
  Real g = a.gravity;
  acc0 = M6{a0.w, a0.v + V3{-g * R.m20, -g * R.m21, -g * R.m22}};
}

// integrate of spi_dyn.h: the base part runs identically in the four lanes.
__device__ __forceinline__ void integrate_leg(LegState& st, const M6& acc0, const Real* qdd, Real h) {
  st.w = st.w + h * acc0.w;
  st.v = st.v + h * acc0.v;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    st.qd[k] += h * qdd[k];
    st.q[k] += h * st.qd[k];
  }
  G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  st.p = st.p + h * mul(R, st.v);
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

__device__ __forceinline__ void physics_step_leg(const SpiSimArgs& a, const LegParams& lp, const LegK& K, LegState& st,
                                                 const Real* tau_motor) {
  Real h = a.dt / Real(a.substeps);
  for (int k = 0; k < a.substeps; ++k) {
    M6 acc0;
    Real qdd[3];
    forward_dynamics_leg(a, lp, K, st, tau_motor, acc0, qdd);
    integrate_leg(st, acc0, qdd, h);
  }
}

// state_norm of spi_dyn.h (base velocity + all twelve joint speeds), the same in all 4 lanes.
__device__ __forceinline__ Real state_norm_leg(const LegState& st) {
  Real s = fabsf(st.qd[0]) + fabsf(st.qd[1]) + fabsf(st.qd[2]);
  s = sum4(s);
  return s + fabsf(st.w.x) + fabsf(st.w.y) + fabsf(st.w.z) + fabsf(st.v.x) + fabsf(st.v.y) + fabsf(st.v.z);
}

// step_cost of spi_dyn.h: base terms in every lane, joint terms over the four lanes.
template <class TgtT>
__device__ __forceinline__ Real step_cost_leg(const SpiCostArgs& c, const LegState& st, const TgtT* tg, int l) {
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
  for (int k = 0; k < 3; ++k) {
    d = st.q[k] - tg[13 + 3 * l + k]; e4 += d * d;
    d = st.qd[k] - tg[25 + 3 * l + k]; e5 += d * d;
  }
  e4 = sum4(e4);
  e5 = sum4(e5);
  if (c.mode == SPI_COST_TRACKING)
    return c.w[0] * e0 + c.w[1] * e1 + c.w[2] * e2 + c.w[3] * e3 + c.w[4] * e4 + c.w[5] * e5;
  return c.w[0] * sqrtf(e0) + c.w[1] * sqrtf(e1) + c.w[4] * sqrtf(e4);
}

// Root-layout (IsaacGym) state of env row `src` into this lane's LegState.
__device__ __forceinline__ void leg_state_from_root(LegState& st, const float* src, int l) {
  st.p = V3{src[0], src[1], src[2]};
  st.qx = src[3]; st.qy = src[4]; st.qz = src[5]; st.qw = src[6];
  G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  st.v = mulT(R, V3{src[7], src[8], src[9]});
  st.w = mulT(R, V3{src[10], src[11], src[12]});
#pragma unroll
  for (int k = 0; k < 3; ++k) { st.q[k] = src[13 + 3 * l + k]; st.qd[k] = src[25 + 3 * l + k]; }
}
// Root row (IsaacGym layout: world-axis velocities) of the base state.
template <class OutT> __device__ __forceinline__ void leg_root(const LegState& st, OutT* dst) {
  G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  V3 vw = mul(R, st.v), ww = mul(R, st.w);
  dst[0] = st.p.x; dst[1] = st.p.y; dst[2] = st.p.z;
  dst[3] = st.qx; dst[4] = st.qy; dst[5] = st.qz; dst[6] = st.qw;
  dst[7] = vw.x; dst[8] = vw.y; dst[9] = vw.z;
  dst[10] = ww.x; dst[11] = ww.y; dst[12] = ww.z;
}
// Lane 0 of the env writes the root, every lane its three joints.
template <class OutT> __device__ __forceinline__ void leg_state_to_root(const LegState& st, OutT* dst, int l) {
  if (l == 0) leg_root(st, dst);
#pragma unroll
  for (int k = 0; k < 3; ++k) { dst[13 + 3 * l + k] = st.q[k]; dst[25 + 3 * l + k] = st.qd[k]; }
}

}  // namespace SPI_DYN_NS
