// ORACLE — test infrastructure only.  Nothing in the product links or calls this;
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the checker (or as the timed CPU baseline).
//
// CPU restatement of the SPI-Active sysID hot path for the Go2:
//   * PD law + action clip + motor models   (reference: legged_robot_base.py:185-198, 529-560;
//     go2_omni.py:423-465; active_sysid_openloop.py:174-187, 356-400)
//   * rigid-body dynamics of the URDF tree   (replaces PhysX behind isaacgym.py:598-626;
//     the reference's physics is a closed binary, so this is new numerics:
//     Featherstone ABA on the UNFOLDED URDF tree with fixed joints, dense 6x6 spatial
//     algebra; RNEA and CRBA on the same tree as independent cross-checks)
//   * penalty ground contact on the foot spheres (go2.urdf:360-365 sphere r = 0.022)
//   * semi-implicit Euler with quaternion exponential update
//   * tracking cost (scripts/eval.py:287-310, mass_landscape.py:32-36, 161-167)
//   * finite-difference FIM trace (active_sysid_openloop.py:402-426)
//   * CEM with Philox4x32-10 + Box-Muller (Salmon et al., SC'11) — replaces the
//     reference's Optuna samplers (algo/active_sysid.yaml:20, mass_opt.py:216).
//
// Templated on the scalar so the same code runs in double (parity), float
// (CPU baseline timing) and an op-counting scalar (FLOP census).
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

// SPIO_MUTATION plants one known mistake in the time stepping, so that the closed-form
// integrator tests can show that each of them fails (built only by `make mutants` into
// lib/libspi_oracle_mut<k>.so; no parity test loads these):
//   1  explicit Euler: positions advanced with the velocities from before the update
//   2  world-frame angular velocity in the right-multiplied quaternion increment
//   3  body-frame angular velocity in a left-multiplied (world-side) increment
//   4  sub-interval length dt instead of dt / substeps
//   5  torque computed once per control step instead of once per physics step
#ifndef SPIO_MUTATION
#define SPIO_MUTATION 0
#endif

namespace spio {

using std::size_t;

// ---------------------------------------------------------------- scalar helpers
template <class T> inline T s_sqrt(T x) { using std::sqrt; return sqrt(x); }
template <class T> inline T s_sin(T x) { using std::sin; return sin(x); }
template <class T> inline T s_cos(T x) { using std::cos; return cos(x); }
template <class T> inline T s_tanh(T x) { using std::tanh; return tanh(x); }
template <class T> inline bool s_pos(T x) { return x > T(0); }

template <class T> using V3 = std::array<T, 3>;
template <class T> using V6 = std::array<T, 6>;
template <class T> using M3 = std::array<std::array<T, 3>, 3>;
template <class T> using M6 = std::array<std::array<T, 6>, 6>;

template <class T> V3<T> cross(const V3<T>& a, const V3<T>& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
template <class T> T dot3(const V3<T>& a, const V3<T>& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <class T> V3<T> mv3(const M3<T>& m, const V3<T>& v) {
  V3<T> r;
  for (int i = 0; i < 3; ++i) r[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
  return r;
}
template <class T> V3<T> mtv3(const M3<T>& m, const V3<T>& v) {  // m^T v
  V3<T> r;
  for (int i = 0; i < 3; ++i) r[i] = m[0][i] * v[0] + m[1][i] * v[1] + m[2][i] * v[2];
  return r;
}
template <class T> M3<T> mm3(const M3<T>& a, const M3<T>& b) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
  return r;
}
template <class T> M3<T> tr3(const M3<T>& a) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = a[j][i];
  return r;
}
template <class T> M3<T> eye3() {
  M3<T> r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = T(i == j ? 1 : 0);
  return r;
}
template <class T> M3<T> skew(const V3<T>& v) {
  M3<T> r;
  r[0] = {T(0), -v[2], v[1]};
  r[1] = {v[2], T(0), -v[0]};
  r[2] = {-v[1], v[0], T(0)};
  return r;
}
// URDF rpy: R = Rz(y) Ry(p) Rx(r)
template <class T> M3<T> rpy_matrix(double r, double p, double y) {
  double cr = std::cos(r), sr = std::sin(r), cp = std::cos(p), sp = std::sin(p), cy = std::cos(y), sy = std::sin(y);
  M3<T> m;
  m[0] = {T(cy * cp), T(cy * sp * sr - sy * cr), T(cy * sp * cr + sy * sr)};
  m[1] = {T(sy * cp), T(sy * sp * sr + cy * cr), T(sy * sp * cr - cy * sr)};
  m[2] = {T(-sp), T(cp * sr), T(cp * cr)};
  return m;
}
// Rotation matrix (body->world) of a unit quaternion stored x, y, z, w
template <class T> M3<T> quat_matrix(const T* q) {
  T x = q[0], y = q[1], z = q[2], w = q[3];
  M3<T> m;
  m[0] = {T(1) - T(2) * (y * y + z * z), T(2) * (x * y - z * w), T(2) * (x * z + y * w)};
  m[1] = {T(2) * (x * y + z * w), T(1) - T(2) * (x * x + z * z), T(2) * (y * z - x * w)};
  m[2] = {T(2) * (x * z - y * w), T(2) * (y * z + x * w), T(1) - T(2) * (x * x + y * y)};
  return m;
}
// Hamilton product, x y z w storage
template <class T> void quat_mul(const T* a, const T* b, T* out) {
  T ax = a[0], ay = a[1], az = a[2], aw = a[3];
  T bx = b[0], by = b[1], bz = b[2], bw = b[3];
  out[0] = aw * bx + ax * bw + ay * bz - az * by;
  out[1] = aw * by - ax * bz + ay * bw + az * bx;
  out[2] = aw * bz + ax * by - ay * bx + az * bw;
  out[3] = aw * bw - ax * bx - ay * by - az * bz;
}

// ---------------------------------------------------------------- spatial algebra (dense)
// Motion vectors [w; v], force vectors [n; f] (angular first).
template <class T> V6<T> crm(const V6<T>& a, const V6<T>& b) {  // a x b (motion)
  V3<T> aw{a[0], a[1], a[2]}, av{a[3], a[4], a[5]}, bw{b[0], b[1], b[2]}, bv{b[3], b[4], b[5]};
  V3<T> r1 = cross(aw, bw), r2 = cross(aw, bv), r3 = cross(av, bw);
  return {r1[0], r1[1], r1[2], r2[0] + r3[0], r2[1] + r3[1], r2[2] + r3[2]};
}
template <class T> V6<T> crf(const V6<T>& a, const V6<T>& f) {  // a x* f (force)
  V3<T> aw{a[0], a[1], a[2]}, av{a[3], a[4], a[5]}, fn{f[0], f[1], f[2]}, ff{f[3], f[4], f[5]};
  V3<T> r1 = cross(aw, fn), r2 = cross(av, ff), r3 = cross(aw, ff);
  return {r1[0] + r2[0], r1[1] + r2[1], r1[2] + r2[2], r3[0], r3[1], r3[2]};
}
template <class T> V6<T> m6v(const M6<T>& m, const V6<T>& v) {
  V6<T> r;
  for (int i = 0; i < 6; ++i) {
    T s = T(0);
    for (int j = 0; j < 6; ++j) s += m[i][j] * v[j];
    r[i] = s;
  }
  return r;
}
template <class T> V6<T> m6tv(const M6<T>& m, const V6<T>& v) {
  V6<T> r;
  for (int i = 0; i < 6; ++i) {
    T s = T(0);
    for (int j = 0; j < 6; ++j) s += m[j][i] * v[j];
    r[i] = s;
  }
  return r;
}
template <class T> M6<T> m6m(const M6<T>& a, const M6<T>& b) {
  M6<T> r;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      T s = T(0);
      for (int k = 0; k < 6; ++k) s += a[i][k] * b[k][j];
      r[i][j] = s;
    }
  return r;
}
template <class T> M6<T> m6tm(const M6<T>& a, const M6<T>& b) {  // a^T b
  M6<T> r;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      T s = T(0);
      for (int k = 0; k < 6; ++k) s += a[k][i] * b[k][j];
      r[i][j] = s;
    }
  return r;
}
// Plücker transform from frame A to frame B: E rotates A coords into B coords,
// r = origin of B relative to A in A coords.  X = [E 0; -E r^ E].
template <class T> M6<T> plux(const M3<T>& E, const V3<T>& r) {
  M6<T> X{};
  M3<T> Erx = mm3(E, skew(r));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      X[i][j] = E[i][j];
      X[i][j + 3] = T(0);
      X[i + 3][j] = -Erx[i][j];
      X[i + 3][j + 3] = E[i][j];
    }
  return X;
}
// Rigid-body spatial inertia about the frame origin from mass, CoM and inertia about CoM.
template <class T> M6<T> mcI(T m, const V3<T>& c, const M3<T>& Ic) {
  M6<T> I{};
  M3<T> C = skew(c);
  M3<T> CCt = mm3(C, tr3(C));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      I[i][j] = Ic[i][j] + m * CCt[i][j];
      I[i][j + 3] = m * C[i][j];
      I[i + 3][j] = m * C[j][i];
      I[i + 3][j + 3] = (i == j) ? m : T(0);
    }
  return I;
}

// ---------------------------------------------------------------- model
enum JointType : int { J_FREE = 0, J_REVOLUTE = 1, J_FIXED = 2 };

// One link of the (unfolded) tree.  Links are in topological order (parent < child).
struct LinkDesc {
  int parent;          // -1 for the floating base
  int jtype;           // JointType
  int dof;             // index into q for revolute joints, -1 otherwise
  double xyz[3];       // joint origin in parent frame
  double rpy[3];       // joint origin rotation
  double axis[3];      // joint axis (child frame)
  double mass;
  double com[3];       // CoM in link frame
  double inertia[6];   // about CoM, link axes: xx yy zz xy xz yz
};

struct ContactDesc {
  int link;            // link index in the tree
  double pos[3];       // sphere centre in link frame
  double radius;
};

// Physical parameters that the identification varies (one lane's worth).
// Layout matches the product's parameter vector: see PARAM_* below.
enum ParamIndex : int {
  P_MASS = 0, P_COMX, P_COMY, P_COMZ, P_IXX, P_IYY, P_IZZ, P_IXY, P_IXZ, P_IYZ,
  P_KP0 = 10, P_KD0 = 14, P_FRIC0 = 18, P_MOTOR0 = 22, P_COUNT = 25
};

enum ControlMode : int { CTRL_LEGGED_BASE = 0, CTRL_GO2_OMNI = 1 };
enum MotorModel : int { MOTOR_NONE = 0, MOTOR_SCALAR = 1, MOTOR_VEC3 = 2, MOTOR_VEC3_TANH = 3 };

struct SimConfig {
  double dt;             // physics step (1/fps); torques are recomputed once per physics step
  int decimation;        // physics steps per control step
  int substeps;          // integration sub-intervals per physics step (torque held), the role of
                         // sim.substeps in isaacgym_active_sysid.yaml:17
  double gravity;        // 9.81, along -z
  double kp[12], kd[12]; // PD gains
  double q0[12];         // default joint angles
  double effort[12];     // torque clip
  double action_scale;   // 0.25
  double action_clip;    // 20
  int control_mode;      // ControlMode
  int motor_model;       // MotorModel
  double k_contact, c_contact, mu, c_tangent;
  double fric_vel;       // velocity scale of the smooth Coulomb joint friction
};

// ---------------------------------------------------------------- state
// Root state in IsaacGym layout: pos(3, world) quat(4, xyzw) linvel(3, world) angvel(3, world);
// linvel is the velocity of the base frame origin.
template <class T> struct State {
  T root[13];
  T q[12];
  T qd[12];
};

template <class T> struct Tree {
  int nl = 0;
  std::vector<LinkDesc> links;
  std::vector<ContactDesc> contacts;
  // per-link constants
  std::vector<M3<T>> E_T;     // rotation of the joint origin (parent->child coords, at q = 0)
  std::vector<V3<T>> r_T;     // joint origin in parent coords
  std::vector<V6<T>> S;       // motion subspace (revolute: [axis;0])
  std::vector<M6<T>> I;       // spatial inertia (link frame)
  int base_param_link = 0;    // which link the base parameters overwrite

  void build(const LinkDesc* ls, int n, const ContactDesc* cs, int nc) {
    nl = n;
    links.assign(ls, ls + n);
    contacts.assign(cs, cs + nc);
    E_T.resize(n); r_T.resize(n); S.resize(n); I.resize(n);
    for (int i = 0; i < n; ++i) {
      const LinkDesc& L = links[i];
      E_T[i] = tr3(rpy_matrix<T>(L.rpy[0], L.rpy[1], L.rpy[2]));
      r_T[i] = {T(L.xyz[0]), T(L.xyz[1]), T(L.xyz[2])};
      S[i] = {T(L.axis[0]), T(L.axis[1]), T(L.axis[2]), T(0), T(0), T(0)};
      set_inertia(i, T(L.mass), L.com, L.inertia);
    }
  }
  void set_inertia(int i, T m, const double* com, const double* in6) {
    V3<T> c{T(com[0]), T(com[1]), T(com[2])};
    M3<T> Ic;
    Ic[0] = {T(in6[0]), T(in6[3]), T(in6[4])};
    Ic[1] = {T(in6[3]), T(in6[1]), T(in6[5])};
    Ic[2] = {T(in6[4]), T(in6[5]), T(in6[2])};
    I[i] = mcI(m, c, Ic);
  }
  void set_inertia_T(int i, T m, const V3<T>& c, const M3<T>& Ic) { I[i] = mcI(m, c, Ic); }
};

// Joint rotation for a revolute joint about unit axis a by angle q, as a coordinate
// transform parent->child (transpose of the child-to-parent rotation).
template <class T> M3<T> axis_rot_T(const double* a, T q) {
  T c = s_cos(q), s = s_sin(q), C = T(1) - c;
  T x = T(a[0]), y = T(a[1]), z = T(a[2]);
  M3<T> R;  // Rodrigues, child->parent
  R[0] = {c + x * x * C, x * y * C - z * s, x * z * C + y * s};
  R[1] = {y * x * C + z * s, c + y * y * C, y * z * C - x * s};
  R[2] = {z * x * C - y * s, z * y * C + x * s, c + z * z * C};
  return tr3(R);
}

// Kinematics of the whole tree for one state.
template <class T> struct Kin {
  std::vector<M6<T>> Xup;     // parent->link Plücker transforms
  std::vector<V6<T>> v;       // link spatial velocities (link coords)
  std::vector<M3<T>> Rw;      // link->world rotation
  std::vector<V3<T>> pw;      // link origin, world
};

template <class T> void kinematics(const Tree<T>& tr, const State<T>& st, const V6<T>& v0, Kin<T>& k) {
  int n = tr.nl;
  k.Xup.resize(n); k.v.resize(n); k.Rw.resize(n); k.pw.resize(n);
  for (int i = 0; i < n; ++i) {
    const LinkDesc& L = tr.links[i];
    if (L.parent < 0) {
      k.Xup[i] = M6<T>{};
      k.v[i] = v0;
      k.Rw[i] = quat_matrix(st.root + 3);
      k.pw[i] = {st.root[0], st.root[1], st.root[2]};
      continue;
    }
    M3<T> E = tr.E_T[i];
    if (L.jtype == J_REVOLUTE) E = mm3(axis_rot_T<T>(L.axis, st.q[L.dof]), E);
    k.Xup[i] = plux(E, tr.r_T[i]);
    V6<T> vj{};
    if (L.jtype == J_REVOLUTE)
      for (int d = 0; d < 6; ++d) vj[d] = tr.S[i][d] * st.qd[L.dof];
    V6<T> vp = m6v(k.Xup[i], k.v[L.parent]);
    for (int d = 0; d < 6; ++d) k.v[i][d] = vp[d] + vj[d];
    k.Rw[i] = mm3(k.Rw[L.parent], tr3(E));
    k.pw[i] = mv3(k.Rw[L.parent], tr.r_T[i]);
    for (int d = 0; d < 3; ++d) k.pw[i][d] += k.pw[L.parent][d];
  }
}

// Base spatial velocity in base coords from the IsaacGym root state.
template <class T> V6<T> base_body_velocity(const State<T>& st) {
  M3<T> R = quat_matrix(st.root + 3);
  V3<T> w{st.root[10], st.root[11], st.root[12]}, v{st.root[7], st.root[8], st.root[9]};
  V3<T> wb = mtv3(R, w), vb = mtv3(R, v);
  return {wb[0], wb[1], wb[2], vb[0], vb[1], vb[2]};
}

// Penalty contact forces, returned as spatial forces in link coordinates (fext[link]).
// Ground plane z = 0.  Normal: Fn = k*pen*(1 - tanh(c*vz / (k*pen))), a spring-damper
// (k*pen - c*vz where |c*vz| << k*pen) whose damping force saturates at the spring force, so
// 0 <= Fn <= 2*k*pen and Fn -> 0 at the contact onset whatever the approach speed; tangential:
// viscous with slope c_tangent, capped by the Coulomb cone mu*Fn.
template <class T> void contact_forces(const Tree<T>& tr, const Kin<T>& k, const SimConfig& cfg,
                                       std::vector<V6<T>>& fext, T* foot_force_world = nullptr) {
  for (auto& f : fext) f = V6<T>{};
  for (size_t ci = 0; ci < tr.contacts.size(); ++ci) {
    const ContactDesc& C = tr.contacts[ci];
    int l = C.link;
    const M3<T>& R = k.Rw[l];
    V3<T> pl{T(C.pos[0]), T(C.pos[1]), T(C.pos[2])};
    V3<T> centre = mv3(R, pl);
    for (int d = 0; d < 3; ++d) centre[d] += k.pw[l][d];
    T pen = T(C.radius) - centre[2];
    V3<T> fw{T(0), T(0), T(0)};
    if (s_pos(pen)) {
      // contact point = lowest point of the sphere; local coordinates of that material point
      V3<T> down{T(0), T(0), -T(C.radius)};
      V3<T> off = mtv3(R, down);
      V3<T> pc{pl[0] + off[0], pl[1] + off[1], pl[2] + off[2]};
      V3<T> w{k.v[l][0], k.v[l][1], k.v[l][2]}, vo{k.v[l][3], k.v[l][4], k.v[l][5]};
      V3<T> wxp = cross(w, pc);
      V3<T> vl{vo[0] + wxp[0], vo[1] + wxp[1], vo[2] + wxp[2]};
      V3<T> vc = mv3(R, vl);
      const T kpen = T(cfg.k_contact) * pen;
      T fn = kpen * (T(1) - s_tanh(T(cfg.c_contact) * vc[2] / kpen));   // damping saturates at k*pen
      T vt2 = vc[0] * vc[0] + vc[1] * vc[1];
      T ft = T(0), sc = T(0);
      if (s_pos(vt2)) {
        T vt = s_sqrt(vt2);
        ft = T(cfg.c_tangent) * vt;
        T cap = T(cfg.mu) * fn;
        if (ft > cap) ft = cap;
        sc = ft / vt;
      }
      fw = {-sc * vc[0], -sc * vc[1], fn};
      V3<T> fl = mtv3(R, fw);
      V3<T> nl = cross(pc, fl);
      V6<T> f6{nl[0], nl[1], nl[2], fl[0], fl[1], fl[2]};
      for (int d = 0; d < 6; ++d) fext[l][d] += f6[d];
    }
    if (foot_force_world)
      for (int d = 0; d < 3; ++d) foot_force_world[3 * ci + d] = fw[d];
  }
}

template <class T> V6<T> solve6(M6<T> A, V6<T> b);

// Featherstone ABA, floating base, gravity via the fictitious base acceleration.
// Inputs: joint torques tau[12], external forces per link (link coords).
// Outputs: base spatial acceleration a0 (base coords, true acceleration incl. gravity), qdd.
template <class T> void aba(const Tree<T>& tr, const Kin<T>& k, const T* qd, const T* tau,
                            const std::vector<V6<T>>& fext, const V3<T>& g_base, V6<T>& a0, T* qdd) {
  int n = tr.nl;
  std::vector<M6<T>> IA(n);
  std::vector<V6<T>> pA(n), c(n), U(n);
  std::vector<T> D(n), u(n);
  for (int i = 0; i < n; ++i) {
    IA[i] = tr.I[i];
    V6<T> h = m6v(tr.I[i], k.v[i]);
    V6<T> b = crf(k.v[i], h);
    for (int d = 0; d < 6; ++d) pA[i][d] = b[d] - fext[i][d];
    const LinkDesc& L = tr.links[i];
    if (L.parent >= 0 && L.jtype == J_REVOLUTE) {
      V6<T> sq;  // S qd
      for (int d = 0; d < 6; ++d) sq[d] = tr.S[i][d] * qd[L.dof];
      c[i] = crm(k.v[i], sq);
    } else {
      c[i] = V6<T>{};
    }
  }
  for (int i = n - 1; i >= 1; --i) {
    const LinkDesc& L = tr.links[i];
    int p = L.parent;
    M6<T> Ia = IA[i];
    V6<T> pa = pA[i];
    if (L.jtype == J_REVOLUTE) {
      U[i] = m6v(IA[i], tr.S[i]);
      T Dv = T(0);
      for (int d = 0; d < 6; ++d) Dv += tr.S[i][d] * U[i][d];
      D[i] = Dv;
      T sp = T(0);
      for (int d = 0; d < 6; ++d) sp += tr.S[i][d] * pA[i][d];
      u[i] = tau[L.dof] - sp;
      for (int r = 0; r < 6; ++r)
        for (int s = 0; s < 6; ++s) Ia[r][s] -= U[i][r] * U[i][s] / Dv;
      V6<T> Iac = m6v(Ia, c[i]);
      for (int d = 0; d < 6; ++d) pa[d] = pA[i][d] + Iac[d] + U[i][d] * u[i] / Dv;
    }
    // IA[p] += X^T Ia X ; pA[p] += X^T pa
    M6<T> t = m6m(Ia, k.Xup[i]);
    M6<T> XtIX = m6tm(k.Xup[i], t);
    for (int r = 0; r < 6; ++r)
      for (int s = 0; s < 6; ++s) IA[p][r][s] += XtIX[r][s];
    V6<T> Xtp = m6tv(k.Xup[i], pa);
    for (int d = 0; d < 6; ++d) pA[p][d] += Xtp[d];
  }
  // base: IA0 a0' = -pA0, solved by Gaussian elimination with partial pivoting
  M6<T> A = IA[0];
  V6<T> b;
  for (int d = 0; d < 6; ++d) b[d] = -pA[0][d];
  V6<T> x;
  if (tr.links[0].jtype == J_FIXED) {
    // fixed base: true acceleration zero, relative (gravity-free) acceleration = -a_g
    x = V6<T>{T(0), T(0), T(0), -g_base[0], -g_base[1], -g_base[2]};
  } else {
    x = solve6(A, b);
  }
  std::vector<V6<T>> a(n);
  a[0] = x;  // relative (gravity-free) acceleration
  for (int i = 1; i < n; ++i) {
    const LinkDesc& L = tr.links[i];
    V6<T> ap = m6v(k.Xup[i], a[L.parent]);
    for (int d = 0; d < 6; ++d) a[i][d] = ap[d] + c[i][d];
    if (L.jtype == J_REVOLUTE) {
      T ua = T(0);
      for (int d = 0; d < 6; ++d) ua += U[i][d] * a[i][d];
      T qa = (u[i] - ua) / D[i];
      qdd[L.dof] = qa;
      for (int d = 0; d < 6; ++d) a[i][d] += tr.S[i][d] * qa;
    }
  }
  a0 = x;
  a0[3] += g_base[0];
  a0[4] += g_base[1];
  a0[5] += g_base[2];
}

template <class T> V6<T> solve6(M6<T> A, V6<T> b) {
  for (int col = 0; col < 6; ++col) {
    int piv = col;
    for (int r = col + 1; r < 6; ++r) {
      using std::abs;
      if (abs(A[r][col]) > abs(A[piv][col])) piv = r;
    }
    if (piv != col) { std::swap(A[piv], A[col]); std::swap(b[piv], b[col]); }
    for (int r = col + 1; r < 6; ++r) {
      T f = A[r][col] / A[col][col];
      for (int s = col; s < 6; ++s) A[r][s] -= f * A[col][s];
      b[r] -= f * b[col];
    }
  }
  V6<T> x;
  for (int r = 5; r >= 0; --r) {
    T s = b[r];
    for (int c2 = r + 1; c2 < 6; ++c2) s -= A[r][c2] * x[c2];
    x[r] = s / A[r][r];
  }
  return x;
}

// RNEA: generalized forces for given generalized accelerations (base a0 is the TRUE
// spatial acceleration in base coords, gravity g_base is applied as a body force).
// Returns tau_base (6) and tau joints (12).
template <class T> void rnea(const Tree<T>& tr, const Kin<T>& k, const V6<T>& a0, const T* qdd,
                             const std::vector<V6<T>>& fext, const V3<T>& g_base, T* tau_base, T* tau) {
  int n = tr.nl;
  std::vector<V6<T>> a(n), f(n), ag(n);
  // relative acceleration = a - a_g
  a[0] = a0;
  a[0][3] -= g_base[0];
  a[0][4] -= g_base[1];
  a[0][5] -= g_base[2];
  for (int i = 0; i < n; ++i) {
    const LinkDesc& L = tr.links[i];
    if (i > 0) {
      V6<T> ap = m6v(k.Xup[i], a[L.parent]);
      V6<T> vp = m6v(k.Xup[i], k.v[L.parent]);
      T qdv = T(0);
      if (L.jtype == J_REVOLUTE)
        for (int d = 0; d < 6; ++d) qdv += tr.S[i][d] * (k.v[i][d] - vp[d]);
      V6<T> sq;
      for (int d = 0; d < 6; ++d) sq[d] = tr.S[i][d] * qdv;
      V6<T> cc = (L.jtype == J_REVOLUTE) ? crm(k.v[i], sq) : V6<T>{};
      for (int d = 0; d < 6; ++d) a[i][d] = ap[d] + cc[d];
      if (L.jtype == J_REVOLUTE)
        for (int d = 0; d < 6; ++d) a[i][d] += tr.S[i][d] * qdd[L.dof];
    }
    V6<T> Ia = m6v(tr.I[i], a[i]);
    V6<T> h = m6v(tr.I[i], k.v[i]);
    V6<T> b = crf(k.v[i], h);
    for (int d = 0; d < 6; ++d) f[i][d] = Ia[d] + b[d] - fext[i][d];
  }
  for (int i = n - 1; i >= 1; --i) {
    const LinkDesc& L = tr.links[i];
    if (L.jtype == J_REVOLUTE) {
      T s = T(0);
      for (int d = 0; d < 6; ++d) s += tr.S[i][d] * f[i][d];
      tau[L.dof] = s;
    }
    V6<T> fp = m6tv(k.Xup[i], f[i]);
    for (int d = 0; d < 6; ++d) f[L.parent][d] += fp[d];
  }
  for (int d = 0; d < 6; ++d) tau_base[d] = f[0][d];
}

// CRBA: joint-space inertia matrix H (18x18, base block first) of the tree.
template <class T> void crba(const Tree<T>& tr, const Kin<T>& k, T* H /*18*18*/) {
  int n = tr.nl;
  std::vector<M6<T>> Ic(tr.I.begin(), tr.I.end());
  for (int i = n - 1; i >= 1; --i) {
    M6<T> t = m6m(Ic[i], k.Xup[i]);
    M6<T> XtIX = m6tm(k.Xup[i], t);
    int p = tr.links[i].parent;
    for (int r = 0; r < 6; ++r)
      for (int s = 0; s < 6; ++s) Ic[p][r][s] += XtIX[r][s];
  }
  for (int i = 0; i < 18 * 18; ++i) H[i] = T(0);
  auto col_of = [&](int i) { return 6 + tr.links[i].dof; };
  for (int i = 1; i < n; ++i) {
    if (tr.links[i].jtype != J_REVOLUTE) continue;
    V6<T> F = m6v(Ic[i], tr.S[i]);
    int ci = col_of(i);
    T hii = T(0);
    for (int d = 0; d < 6; ++d) hii += tr.S[i][d] * F[d];
    H[ci * 18 + ci] = hii;
    int j = i;
    while (tr.links[j].parent >= 0) {
      F = m6tv(k.Xup[j], F);
      j = tr.links[j].parent;
      if (tr.links[j].parent < 0) {
        for (int d = 0; d < 6; ++d) { H[ci * 18 + d] = F[d]; H[d * 18 + ci] = F[d]; }
      } else if (tr.links[j].jtype == J_REVOLUTE) {
        int cj = col_of(j);
        T s = T(0);
        for (int d = 0; d < 6; ++d) s += tr.S[j][d] * F[d];
        H[ci * 18 + cj] = s;
        H[cj * 18 + ci] = s;
      }
    }
  }
  for (int r = 0; r < 6; ++r)
    for (int s = 0; s < 6; ++s) H[r * 18 + s] = Ic[0][r][s];
}

// ---------------------------------------------------------------- control
// PD + clip + motor model for one env.  action: raw policy/data action (12).
// Follows legged_robot_base.py:185-198 (action clip), :542-560 (legged_base PD and clip),
// go2_omni.py:436-462 (hip x0.5, PD, clip), active_sysid_openloop.py:174-187, 356-400
// (motor model applied to the clipped torque).
template <class T> void control_torque(const SimConfig& cfg, const T* action, const T* q, const T* qd,
                                       const T* par, T* tau_out) {
  for (int j = 0; j < 12; ++j) {
    int leg = j / 3, grp = j % 3;
    T a = action[j];
    T clip = T(cfg.action_clip);
    if (a > clip) a = clip;
    if (a < -clip) a = -clip;
    T as = a * T(cfg.action_scale);
    if (cfg.control_mode == CTRL_GO2_OMNI && grp == 0) as = as * T(0.5);
    T kp = par[P_KP0 + leg] * T(cfg.kp[j]);
    T kd = par[P_KD0 + leg] * T(cfg.kd[j]);
    T tau = kp * (as + T(cfg.q0[j]) - q[j]) - kd * qd[j];
    T lim = T(cfg.effort[j]);
    if (tau > lim) tau = lim;
    if (tau < -lim) tau = -lim;
    switch (cfg.motor_model) {
      case MOTOR_SCALAR: tau = tau * par[P_MOTOR0]; break;
      case MOTOR_VEC3: tau = tau * par[P_MOTOR0 + grp]; break;
      case MOTOR_VEC3_TANH: {
        T A = par[P_MOTOR0 + grp];
        T B = T(1) / A;
        tau = A * s_tanh(B * tau);
        break;
      }
      default: break;
    }
    tau_out[j] = tau;
  }
}

// ---------------------------------------------------------------- one physics step
template <class T> struct Engine {
  Tree<T> tree;
  SimConfig cfg;
  int base_link = 0;
  // constant inertia rigidly attached to the base link (head links), folded by the caller
  // into the base link's inertia when parameters are applied.
  void apply_base_params(const T* par) {
    // base link inertia := parameters (mass, com, I about com); attached fixed links stay
    // separate links of the unfolded tree, exactly as IsaacGym keeps them.
    V3<T> c{par[P_COMX], par[P_COMY], par[P_COMZ]};
    M3<T> Ic;
    Ic[0] = {par[P_IXX], par[P_IXY], par[P_IXZ]};
    Ic[1] = {par[P_IXY], par[P_IYY], par[P_IYZ]};
    Ic[2] = {par[P_IXZ], par[P_IYZ], par[P_IZZ]};
    tree.set_inertia_T(base_link, par[P_MASS], c, Ic);
  }

  // Forward dynamics for given joint torques (motor output, before joint friction).
  void forward(const State<T>& st, const T* tau_motor, const T* par, V6<T>& a0, T* qdd,
               T* foot_force = nullptr) const {
    V6<T> v0 = base_body_velocity(st);
    Kin<T> k;
    kinematics(tree, st, v0, k);
    std::vector<V6<T>> fext(tree.nl);
    contact_forces(tree, k, cfg, fext, foot_force);
    T tau[12];
    for (int j = 0; j < 12; ++j) {
      T f = par[P_FRIC0 + j / 3];
      tau[j] = tau_motor[j] - f * s_tanh(st.qd[j] / T(cfg.fric_vel));
    }
    M3<T> R = quat_matrix(st.root + 3);
    V3<T> gw{T(0), T(0), -T(cfg.gravity)};
    V3<T> gb = mtv3(R, gw);
    aba(tree, k, st.qd, tau, fext, gb, a0, qdd);
  }

  // Semi-implicit Euler: velocities first, then positions with the new velocities.
  void integrate(State<T>& st, const V6<T>& a0, const T* qdd) const {
#if SPIO_MUTATION == 4
    T dt = T(cfg.dt);
#else
    T dt = T(cfg.dt) / T(cfg.substeps);
#endif
    V6<T> v0 = base_body_velocity(st);
#if SPIO_MUTATION == 1
    const V6<T> v_old = v0;
#endif
    for (int d = 0; d < 6; ++d) v0[d] += dt * a0[d];
    for (int j = 0; j < 12; ++j) {
#if SPIO_MUTATION == 1
      st.q[j] += dt * st.qd[j];
      st.qd[j] += dt * qdd[j];
#else
      st.qd[j] += dt * qdd[j];
      st.q[j] += dt * st.qd[j];
#endif
    }
    M3<T> R = quat_matrix(st.root + 3);
    V3<T> vb{v0[3], v0[4], v0[5]}, wb{v0[0], v0[1], v0[2]};
#if SPIO_MUTATION == 1
    V3<T> vw = mv3(R, V3<T>{v_old[3], v_old[4], v_old[5]});
#else
    V3<T> vw = mv3(R, vb);
#endif
    for (int d = 0; d < 3; ++d) st.root[d] += dt * vw[d];
    // q <- q (x) exp(w dt / 2), body-frame angular velocity
#if SPIO_MUTATION == 2
    V3<T> wq = mv3(R, wb);
#else
    V3<T> wq = wb;
#endif
    T th2 = (wq[0] * wq[0] + wq[1] * wq[1] + wq[2] * wq[2]) * dt * dt;
    T dq[4];
    if (s_pos(th2)) {
      T th = s_sqrt(th2);
      T s = s_sin(th * T(0.5)) / th;
      dq[0] = wq[0] * dt * s; dq[1] = wq[1] * dt * s; dq[2] = wq[2] * dt * s; dq[3] = s_cos(th * T(0.5));
    } else {
      dq[0] = dq[1] = dq[2] = T(0); dq[3] = T(1);
    }
    T qn[4];
#if SPIO_MUTATION == 3
    quat_mul(dq, st.root + 3, qn);
#else
    quat_mul(st.root + 3, dq, qn);
#endif
    T nrm = s_sqrt(qn[0] * qn[0] + qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]);
    for (int d = 0; d < 4; ++d) st.root[3 + d] = qn[d] / nrm;
    M3<T> Rn = quat_matrix(st.root + 3);
    V3<T> vwn = mv3(Rn, vb), wwn = mv3(Rn, wb);
    for (int d = 0; d < 3; ++d) { st.root[7 + d] = vwn[d]; st.root[10 + d] = wwn[d]; }
  }

  void physics_step(State<T>& st, const T* tau_motor, const T* par, T* foot_force = nullptr) const {
    for (int k = 0; k < cfg.substeps; ++k) {
      V6<T> a0;
      T qdd[12];
      forward(st, tau_motor, par, a0, qdd, foot_force);
      integrate(st, a0, qdd);
    }
  }

  // One control step: torques recomputed at every physics step from the current state
  // (legged_robot_base.py:201-209).
  void control_step(State<T>& st, const T* action, const T* par) const {
#if SPIO_MUTATION == 5
    T tau[12];
    control_torque(cfg, action, st.q, st.qd, par, tau);
    for (int s = 0; s < cfg.decimation; ++s) physics_step(st, tau, par);
#else
    for (int s = 0; s < cfg.decimation; ++s) {
      T tau[12];
      control_torque(cfg, action, st.q, st.qd, par, tau);
      physics_step(st, tau, par);
    }
#endif
  }
};

// ---------------------------------------------------------------- cost
// Tracking cost of one (candidate, segment) rollout.
//  mode 0 (tracking): sum over steps t = 1..H of w . squared errors of
//        [pos, quat(raw xyzw), linvel, angvel, q, qd]
//  mode 1 (reference): scripts/eval.py:287-292 — at the horizon end only,
//        w_pos*||dpos|| + w_quat*||dquat_xyzw|| + w_q*||dq|| (L2 norms, not squared)
enum CostMode : int { COST_TRACKING = 0, COST_REFERENCE = 1 };

template <class T> T step_cost(int mode, const double* w, const State<T>& st, const float* tgt /*37*/) {
  T e[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int d = 0; d < 3; ++d) { T x = st.root[d] - T(tgt[d]); e[0] += x * x; }
  for (int d = 0; d < 4; ++d) { T x = st.root[3 + d] - T(tgt[3 + d]); e[1] += x * x; }
  for (int d = 0; d < 3; ++d) { T x = st.root[7 + d] - T(tgt[7 + d]); e[2] += x * x; }
  for (int d = 0; d < 3; ++d) { T x = st.root[10 + d] - T(tgt[10 + d]); e[3] += x * x; }
  for (int j = 0; j < 12; ++j) { T x = st.q[j] - T(tgt[13 + j]); e[4] += x * x; }
  for (int j = 0; j < 12; ++j) { T x = st.qd[j] - T(tgt[25 + j]); e[5] += x * x; }
  if (mode == COST_TRACKING) {
    T s = T(0);
    for (int g = 0; g < 6; ++g) s += T(w[g]) * e[g];
    return s;
  }
  return T(w[0]) * s_sqrt(e[0]) + T(w[1]) * s_sqrt(e[1]) + T(w[4]) * s_sqrt(e[4]);
}

}  // namespace spio
