// ORACLE — test infrastructure only (see spi_oracle.hpp).  C entry points for ctypes.
#include "spi_oracle.hpp"
#include "spi_dual64.hpp"
#include "spi_philox.hpp"

#include <cstdlib>
#include <new>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace spio;

namespace {
template <class T> struct Handle {
  Engine<T> eng;
};

template <class T> void load_state(State<T>& st, const double* root, const double* q, const double* qd) {
  for (int i = 0; i < 13; ++i) st.root[i] = T(root[i]);
  for (int j = 0; j < 12; ++j) { st.q[j] = T(q[j]); st.qd[j] = T(qd[j]); }
}
template <class T> void store_state(const State<T>& st, double* root, double* q, double* qd) {
  for (int i = 0; i < 13; ++i) root[i] = double(st.root[i]);
  for (int j = 0; j < 12; ++j) { q[j] = double(st.q[j]); qd[j] = double(st.qd[j]); }
}
template <class T> void load37(State<T>& st, const float* s) {
  for (int i = 0; i < 13; ++i) st.root[i] = T(s[i]);
  for (int j = 0; j < 12; ++j) { st.q[j] = T(s[13 + j]); st.qd[j] = T(s[25 + j]); }
}
template <class T> void store37(const State<T>& st, float* s) {
  for (int i = 0; i < 13; ++i) s[i] = float(st.root[i]);
  for (int j = 0; j < 12; ++j) { s[13 + j] = float(st.q[j]); s[25 + j] = float(st.qd[j]); }
}
template <class T> bool finite_state(const State<T>& st) {
  for (int i = 0; i < 13; ++i) if (!std::isfinite(double(st.root[i]))) return false;
  for (int j = 0; j < 12; ++j) if (!std::isfinite(double(st.q[j])) || !std::isfinite(double(st.qd[j]))) return false;
  return true;
}
}  // namespace

extern "C" {

// ---------------------------------------------------------------- handles
void* spio_new(const LinkDesc* links, int nl, const ContactDesc* cs, int nc, int base_link, const SimConfig* cfg) {
  auto* h = new (std::nothrow) Handle<double>();
  if (!h) return nullptr;
  h->eng.tree.build(links, nl, cs, nc);
  h->eng.cfg = *cfg;
  h->eng.base_link = base_link;
  return h;
}
void spio_free(void* h) { delete static_cast<Handle<double>*>(h); }

void* spio_new_f32(const LinkDesc* links, int nl, const ContactDesc* cs, int nc, int base_link, const SimConfig* cfg) {
  auto* h = new (std::nothrow) Handle<float>();
  if (!h) return nullptr;
  h->eng.tree.build(links, nl, cs, nc);
  h->eng.cfg = *cfg;
  h->eng.base_link = base_link;
  return h;
}
void spio_free_f32(void* h) { delete static_cast<Handle<float>*>(h); }

int spio_sizeof_config() { return int(sizeof(SimConfig)); }
int spio_sizeof_link() { return int(sizeof(LinkDesc)); }
int spio_sizeof_contact() { return int(sizeof(ContactDesc)); }

// ---------------------------------------------------------------- dynamics pieces (double)
// Forward dynamics: base spatial acceleration (base coords) and qdd for a given motor torque.
// tau_applied (optional) receives the joint torque after joint friction.
int spio_forward(void* hp, const double* root, const double* q, const double* qd, const double* tau_motor,
                 const double* par, double* a0_out, double* qdd_out, double* foot_force, double* tau_applied) {
  auto* h = static_cast<Handle<double>*>(hp);
  h->eng.apply_base_params(par);
  State<double> st;
  load_state(st, root, q, qd);
  V6<double> a0;
  h->eng.forward(st, tau_motor, par, a0, qdd_out, foot_force);
  for (int d = 0; d < 6; ++d) a0_out[d] = a0[d];
  if (tau_applied)
    for (int j = 0; j < 12; ++j)
      tau_applied[j] = tau_motor[j] - par[P_FRIC0 + j / 3] * std::tanh(qd[j] / h->eng.cfg.fric_vel);
  return 0;
}

// RNEA with the same contact forces and gravity: generalized forces for (a0, qdd).
int spio_rnea(void* hp, const double* root, const double* q, const double* qd, const double* a0_in,
              const double* qdd, const double* par, double* tau_base, double* tau) {
  auto* h = static_cast<Handle<double>*>(hp);
  h->eng.apply_base_params(par);
  State<double> st;
  load_state(st, root, q, qd);
  V6<double> v0 = base_body_velocity(st);
  Kin<double> k;
  kinematics(h->eng.tree, st, v0, k);
  std::vector<V6<double>> fext(h->eng.tree.nl);
  contact_forces(h->eng.tree, k, h->eng.cfg, fext);
  M3<double> R = quat_matrix(st.root + 3);
  V3<double> gb = mtv3(R, V3<double>{0.0, 0.0, -h->eng.cfg.gravity});
  V6<double> a0;
  for (int d = 0; d < 6; ++d) a0[d] = a0_in[d];
  rnea(h->eng.tree, k, a0, qdd, fext, gb, tau_base, tau);
  return 0;
}

// CRBA joint-space inertia (18x18, base block first, base velocity in base coords).
int spio_crba(void* hp, const double* root, const double* q, const double* par, double* H) {
  auto* h = static_cast<Handle<double>*>(hp);
  h->eng.apply_base_params(par);
  State<double> st;
  double zero[12] = {0};
  load_state(st, root, q, zero);
  for (int d = 7; d < 13; ++d) st.root[d] = 0.0;
  Kin<double> k;
  kinematics(h->eng.tree, st, V6<double>{}, k);
  crba(h->eng.tree, k, H);
  return 0;
}

// World-frame spatial momentum about the world origin [angular; linear] and kinetic energy.
int spio_momentum(void* hp, const double* root, const double* q, const double* qd, const double* par,
                  double* h_world, double* kinetic) {
  auto* h = static_cast<Handle<double>*>(hp);
  h->eng.apply_base_params(par);
  State<double> st;
  load_state(st, root, q, qd);
  V6<double> v0 = base_body_velocity(st);
  Kin<double> k;
  kinematics(h->eng.tree, st, v0, k);
  double L[3] = {0, 0, 0}, P[3] = {0, 0, 0}, E = 0.0;
  for (int i = 0; i < h->eng.tree.nl; ++i) {
    V6<double> hb = m6v(h->eng.tree.I[i], k.v[i]);
    for (int d = 0; d < 6; ++d) E += 0.5 * k.v[i][d] * hb[d];
    V3<double> n{hb[0], hb[1], hb[2]}, f{hb[3], hb[4], hb[5]};
    V3<double> nw = mv3(k.Rw[i], n), fw = mv3(k.Rw[i], f);
    V3<double> pxf = cross(k.pw[i], fw);
    for (int d = 0; d < 3; ++d) { L[d] += nw[d] + pxf[d]; P[d] += fw[d]; }
  }
  for (int d = 0; d < 3; ++d) { h_world[d] = L[d]; h_world[3 + d] = P[d]; }
  *kinetic = E;
  return 0;
}

// World positions of every link origin (n_links x 3) and foot-sphere centres (nc x 3).
int spio_fk(void* hp, const double* root, const double* q, double* link_pos, double* foot_pos) {
  auto* h = static_cast<Handle<double>*>(hp);
  State<double> st;
  double zero[12] = {0};
  load_state(st, root, q, zero);
  Kin<double> k;
  kinematics(h->eng.tree, st, V6<double>{}, k);
  for (int i = 0; i < h->eng.tree.nl; ++i)
    for (int d = 0; d < 3; ++d) link_pos[3 * i + d] = k.pw[i][d];
  for (size_t c = 0; c < h->eng.tree.contacts.size(); ++c) {
    const ContactDesc& C = h->eng.tree.contacts[c];
    V3<double> pl{C.pos[0], C.pos[1], C.pos[2]};
    V3<double> p = mv3(k.Rw[C.link], pl);
    for (int d = 0; d < 3; ++d) foot_pos[3 * c + d] = p[d] + k.pw[C.link][d];
  }
  return 0;
}

int spio_physics_step(void* hp, double* root, double* q, double* qd, const double* tau_motor, const double* par) {
  auto* h = static_cast<Handle<double>*>(hp);
  h->eng.apply_base_params(par);
  State<double> st;
  load_state(st, root, q, qd);
  h->eng.physics_step(st, tau_motor, par);
  store_state(st, root, q, qd);
  return 0;
}

int spio_control_torque(const SimConfig* cfg, const double* action, const double* q, const double* qd,
                        const double* par, double* tau) {
  control_torque<double>(*cfg, action, q, qd, par, tau);
  return 0;
}

// Simulate n_steps control steps from an initial 37-state, recording the state after each.
int spio_simulate(void* hp, const double* par, const float* init37, const float* actions, int n_steps,
                  float* states_out) {
  auto* h = static_cast<Handle<double>*>(hp);
  h->eng.apply_base_params(par);
  State<double> st;
  load37(st, init37);
  for (int t = 0; t < n_steps; ++t) {
    double a[12];
    for (int j = 0; j < 12; ++j) a[j] = actions[12 * t + j];
    h->eng.control_step(st, a, par);
    store37(st, states_out + 37 * t);
  }
  return finite_state(st) ? 0 : 1;
}

// Rollout cost for n_cand parameter vectors (rows of par[n_cand][25]) on n_seg segments.
// seg_init [n_seg][37], seg_actions [n_seg][H][12], seg_targets [n_seg][H][37].
// cost_out [n_cand][n_seg] (per-segment cost; tracking mode sums over steps).
int spio_rollout_cost(void* hp, const double* par, int n_cand, const float* seg_init, const float* seg_actions,
                      const float* seg_targets, int n_seg, int H, int cost_mode, const double* w, double* cost_out) {
  auto* h = static_cast<Handle<double>*>(hp);
  int bad = 0;
  for (int c = 0; c < n_cand; ++c) {
    const double* p = par + P_COUNT * c;
    h->eng.apply_base_params(p);
    for (int s = 0; s < n_seg; ++s) {
      State<double> st;
      load37(st, seg_init + 37 * s);
      double acc = 0.0;
      for (int t = 0; t < H; ++t) {
        double a[12];
        for (int j = 0; j < 12; ++j) a[j] = seg_actions[(size_t(s) * H + t) * 12 + j];
        h->eng.control_step(st, a, p);
        const float* tg = seg_targets + (size_t(s) * H + t) * 37;
        if (cost_mode == COST_TRACKING || t == H - 1) acc += step_cost<double>(cost_mode, w, st, tg);
      }
      if (!finite_state(st) || !std::isfinite(acc)) { acc = INFINITY; ++bad; }
      cost_out[size_t(c) * n_seg + s] = acc;
    }
  }
  return bad;
}

// FP32 engine over OpenMP threads: the timed host-CPU baseline (same algorithm as above).
int spio_rollout_cost_f32(void* hp, const float* par, int n_cand, const float* seg_init, const float* seg_actions,
                          const float* seg_targets, int n_seg, int H, int cost_mode, const double* w, float* cost_out,
                          int n_threads) {
  auto* h = static_cast<Handle<float>*>(hp);
  int total = n_cand * n_seg;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
#pragma omp parallel
  {
    Engine<float> eng = h->eng;  // per-thread copy (base inertia is overwritten per lane)
#pragma omp for schedule(dynamic, 4)
    for (int lane = 0; lane < total; ++lane) {
      int c = lane / n_seg, s = lane % n_seg;
      const float* p = par + P_COUNT * c;
      eng.apply_base_params(p);
      State<float> st;
      load37(st, seg_init + 37 * s);
      float acc = 0.0f;
      for (int t = 0; t < H; ++t) {
        float a[12];
        for (int j = 0; j < 12; ++j) a[j] = seg_actions[(size_t(s) * H + t) * 12 + j];
        eng.control_step(st, a, p);
        const float* tg = seg_targets + (size_t(s) * H + t) * 37;
        if (cost_mode == COST_TRACKING || t == H - 1) acc += step_cost<float>(cost_mode, w, st, tg);
      }
      cost_out[size_t(c) * n_seg + s] = acc;
    }
  }
  return 0;
}

// ---------------------------------------------------------------- FP64 derivatives
// Forward-mode derivatives of the FP64 engine with respect to P <= 25 parameter rows idx[].
using DualP = spio_ad::DualD<P_COUNT>;

static void dual_engine(const Handle<double>* h, Engine<DualP>& eng) {
  eng.tree.build(h->eng.tree.links.data(), h->eng.tree.nl, h->eng.tree.contacts.data(),
                 int(h->eng.tree.contacts.size()));
  eng.cfg = h->eng.cfg;
  eng.base_link = h->eng.base_link;
}
static void dual_params(const double* par, const int* idx, int P, DualP* p) {
  for (int k = 0; k < P_COUNT; ++k) p[k] = DualP(par[k]);
  for (int m = 0; m < P; ++m) p[idx[m]].d[m] = 1.0;
}

// Trajectory and sensitivities: after control step t, y = [root (13) | q (12) | qd (12)][:n_rows]
// and J[t][i][m] = d y[i] / d par[idx[m]].  y_out [T][n_rows], J_out [T][n_rows][P], n_rows 25 or 37.
int spio_traj_sens(void* hp, const double* par, const int* idx, int P, const float* init37, const float* actions,
                   int n_steps, int n_rows, double* y_out, double* J_out) {
  if (P < 1 || P > P_COUNT || n_rows < 1 || n_rows > 37) return -1;
  auto* h = static_cast<Handle<double>*>(hp);
  Engine<DualP> eng;
  dual_engine(h, eng);
  DualP p[P_COUNT];
  dual_params(par, idx, P, p);
  eng.apply_base_params(p);
  State<DualP> st;
  load37(st, init37);
  for (int t = 0; t < n_steps; ++t) {
    DualP a[12];
    for (int j = 0; j < 12; ++j) a[j] = DualP(double(actions[12 * t + j]));
    eng.control_step(st, a, p);
    for (int i = 0; i < n_rows; ++i) {
      const DualP& y = i < 13 ? st.root[i] : (i < 25 ? st.q[i - 13] : st.qd[i - 25]);
      y_out[size_t(t) * n_rows + i] = y.v;
      for (int m = 0; m < P; ++m) J_out[(size_t(t) * n_rows + i) * P + m] = y.d[m];
    }
  }
  return finite_state(st) ? 0 : 1;
}

// spio_traj_sens over n_traj trajectories in parallel (OpenMP; n_threads <= 0: the runtime's
// default): init37 [n_traj][37], actions [n_traj][n_steps][12], y_out [n_traj][n_steps][n_rows],
// J_out [n_traj][n_steps][n_rows][P].  Returns the number of trajectories that diverged.
int spio_traj_sens_batch(void* hp, const double* par, const int* idx, int P, const float* init37,
                         const float* actions, int n_traj, int n_steps, int n_rows, double* y_out, double* J_out,
                         int n_threads) {
  if (P < 1 || P > P_COUNT || n_rows < 1 || n_rows > 37 || n_traj < 0) return -1;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
  for (int i = 0; i < n_traj; ++i) {
    const size_t ys = size_t(n_steps) * n_rows;
    bad += spio_traj_sens(hp, par, idx, P, init37 + size_t(i) * 37, actions + size_t(i) * n_steps * 12, n_steps,
                          n_rows, y_out + i * ys, J_out + i * ys * P) != 0;
  }
  return bad;
}

// spio_simulate of one parameter vector from n_traj initial states in parallel (OpenMP):
// init37 [n_traj][37], actions [n_traj][n_steps][12], states_out [n_traj][n_steps][37] in double
// (root 13, q 12, qd 12 after each control step).  Returns the number of trajectories that left
// the finite range.
int spio_simulate_batch(void* hp, const double* par, const float* init37, const float* actions, int n_traj,
                        int n_steps, double* states_out, int n_threads) {
  auto* h = static_cast<Handle<double>*>(hp);
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
  int bad = 0;
#pragma omp parallel reduction(+ : bad)
  {
    Engine<double> eng = h->eng;
    eng.apply_base_params(par);
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < n_traj; ++i) {
      State<double> st;
      load37(st, init37 + size_t(i) * 37);
      for (int t = 0; t < n_steps; ++t) {
        double a[12];
        for (int j = 0; j < 12; ++j) a[j] = actions[(size_t(i) * n_steps + t) * 12 + j];
        eng.control_step(st, a, par);
        double* o = states_out + (size_t(i) * n_steps + t) * 37;
        store_state(st, o, o + 13, o + 25);
      }
      bad += !finite_state(st);
    }
  }
  return bad;
}

// FP64 rollout cost of n_cand parameter vectors on n_seg segments with the (candidate, segment)
// pairs spread over OpenMP threads; same arguments and result as spio_rollout_cost.
int spio_rollout_cost_par(void* hp, const double* par, int n_cand, const float* seg_init, const float* seg_actions,
                          const float* seg_targets, int n_seg, int H, int cost_mode, const double* w,
                          double* cost_out, int n_threads) {
  auto* h = static_cast<Handle<double>*>(hp);
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
  int bad = 0;
  const int total = n_cand * n_seg;
#pragma omp parallel reduction(+ : bad)
  {
    Engine<double> eng = h->eng;  // per-thread copy (base inertia is overwritten per candidate)
#pragma omp for schedule(dynamic, 2)
    for (int lane = 0; lane < total; ++lane) {
      const int c = lane / n_seg, s = lane % n_seg;
      const double* p = par + P_COUNT * c;
      eng.apply_base_params(p);
      State<double> st;
      load37(st, seg_init + 37 * s);
      double acc = 0.0;
      for (int t = 0; t < H; ++t) {
        double a[12];
        for (int j = 0; j < 12; ++j) a[j] = seg_actions[(size_t(s) * H + t) * 12 + j];
        eng.control_step(st, a, p);
        const float* tg = seg_targets + (size_t(s) * H + t) * 37;
        if (cost_mode == COST_TRACKING || t == H - 1) acc += step_cost<double>(cost_mode, w, st, tg);
      }
      if (!finite_state(st) || !std::isfinite(acc)) { acc = INFINITY; ++bad; }
      cost_out[size_t(c) * n_seg + s] = acc;
    }
  }
  return bad;
}

// Cost of one parameter vector on n_seg segments and its gradient with respect to idx[]:
// cost_out [n_seg], grad_out [n_seg][P].
int spio_cost_grad(void* hp, const double* par, const int* idx, int P, const float* seg_init, const float* seg_actions,
                   const float* seg_targets, int n_seg, int H, int cost_mode, const double* w, double* cost_out,
                   double* grad_out) {
  if (P < 1 || P > P_COUNT) return -1;
  auto* h = static_cast<Handle<double>*>(hp);
  Engine<DualP> eng;
  dual_engine(h, eng);
  DualP p[P_COUNT];
  dual_params(par, idx, P, p);
  eng.apply_base_params(p);
  int bad = 0;
  for (int s = 0; s < n_seg; ++s) {
    State<DualP> st;
    load37(st, seg_init + 37 * s);
    DualP acc(0.0);
    for (int t = 0; t < H; ++t) {
      DualP a[12];
      for (int j = 0; j < 12; ++j) a[j] = DualP(double(seg_actions[(size_t(s) * H + t) * 12 + j]));
      eng.control_step(st, a, p);
      const float* tg = seg_targets + (size_t(s) * H + t) * 37;
      if (cost_mode == COST_TRACKING || t == H - 1) acc += step_cost<DualP>(cost_mode, w, st, tg);
    }
    const bool ok = finite_state(st) && std::isfinite(acc.v);
    cost_out[s] = ok ? acc.v : INFINITY;
    for (int m = 0; m < P; ++m) grad_out[size_t(s) * P + m] = ok ? acc.d[m] : 0.0;
    bad += !ok;
  }
  return bad;
}

// ---------------------------------------------------------------- RNG and CEM restatement
void spio_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  Philox4 c;
  for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  Philox4 r = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

double spio_philox_normal(unsigned long long seed, unsigned int iteration, unsigned long long idx) {
  return philox_normal(seed, iteration, idx);
}

// One CEM iteration on the host: sample theta[p][k] = clamp(mean + std z), evaluate is done by the
// caller; here only sampling into theta [n_param][n].
void spio_cem_sample(const double* mean, const double* stdv, const double* lo, const double* hi, int n_param, int n,
                     unsigned long long seed, unsigned int iteration, double* theta) {
  for (int p = 0; p < n_param; ++p)
    for (int k = 0; k < n; ++k) {
      double z = philox_normal(seed, iteration, (unsigned long long)p * n + k);
      double v = mean[p] + stdv[p] * z;
      theta[(size_t)p * n + k] = std::min(std::max(v, lo[p]), hi[p]);
    }
}

int spio_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
