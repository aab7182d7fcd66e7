"""Policy-in-the-loop active exploration: a locomotion policy tracks each sampled command sequence
while the finite-difference Fisher information of (1 main + P aux) env groups accumulates.

Reference: ``ActiveSysId.evaluate_policy`` (spigym/agents/sysid/active_sysid.py:527-600) over
``ActiveSysId_OpenLoop`` (spigym/envs/sysid/active_sysid_openloop.py) with the go2_omni actor
(observation config/obs/loco/go2_omni.yaml, MLP 512-256-128 ELU, config/algo/ppo.yaml:32-47).
Everything per step runs on the device: ``spi_mlp_forward`` (actor), ``spi_policy_step``
(physics), ``spi_policy_reward`` (termination + FIM reward), ``spi_policy_obs`` (gait clock,
observation, history).  The whole rollout can be captured once as a HIP graph and replayed.

Step order (one reset step with zero actions, then total_steps - 2 policy steps,
base_task.py:90-100 and active_sysid.py:539-576):
  * the gait clock of a step uses the commands the step ran with, the observation the commands
    after the playback pointer advanced (go2_omni.py:348-379, active_sysid_openloop.py:196-201);
  * termination (gravity, group OR) and the FIM reward use the post-step state; a terminated
    group earns ``termination_rew`` and gets zero actions at the next step;
  * aux envs are synced to their main env after steps where step_idx % ksync == 1
    (every step for ksync <= 1), step_idx counting the reset step as 1;
  * the result is the sum of rewards over the policy steps divided by total_steps - 1.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, Optional, Sequence

import numpy as np
import torch

from .. import _lib
from ..params import PARAM_NAMES
from ..rollout import C_ref, RolloutEngine
from .active import KSYNC_MODES, _alias

# Up to this many envs the four-lanes-per-env physics runs faster (one env per lane leaves
# SIMDs idle below ~64 K envs and has ~3x the instructions per lane on the critical path).
LEG_PARALLEL_MAX_ENVS = 16384
OBS_TERMS = 60
OBS_HISTORY = 14
ACTOR_HIDDEN = (512, 256, 128)


class PolicyMLP:
    """The actor network on the device: Linear layers with ELU between them, weights packed per
    layer as W [out][in] then b [out] (the torch Linear layout)."""

    def __init__(self, layers: Sequence[tuple], device="cuda"):
        _lib.require_gpu()
        self.device = torch.device(device)
        dims = [int(layers[0][0].shape[1])]
        parts = []
        for W, b in layers:
            W = torch.as_tensor(W, dtype=torch.float32)
            b = torch.as_tensor(b, dtype=torch.float32)
            if W.dim() != 2 or W.shape[1] != dims[-1] or b.shape != (W.shape[0],):
                raise ValueError("layer shapes do not chain")
            dims.append(int(W.shape[0]))
            parts += [W.reshape(-1), b]
        self.dims = dims
        self._dims_c = (C.c_int * len(dims))(*dims)
        self.n_layers = len(dims) - 1
        n = _lib.lib().spi_mlp_param_floats(self.n_layers, self._dims_c)
        if n < 0:
            raise ValueError("unsupported layer stack")
        self.params = torch.cat([p.cpu() for p in parts]).to(self.device).contiguous()
        assert self.params.numel() == n
        self._work: Optional[torch.Tensor] = None

    @classmethod
    def from_state_dict(cls, sd: Dict[str, torch.Tensor], prefix: str = "actor_module.module.", device="cuda"):
        """From the reference's actor state dict (``loaded_dict["actor_model_state_dict"]``,
        ppo.py:138-143): Linear layers at the even indices of the nn.Sequential."""
        idx = sorted({int(k[len(prefix):].split(".")[0]) for k in sd if k.startswith(prefix) and k.endswith("weight")})
        if not idx:
            raise KeyError(f"no '{prefix}*.weight' entries in the state dict")
        return cls([(sd[f"{prefix}{i}.weight"], sd[f"{prefix}{i}.bias"]) for i in idx], device)

    @classmethod
    def from_checkpoint(cls, path: str, device="cuda"):
        """Load a reference PPO checkpoint (tensors only: weights_only=True)."""
        ck = torch.load(path, map_location="cpu", weights_only=True)
        return cls.from_state_dict(ck["actor_model_state_dict"], device=device)

    @classmethod
    def random(cls, in_dim: int, out_dim: int = 12, hidden: Sequence[int] = ACTOR_HIDDEN, seed: int = 0,
               device="cuda"):
        """Random weights with torch.nn.Linear's default initialisation."""
        g = torch.Generator().manual_seed(seed)
        dims = [in_dim, *hidden, out_dim]
        layers = []
        for i, o in zip(dims[:-1], dims[1:]):
            bound = 1.0 / np.sqrt(i)
            layers.append(((torch.rand(o, i, generator=g) * 2 - 1) * bound, (torch.rand(o, generator=g) * 2 - 1) * bound))
        return cls(layers, device)

    def layers(self):
        """(W, b) per layer, views into the packed device buffer."""
        out, off = [], 0
        for i, o in zip(self.dims[:-1], self.dims[1:]):
            W = self.params[off:off + o * i].view(o, i)
            off += o * i
            out.append((W, self.params[off:off + o]))
            off += o
        return out

    def _work_for(self, M: int) -> torch.Tensor:
        need = _lib.lib().spi_mlp_work_floats(M, self.n_layers, self._dims_c)
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(max(need, 1), dtype=torch.float32, device=self.device)
        return self._work

    def work_floats(self, M: int) -> int:
        return int(_lib.lib().spi_mlp_work_floats(M, self.n_layers, self._dims_c))

    def forward_ring(self, ring: torch.Tensor, head: int, kmap: torch.Tensor, clip: float, M: int,
                     out: Optional[torch.Tensor] = None, row0: int = 0,
                     work: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Input rows row0 .. row0 + M - 1 gathered from a term ring [n_slots, n_terms, ld]
        (spi_mlp_forward_ring); work: a private scratch buffer (for concurrent calls)."""
        S, n_terms, ld = ring.shape
        if kmap.shape != (self.dims[0], 2) or kmap.dtype != torch.int32:
            raise ValueError(f"kmap must be int32 [{self.dims[0]}, 2]")
        if not (ring.is_cuda and ring.is_contiguous() and ring.dtype == torch.float32) or row0 + M > ld:
            raise ValueError("ring must be a contiguous float32 GPU tensor covering the rows")
        if work is None:
            work = self._work_for(M)
        if work.numel() < self.work_floats(M):
            raise ValueError("work buffer too small")
        if out is None:
            out = torch.empty((M, self.dims[-1]), dtype=torch.float32, device=self.device)
        rc = _lib.lib().spi_mlp_forward_ring(ring.data_ptr() + 4 * row0, ld, S, n_terms, int(head),
                                             _lib.ptr(kmap, torch.int32), float(clip), M, self.n_layers,
                                             self._dims_c, _lib.ptr(self.params), _lib.ptr(work), work.numel(),
                                             _lib.ptr(out), _lib.stream())
        _lib.check(rc, "spi_mlp_forward_ring")
        return out

    def __call__(self, x: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        if x.dim() != 2 or x.shape[1] != self.dims[0]:
            raise ValueError(f"input must be [M, {self.dims[0]}]")
        M = x.shape[0]
        work = self._work_for(M)
        if out is None:
            out = torch.empty((M, self.dims[-1]), dtype=torch.float32, device=self.device)
        rc = _lib.lib().spi_mlp_forward(_lib.ptr(x), M, self.n_layers, self._dims_c, _lib.ptr(self.params),
                                        _lib.ptr(work), work.numel(), _lib.ptr(out), _lib.stream())
        _lib.check(rc, "spi_mlp_forward")
        return out


def reset_states(n_env: int, q_default: np.ndarray, seed: int = 0, base_height: float = 0.34,
                 vel_range: float = 0.5) -> np.ndarray:
    """Reset states drawn like the reference's reset of every env: joint positions
    q_default * U[0.5, 1.5], zero joint velocities (active_sysid_openloop.py:274-292); base at
    (0, 0, base_height), identity orientation, linear and angular velocity U[-0.5, 0.5]
    (legged_robot_base.py:761-784, go2.yaml init_state).  [n_env, 37] float32."""
    rng = np.random.default_rng(seed)
    s = np.zeros((n_env, 37), np.float32)
    s[:, 2] = base_height
    s[:, 6] = 1.0
    s[:, 7:13] = rng.uniform(-vel_range, vel_range, size=(n_env, 6))
    s[:, 13:25] = np.asarray(q_default, np.float32)[None] * rng.uniform(0.5, 1.5, size=(n_env, 12))
    return s


def reset_states_device(n_env: int, q_default: np.ndarray, seed: int = 0, base_height: float = 0.34,
                        vel_range: float = 0.5, q_range=(0.5, 1.5), origins: Optional[torch.Tensor] = None,
                        out: Optional[torch.Tensor] = None, device="cuda") -> torch.Tensor:
    """The reset draw of ``reset_states`` on the device (``spi_reset_states``): joint positions
    q_default * U[0.5, 1.5], zero joint velocities (active_sysid_openloop.py:274-292), base at
    (0, 0, base_height) (+ origins [n_env, 3]) with identity orientation and linear and angular
    velocity U[-0.5, 0.5] (legged_robot_base.py:761-784).  Philox4x32-10 under ``seed``: env e's
    draw depends only on (seed, e), so a draw can be repeated for any subset.  [n_env, 37]."""
    spec = _lib.SpiResetSpec()
    spec.q_default[:] = [float(x) for x in np.asarray(q_default, np.float64).reshape(12)]
    spec.root0[:] = [0.0, 0.0, float(base_height), 0.0, 0.0, 0.0, 1.0]
    spec.q_lo, spec.q_hi = float(q_range[0]), float(q_range[1])
    spec.vel_range = float(vel_range)
    if out is None:
        out = torch.empty((n_env, 37), dtype=torch.float32, device=device)
    if out.shape != (n_env, 37) or not out.is_contiguous() or out.dtype != torch.float32:
        raise ValueError(f"out must be a contiguous float32 [{n_env}, 37] tensor")
    if origins is not None and (origins.shape != (n_env, 3) or not origins.is_contiguous()
                                or origins.dtype != torch.float32):
        raise ValueError(f"origins must be a contiguous float32 [{n_env}, 3] tensor")
    rc = _lib.lib().spi_reset_states(C.byref(spec), n_env, _lib.ptr(origins), int(seed), _lib.ptr(out),
                                     _lib.stream())
    _lib.check(rc, "spi_reset_states")
    return out


class PolicyFIMRollout:
    """Scores command sequences by the FD Fisher information of a policy-driven rollout."""

    def __init__(self, engine: RolloutEngine, policy: PolicyMLP, theta: np.ndarray, explore: Sequence[str],
                 delta: float = 0.1, ksync_steps: int = 5, ksync_mode: str = "reference",
                 termination_rew: float = 0.0, gravity_limits=(0.8, 0.8), height_limits=(-0.25, 0.15),
                 obs_clip: float = 100.0, evaluating: bool = True, obs_layout: str = "ring", n_streams: int = 1,
                 physics: str = "auto"):
        if obs_layout not in ("ring", "dense"):
            raise ValueError("obs_layout must be 'ring' or 'dense'")
        self.obs_layout = obs_layout
        # one env per lane, or four lanes per env (one per leg) when the batch leaves SIMDs idle
        if physics not in ("auto", "lanes", "legs"):
            raise ValueError("physics must be 'auto', 'lanes' or 'legs'")
        self.physics = physics
        # independent slices of the groups on their own streams: one slice's physics (latency-
        # bound, a few wavefronts per SIMD) overlaps another's actor GEMMs (matrix cores)
        self.n_streams = max(1, int(n_streams)) if obs_layout == "ring" else 1
        self._streams = [torch.cuda.Stream() for _ in range(self.n_streams - 1)]
        self.engine = engine
        self.policy = policy
        obs_dim = _lib.lib().spi_policy_obs_dim()
        if policy.dims[0] != obs_dim or policy.dims[-1] != 12:
            raise ValueError(f"policy must map {obs_dim} observations to 12 actions")
        self.obs_dim = obs_dim
        self.theta = np.asarray(theta, np.float64)
        self.idx = [PARAM_NAMES.index(p) if p in PARAM_NAMES else _alias(p) for p in explore]
        self.P = len(self.idx)
        self.G = self.P + 1
        self.delta = float(delta)
        self.ksync = int(ksync_steps)
        self.ksync_mode = KSYNC_MODES[ksync_mode]
        self.termination_rew = float(termination_rew)
        self.grav = tuple(float(g) for g in gravity_limits)
        self.height = tuple(float(h) for h in height_limits)
        self.obs_clip = float(obs_clip)
        self.evaluating = int(bool(evaluating))
        sim = engine.sim
        self.dt_ctrl = float(sim.dt * sim.decimation)
        self.q_default = torch.as_tensor(np.asarray(sim.q0, np.float32), device=engine.device)
        self._bufs: Optional[dict] = None
        self._graph = None

    # ------------------------------------------------------------------ layout
    def params(self, n_group: int) -> torch.Tensor:
        """SoA [25, n_env]: aux env k carries theta[idx[k-1]] + delta (active_sysid_openloop.py:62-68)."""
        rows = np.repeat(self.theta[None], n_group * self.G, axis=0)
        for k, p in enumerate(self.idx):
            rows[k + 1::self.G, p] += self.delta
        return torch.from_numpy(np.ascontiguousarray(rows.T.astype(np.float32))).to(self.engine.device)

    def sync_after(self, step_idx: int) -> bool:
        """active_sysid_openloop.py:246-252."""
        return self.ksync <= 1 or step_idx % self.ksync == 1

    def _alloc(self, M: int, T: int) -> dict:
        b = self._bufs
        if b is not None and b["M"] == M and b["T"] == T:
            return b
        dev, n = self.engine.device, M * self.G
        f = lambda *s: torch.zeros(s, dtype=torch.float32, device=dev)
        i = lambda *s: torch.zeros(s, dtype=torch.int32, device=dev)
        b = {"M": M, "T": T, "n": n, "par": self.params(M), "s": [f(n, 37), f(n, 37)], "init": f(n, 37),
             "actions": f(n, 12), "applied": f(n, 12), "done": i(n), "bad": i(n), "gait": f(n), "acc": f(M),
             "gbad": i(M), "cmds": f(M, T, 14)}
        if self.obs_layout == "ring":
            L = _lib.lib()
            S = L.spi_policy_ring_slots()
            km = (C.c_int * (2 * self.obs_dim))()
            _lib.check(L.spi_policy_obs_kmap(km), "spi_policy_obs_kmap")
            b["ring"] = f(S, OBS_TERMS, (n + 3) // 4 * 4)   # rows padded to 4: float4 gathers
            b["kmap"] = torch.tensor(list(km), dtype=torch.int32).view(self.obs_dim, 2).to(dev)
        else:
            b["hist"] = f(n, OBS_HISTORY, OBS_TERMS)
            b["obs"] = f(n, self.obs_dim)
        b["parts"] = self._parts(M)
        b["work"] = [f(max(self.policy.work_floats((g1 - g0) * self.G), 1)) for g0, g1 in b["parts"]]
        self._bufs = b
        self._graph = None
        return b

    def _parts(self, M: int):
        """Group ranges of the stream slices; slice starts are multiples of 4 envs when possible
        (16-byte aligned ring rows)."""
        G, P = self.G, self.n_streams
        unit = 4 // math.gcd(G, 4)
        cuts = sorted({0, M} | {min(M, int(round(k * M / P / unit)) * unit) for k in range(1, P)})
        return [(int(a), int(b)) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]

    # ------------------------------------------------------------------ the rollout
    def _run(self, b: dict) -> None:
        """Every launch of one rollout from b['init'] and b['cmds'] (capturable).  Each slice of
        groups runs its own chain of launches on its own stream; the slices never exchange data."""
        L = _lib.lib()
        n, M, T, G = b["n"], b["M"], b["T"], self.G
        a = C_ref(self.engine.args)
        par = b["par"]
        ld = par.shape[1]
        s = b["s"]
        main = torch.cuda.current_stream()
        s[0].copy_(b["init"])
        ring = self.obs_layout == "ring"
        for k in ("actions", "done", "bad", "gait", "acc", "gbad") + (("ring",) if ring else ("hist",)):
            b[k].zero_()
        S = b["ring"].shape[0] if ring else 0
        parts = b["parts"]
        streams = [main] + self._streams[:len(parts) - 1]
        for x in streams[1:]:
            x.wait_stream(main)
        i32 = torch.int32

        legs = self.physics == "legs" or (self.physics == "auto" and n <= LEG_PARALLEL_MAX_ENVS)
        step_fn = L.spi_policy_step_legs if legs else L.spi_policy_step

        def physics(p: int, cur: int, sync_first: bool) -> None:
            g0, g1 = parts[p]
            e0, e1 = g0 * G, g1 * G
            rc = step_fn(a, par.data_ptr() + 4 * e0, ld, e1 - e0, G, _lib.ptr(s[cur][e0:e1]),
                                   _lib.ptr(s[1 - cur][e0:e1]), _lib.ptr(b["actions"][e0:e1]),
                                   _lib.ptr(b["done"][e0:e1], i32), int(sync_first), _lib.ptr(b["applied"][e0:e1]),
                                   _lib.ptr(b["bad"][e0:e1], i32), _lib.stream())
            _lib.check(rc, "spi_policy_step")

        def reward(p: int, cur: int) -> None:
            g0, g1 = parts[p]
            e0, e1 = g0 * G, g1 * G
            rc = L.spi_policy_reward(_lib.ptr(s[cur][e0:e1]), g1 - g0, G, self.delta, self.grav[0], self.grav[1],
                                     self.termination_rew, _lib.ptr(b["done"][e0:e1], i32),
                                     _lib.ptr(b["bad"][e0:e1], i32), _lib.ptr(b["acc"][g0:g1]),
                                     _lib.ptr(b["gbad"][g0:g1], i32), _lib.stream())
            _lib.check(rc, "spi_policy_reward")

        def observe(p: int, cur: int, now: int) -> None:
            g0, g1 = parts[p]
            e0, e1 = g0 * G, g1 * G
            if ring:                     # step `now` writes ring slot now % S
                rg = b["ring"]
                rc = L.spi_policy_terms(e1 - e0, G, _lib.ptr(s[cur][e0:e1]), _lib.ptr(b["cmds"][g0:g1]), T, now,
                                        now + 1, self.dt_ctrl, _lib.ptr(self.q_default), _lib.ptr(b["gait"][e0:e1]),
                                        _lib.ptr(b["applied"][e0:e1]), rg.data_ptr() + 4 * e0, rg.shape[2], now % S,
                                        self.evaluating, self.height[0], self.height[1], _lib.stream())
                _lib.check(rc, "spi_policy_terms")
                return
            rc = L.spi_policy_obs(e1 - e0, G, _lib.ptr(s[cur][e0:e1]), _lib.ptr(b["cmds"][g0:g1]), T, now, now + 1,
                                  self.dt_ctrl, _lib.ptr(self.q_default), _lib.ptr(b["gait"][e0:e1]),
                                  _lib.ptr(b["applied"][e0:e1]), _lib.ptr(b["hist"][e0:e1]),
                                  _lib.ptr(b["obs"][e0:e1]), self.evaluating, self.height[0], self.height[1],
                                  self.obs_clip, _lib.stream())
            _lib.check(rc, "spi_policy_obs")

        def reward_observe(p: int, cur: int, now: int) -> None:
            if not ring:
                reward(p, cur)
                observe(p, cur, now)
                return
            g0, g1 = parts[p]              # one launch: group rewards + this step's ring slot
            e0, e1 = g0 * G, g1 * G
            rg = b["ring"]
            rc = L.spi_policy_reward_terms(
                e1 - e0, G, _lib.ptr(s[cur][e0:e1]), _lib.ptr(b["cmds"][g0:g1]), T, now, now + 1, self.dt_ctrl,
                _lib.ptr(self.q_default), _lib.ptr(b["gait"][e0:e1]), _lib.ptr(b["applied"][e0:e1]),
                rg.data_ptr() + 4 * e0, rg.shape[2], now % S, self.evaluating, self.height[0], self.height[1],
                self.delta, self.grav[0], self.grav[1], self.termination_rew, _lib.ptr(b["done"][e0:e1], i32),
                _lib.ptr(b["bad"][e0:e1], i32), _lib.ptr(b["acc"][g0:g1]), _lib.ptr(b["gbad"][g0:g1], i32),
                _lib.stream())
            _lib.check(rc, "spi_policy_reward_terms")

        def act(p: int, latest: int) -> None:
            g0, g1 = parts[p]
            e0, e1 = g0 * G, g1 * G
            if ring:
                self.policy.forward_ring(b["ring"], latest % S, b["kmap"], self.obs_clip, e1 - e0,
                                         out=b["actions"][e0:e1], row0=e0, work=b["work"][p])
            else:
                self.policy(b["obs"][e0:e1], out=b["actions"][e0:e1])

        def sync(p: int, cur: int, step_idx: int) -> bool:
            if not self.sync_after(step_idx):
                return False
            if self.ksync_mode == KSYNC_MODES["state"]:
                g0, g1 = parts[p]
                x = s[cur][g0 * G:g1 * G]
                _lib.check(L.spi_policy_sync_state(_lib.ptr(x), (g1 - g0) * G, G, _lib.stream()),
                           "spi_policy_sync_state")
                return False
            return True

        pending = [False] * len(parts)
        for p in range(len(parts)):
            with torch.cuda.stream(streams[p]):
                # reset step: zero actions, no reward (base_task.py:90-100)
                physics(p, 0, False)
                observe(p, 1, 0)
                pending[p] = sync(p, 1, 1)
        cur = 1
        for it in range(1, T - 1):
            for p in range(len(parts)):
                with torch.cuda.stream(streams[p]):
                    act(p, it - 1)
                    physics(p, cur, pending[p])
                    reward_observe(p, 1 - cur, it)
                    pending[p] = sync(p, 1 - cur, it + 1)
            cur = 1 - cur
        for x in streams[1:]:
            main.wait_stream(x)
        b["final"] = s[cur]

    def evaluate(self, commands: torch.Tensor, init: torch.Tensor, graph: bool = False) -> torch.Tensor:
        """commands [M, total_steps, 14] (one sequence per main env); init [M * (P + 1), 37] (or
        [37]).  Returns the average per-step reward of each group's main env [M] (the value
        active_sysid.py:210-213 tells the study, negated); NaN where a group diverged."""
        if commands.dim() != 3 or commands.shape[2] != 14 or commands.shape[1] < 3:
            raise ValueError("commands must be [M, total_steps >= 3, 14]")
        M, T = commands.shape[0], commands.shape[1]
        b = self._alloc(M, T)
        if init.dim() == 1:
            init = init[None].expand(b["n"], 37)
        if init.shape != (b["n"], 37):
            raise ValueError(f"init must be [{b['n']}, 37]")
        b["cmds"].copy_(commands)
        b["init"].copy_(init)
        if graph:
            if self._graph is None:
                self._graph = self._capture(b)
            self._graph.replay()
        else:
            self._run(b)
        fim = b["acc"] / float(T - 1)
        return torch.where(b["gbad"] != 0, torch.full_like(fim, float("nan")), fim)

    def _capture(self, b: dict):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._run(b)            # warm-up outside the graph (allocations, first launches)
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._run(b)
        return g

    def final_states(self) -> torch.Tensor:
        return self._bufs["final"]

    def flags(self):
        """(done [n_env], bad [n_env], group_bad [M]) after the last evaluation."""
        b = self._bufs
        return b["done"], b["bad"], b["gbad"]


class PolicyActiveSysId:
    """``ActiveSysId.optimize`` (active_sysid.py:165-257) with every stage on the device: each
    iteration samples ``batch_size`` trials of command parameters (device CEM, the role of the
    Optuna sampler), expands them into command sequences (``CommandGenerator``, the role of
    ``sample_commands``), scores them with a policy-driven FIM rollout (``evaluate_policy``) and
    updates the CEM with the negated FIM (``tell(-total_reward)``).  The host only reads the
    best trial at the end."""

    def __init__(self, rollout: PolicyFIMRollout, generator, batch_size: int = 1024, elite_frac: float = 0.1,
                 seed: int = 0, init: Optional[torch.Tensor] = None, graph: bool = True, num_rollouts: int = 1):
        """init: fixed initial states [batch_size * (P + 1), 37] for every evaluation, or None
        (the reference's behaviour): every rollout starts from a fresh reset draw on the device,
        as evaluate_policy resets the envs before each of its num_rollouts_per_command rollouts
        (active_sysid.py:540-543) and averages their rewards (:594-597)."""
        from .cem import CEMConfig, DeviceCEM
        self.rollout, self.gen, self.graph = rollout, generator, graph
        bounds = np.asarray(generator.param_bounds(), np.float64)
        lo, hi = bounds[:, 0], bounds[:, 1]
        self.names = generator.param_names()
        self.cem = DeviceCEM(0.5 * (lo + hi), 0.5 * (hi - lo), lo, hi, np.maximum(1e-3 * (hi - lo), 0.0),
                             CEMConfig(n_samples=batch_size, elite_frac=elite_frac, seed=seed))
        if num_rollouts < 1:
            raise ValueError("num_rollouts must be >= 1")
        self.num_rollouts = int(num_rollouts)
        self.seed = int(seed)
        self.n_env = batch_size * rollout.G
        self.init = init
        self._reset_buf = None
        self.draws = 0                  # reset draws so far; draw k uses Philox key reset_key(k)
        self.history = []

    def reset_key(self, draw: int) -> int:
        """Philox key of reset draw ``draw``: the optimizer's seed in the low word, the draw in the high."""
        return (self.seed & 0xFFFFFFFF) | (int(draw) << 32)

    def _initial_states(self) -> torch.Tensor:
        if self.init is not None:
            return self.init
        if self._reset_buf is None:
            self._reset_buf = torch.empty((self.n_env, 37), dtype=torch.float32, device=self.rollout.engine.device)
        reset_states_device(self.n_env, self.rollout.engine.sim.q0, seed=self.reset_key(self.draws),
                            out=self._reset_buf)
        self.draws += 1
        return self._reset_buf

    def step(self) -> torch.Tensor:
        theta = self.cem.ask()                                 # [n_params, B]
        cmds = self.gen.generate(theta.t())                    # [B, total_steps, 14]
        fim = None
        for _ in range(self.num_rollouts):
            f = self.rollout.evaluate(cmds, self._initial_states(), graph=self.graph)
            fim = f.clone() if fim is None else fim + f
        fim = fim / float(self.num_rollouts)
        cost = torch.nan_to_num(-fim, nan=float("inf"))
        self.cem.tell(cost)
        self.history.append(fim)
        return fim

    def optimize(self, iterations: int, log_dir: Optional[str] = None) -> Dict:
        """``iterations`` optimizer iterations; the best trial's parameters, value (the negated
        FIM, what the reference's study minimises) and command sequence.  With ``log_dir`` the
        sequence is also saved as ``log_dir/best_commands.npz`` (key ``best_commands``), as
        active_sysid.py:229-231 does."""
        import os
        for _ in range(iterations):
            self.step()
        best = self.cem.best.cpu().numpy().astype(np.float64)
        params = dict(zip(self.names, best.tolist()))
        out = {"best_params": params, "best_value": float(self.cem.best_cost.item()),
               "best_commands": self.gen.from_params(params)}
        if log_dir is not None:
            os.makedirs(log_dir, exist_ok=True)
            path = os.path.join(log_dir, "best_commands.npz")
            np.savez(path, best_commands=out["best_commands"])
            out["best_commands_path"] = path
        return out


# params_dict keys of the active-sysID env (env/active_sysid_openloop.yaml default_param) -> parameter
# rows; "inertiay" has no setter in the reference (isaacgym_active_sysid.py:86 defines
# set_inertiaiy), so it is applied only with apply_inertiay.
def _param_rows(apply_inertiay: bool):
    from ..params import P_COMX, P_COMY, P_COMZ, P_IXX, P_IYY, P_IZZ, P_MASS, P_MOTOR0
    rows = {"mass": P_MASS, "comx": P_COMX, "comy": P_COMY, "comz": P_COMZ, "inertiax": P_IXX, "inertiaz": P_IZZ,
            "motor_model_hip_a": P_MOTOR0, "motor_model_thigh_a": P_MOTOR0 + 1, "motor_model_calf_a": P_MOTOR0 + 2}
    if apply_inertiay:
        rows["inertiay"] = P_IYY
    return rows


def theta_from_default_param(default_param: Dict, nominal: Optional[np.ndarray] = None,
                             apply_inertiay: bool = False) -> np.ndarray:
    """The 25-row parameter vector of every env before exploration offsets: the nominal model with
    each ``default_param`` entry (``[body, value]`` or ``[value]``) written to its row."""
    from ..params import nominal_params
    theta = np.array(nominal_params() if nominal is None else nominal, np.float64)
    rows = _param_rows(apply_inertiay)
    for key, spec in default_param.items():
        spec = list(spec)
        if len(spec) == 2 and spec[0] != "base":
            raise NotImplementedError(f"default_param {key!r}: only the base body is identified")
        if key in rows:
            theta[rows[key]] = float(spec[-1])
    return theta


def active_sysid_from_config(cfg: Dict, policy: PolicyMLP, engine: Optional[RolloutEngine] = None,
                             seed: int = 0) -> PolicyActiveSysId:
    """Builds the device optimizer from the reference's composed config (num_envs,
    env.config.{delta_param, ksync_steps, default_param, exploration_params},
    algo.config.command.*, algo.config.optimize.sampler.batch_size), as ``run_active_sysid.py``
    instantiates env and algo from it.  The batch is num_envs / (P + 1) trials."""
    import copy
    from ..params import SimParams
    from .commands import CommandGenerator
    env, algo = cfg["env"]["config"], cfg["algo"]["config"]
    cmd = algo["command"]
    explore = list(env.get("exploration_params", ["mass"]))
    G = len(explore) + 1
    if engine is None:
        sim = copy.deepcopy(SimParams())
        sim.control_mode, sim.motor_model = 1, 3
        engine = RolloutEngine(sim)
    apply_y = bool(cfg.get("simulator", {}).get("config", {}).get("apply_inertiay", False))
    theta = theta_from_default_param(env.get("default_param", {}), apply_inertiay=apply_y)
    rows = _param_rows(apply_y)
    names = [PARAM_NAMES[rows[p]] if p in rows else p for p in explore]
    roll = PolicyFIMRollout(engine, policy, theta, names, delta=float(env.get("delta_param", 0.1)),
                            ksync_steps=int(env.get("ksync_steps", 1)),
                            termination_rew=float(cmd.get("termination_rew", 0.0)))
    dt = engine.sim.dt * engine.sim.decimation
    gen = CommandGenerator(cmd["command_ranges"], cmd["default_command"], cmd["command_sampling_idxs"],
                           cmd.get("command_sampling_mode", "constant"), float(cmd["rollout_length"]),
                           float(cmd["horizon_length"]), dt, int(cmd.get("poly_degree", 3)),
                           int(cmd.get("num_bezier_points", 4)))
    batch = int(cfg.get("num_envs", 2048)) // G
    sampler = algo.get("optimize", {}).get("sampler", {})
    if "batch_size" in sampler:
        batch = int(sampler["batch_size"])
    return PolicyActiveSysId(roll, gen, batch_size=batch, seed=seed,
                             num_rollouts=int(cmd.get("num_rollouts_per_command", 1)))
