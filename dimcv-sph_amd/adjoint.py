"""Differentiation through vortex-ring rollouts (build-defined; the reference
would use ``ti.ad.Tape``, which needs Taichi).

``RingRollout`` owns one ring scene whose ring circulations Gamma are the
parameters: the initial velocity is v0 = sum_k Gamma_k U_k(x0), with U_k the
FP64 velocity field of ring k at unit circulation.  ``rollout(gamma)`` runs
K steps of the native step (``KarmanVortexSolver.step``) from v0 and returns
the final velocities in particle-id order; it is a ``torch.autograd.Function``
whose backward is the adjoint of the rollout.

The adjoint is that of the step with its discrete structure frozen at the
forward's values (neighbour lists, importance samples, Jacobi iteration
counts and every gate; see csrc/adjoint.hip).  The forward keeps a
checkpoint of the whole state every ``checkpoint`` steps.  The backward
replays each segment from its checkpoint through the per-pass forward,
recording every intermediate the VJPs read (the tape), then runs the
hand-written VJP kernels (``dimcv_adj_pass``) backwards through the segment.
Adjoints travel between steps in particle-id order (``particle_id``), so the
grid build's permutation is undone exactly.  Finally
dL/dGamma_k = sum_p v0_bar_p . U_k(p) in FP64.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

import _native as N
from _field import Field
from scenes import (add_ring_particles, jittered_lattice, random_rings, ring_points, ring_scene_config,
                    ring_velocity)


def _ptr(t):
    return t.data_ptr() if t is not None else None


class StepTape:
    """Forward values of one step in the step's slot order (fluid only)."""

    def state(self, ps) -> N.State:
        st = N.State.from_buffer_copy(ps.state())
        st.x = self.x.data_ptr()
        st.m = self.m.data_ptr()
        st.m_V = self.m_V.data_ptr()
        st.fluid_list = self.fluid_list.data_ptr()
        st.nbr_count = self.nbr_count.data_ptr()
        st.nbr = self.nbr.data_ptr()
        st.sample_idx = self.sample_idx.data_ptr()
        st.counters = self.counters.data_ptr()
        st.grid_particles_num = self.gpn0.data_ptr()
        return st


def _clone(field, n):
    return field.t[:n].clone()


def record_step(ps, sv, u_cdf) -> StepTape:
    """One KarmanVortexSolver.step through the per-pass entry points (the
    sequence the native substep fuses, bit-identical to it), recording the
    tape.  Ring scenes: fluid only, no emitters, fixed-M sampling."""
    ps.initialize_particle_system()
    n = ps._pn
    if ps._nfl != n or ps.solid_particle_num[None] != 0:
        raise RuntimeError("the adjoint covers fluid-only scenes")
    T = StepTape()
    T.n = n
    T.ids = ps.particle_id.t[:n].clone().long()
    T.x, T.m, T.m_V = _clone(ps.x, n), ps.m.t.clone(), ps.m_V.t.clone()
    T.fluid_list = ps.fluid_list.t.clone()
    T.nbr_count = ps.nbr_count.t.clone()
    T.nbr = ps.nbr.clone()
    T.gpn0 = ps.grid_particles_num.t[:1].clone()        # cell 0's particles: slots [0, gpn0)
    sv.compute_densities()
    sv.compute_DFSPH_factor()
    T.rho, T.alpha0 = _clone(ps.density, n), _clone(ps.dfsph_factor, n)
    # divergence_solve (dimcv_sph.py:376-393)
    sv.compute_density_change()
    T.vd, T.b = [_clone(ps.v, n)], [_clone(ps.density_adv, n)]
    sv.multiply_time_step(ps.dfsph_factor, sv.inv_dt)
    T.alpha_div = _clone(ps.dfsph_factor, n)
    it = 0
    while it < 1 or it < sv.m_max_iterations_v:
        err = sv.divergence_solver_iteration()
        T.vd.append(_clone(ps.v, n))
        T.b.append(_clone(ps.density_adv, n))
        it += 1
        if err <= sv.inv_dt * sv.max_error_V * 0.01 * sv.density_0:
            break
    T.kv = it
    sv.multiply_time_step(ps.dfsph_factor, sv.dt[None])
    sv.compute_non_pressure_forces()
    T.a = _clone(ps.acceleration, n)
    T.v_div = T.vd[-1]
    # dimcv() (dimcv_sph.py:812-819)
    sv.init()
    T.v_star = _clone(sv.v_star, n)
    T.G = _clone(sv.grad_v, n)
    sv.compute_d_vorticity()
    T.d_vort = _clone(sv.d_vort, n)
    sv.compute_all_d_vort_smoothed()
    T.ws = _clone(sv.d_vort_smoothed, n)
    sv.update_sample()
    T.kvn = _clone(sv.kinematic_vort_num, n)
    sv.compute_kernel_density()
    T.kden = _clone(sv.kernel_density, n)
    T.K = float(sv.kden_sum[None])
    sv.importance_sample(u_cdf)
    T.sample_idx = sv.sample_idx.t.clone()
    T.M = int(sv.num_samples[None])
    sv.vort2vel()
    T.v_vort = _clone(sv.v_vort, n)
    sv.update_particle_life_time()
    sv.predict_velocity()
    # pressure_solve (dimcv_sph.py:597-614)
    sv.compute_density_adv()
    T.vp, T.da = [_clone(ps.v, n)], [_clone(ps.density_adv, n)]
    sv.multiply_time_step(ps.dfsph_factor, sv.inv_dt2)
    T.alpha_p = _clone(ps.dfsph_factor, n)
    it = 0
    while it < 1 or it < sv.m_max_iterations:
        err = sv.pressure_solve_iteration()
        T.vp.append(_clone(ps.v, n))
        T.da.append(_clone(ps.density_adv, n))
        it += 1
        if err <= sv.max_error * 0.01 * sv.density_0:
            break
    T.kp = it
    sv.copy_x_temp()
    sv.compute_vorticity_vis()
    sv.advect()
    T.counters = ps._counters.clone()
    return T


class _Adj:
    """Issues dimcv_adj_pass calls against one tape."""

    def __init__(self, ps, T):
        self.ps, self.T = ps, T
        self.st = T.state(ps)
        self.stream = torch.cuda.current_stream().cuda_stream

    def __call__(self, pass_id, f0=None, f1=None, f2=None, g_in=None, x_bar=None, b0=None, b1=None, b2=None,
                 dscalar=0.0, mode=0):
        a = N.AdjArgs()
        a.f0, a.f1, a.f2, a.g_in = _ptr(f0), _ptr(f1), _ptr(f2), _ptr(g_in)
        a.x_bar, a.b0, a.b1, a.b2 = _ptr(x_bar), _ptr(b0), _ptr(b1), _ptr(b2)
        a.dscalar, a.mode = float(dscalar), int(mode)
        rc = N.lib().dimcv_adj_pass(ctypes.byref(self.st), ctypes.byref(self.ps.params), int(pass_id),
                                    ctypes.byref(a), int(self.T.n), self.stream)
        N.check(rc, "dimcv_adj_pass")


def step_vjp(ps, sv, T: StepTape, xb_id: torch.Tensor, vb_id: torch.Tensor):
    """Adjoint of one step: adjoints of (x', v') after the step -> adjoints of
    (x, v) before it, both in particle-id order."""
    n, dt, rho0 = T.n, float(np.float32(sv.dt[None])), float(np.float32(sv.density_0))
    dev = T.x.device
    z1 = lambda: torch.zeros(n, dtype=torch.float32, device=dev)
    z3 = lambda: torch.zeros((n, 3), dtype=torch.float32, device=dev)
    adj = _Adj(ps, T)
    xb = xb_id[T.ids].contiguous()
    vb = (vb_id[T.ids] + dt * xb).contiguous()          # advect: x' = x + dt v
    rho_b, alpha_p_b = z1(), z1()
    # pressure iterations, last to first
    for t in reversed(range(T.kp)):
        kb = z1()
        adj(N.ADJ_JACOBI, f0=T.da[t], f1=T.alpha_p, g_in=(-vb).contiguous(), x_bar=xb, b0=kb, mode=1)
        alpha_p_b += kb * (T.da[t] - rho0)
        adj(N.ADJ_DIV, f0=T.vp[t], f1=T.da[t], g_in=(kb * T.alpha_p).contiguous(), x_bar=xb, b1=vb, b0=rho_b,
            mode=1)
    # predict: v_p = v_vv + dt a;  vort2vel: v_vv = v_div + gate * v_vort
    ab = dt * vb
    gate = z1()
    adj(N.ADJ_GATE, f0=T.v_vort, f1=T.v_div, b2=gate)
    ub = (gate[:, None] * vb).contiguous()
    S_b, wts, ws_b, kden_b = z1(), z3(), z3(), z1()
    adj(N.ADJ_FAR_TARGETS, f0=T.ws, f1=T.kden, g_in=ub, x_bar=xb, b0=S_b, b2=wts, dscalar=T.K, mode=T.M)
    adj(N.ADJ_FAR_SOURCES, f0=T.ws, f1=T.kden, f2=wts, x_bar=xb, b0=ws_b, b1=kden_b, mode=T.M)
    adj(N.ADJ_NEAR, f0=T.ws, f1=S_b, g_in=ub, x_bar=xb, b0=ws_b, b1=kden_b)
    kden_b += S_b.double().sum().float()                  # S_i = kden_sum - ...
    kvn_b, G_b = z1(), torch.zeros((n, 9), dtype=torch.float32, device=dev)
    adj(N.ADJ_KDEN, f0=T.kvn, g_in=kden_b, x_bar=xb, b1=kvn_b)
    adj(N.ADJ_KVN, f0=T.G, g_in=kvn_b, b1=G_b)
    vdiv_b = vb.clone()
    adj(N.ADJ_GRAD_V, f0=T.v_div, g_in=G_b, x_bar=xb, b1=vdiv_b)
    dv_b = z3()
    adj(N.ADJ_SMOOTH, f0=T.d_vort, g_in=ws_b, x_bar=xb, b1=dv_b)
    # d_vort = curl(v_div) - curl(v*), v* = v_div + dt a
    adj(N.ADJ_CURL, f0=T.v_div, g_in=dv_b, x_bar=xb, b1=vdiv_b)
    vs_b = z3()
    adj(N.ADJ_CURL, f0=T.v_star, g_in=(-dv_b).contiguous(), x_bar=xb, b1=vs_b)
    vdiv_b += vs_b
    ab += dt * vs_b
    adj(N.ADJ_NONPRESSURE, f0=T.v_div, g_in=ab.contiguous(), x_bar=xb, b1=vdiv_b)
    # divergence iterations, last to first
    vb = vdiv_b
    alpha_div_b = z1()
    for t in reversed(range(T.kv)):
        kb = z1()
        adj(N.ADJ_JACOBI, f0=T.b[t], f1=T.alpha_div, f2=T.rho, g_in=(-vb).contiguous(), x_bar=xb, b0=kb, mode=0)
        alpha_div_b += kb * T.b[t] * T.rho
        rho_b += kb * T.b[t] * T.alpha_div
        adj(N.ADJ_DIV, f0=T.vd[t], f1=T.b[t], g_in=(kb * T.alpha_div * T.rho).contiguous(), x_bar=xb, b1=vb,
            mode=0)
    # alpha_div = alpha0 inv_dt, alpha_p = alpha0 inv_dt dt inv_dt2 (f32 scalars)
    f32 = np.float32
    c_div = float(f32(sv.inv_dt))
    c_p = float(f32(sv.inv_dt)) * float(f32(sv.dt[None])) * float(f32(sv.inv_dt2))
    alpha0_b = (c_div * alpha_div_b + c_p * alpha_p_b).contiguous()
    adj(N.ADJ_FACTOR, f0=T.alpha0, g_in=alpha0_b, x_bar=xb)
    adj(N.ADJ_DENSITY, g_in=rho_b, x_bar=xb)
    out_x = torch.zeros_like(xb_id)
    out_v = torch.zeros_like(vb_id)
    out_x[T.ids] = xb
    out_v[T.ids] = vb
    return out_x, out_v


class RingRollout:
    """A ring scene with circulations as parameters and its rollout adjoint."""

    def __init__(self, nx, ny, nz, n_rings=16, steps=32, checkpoint=8, mc_samples=1024, r=0.01, dt=2e-4,
                 viscosity=1e-3, seed=5, K=256, far_field="fast", spacing=1.0, pad=0, device="cuda"):
        """nx x ny x nz jittered lattice of spacing 2 r * spacing, `pad`
        lattice spacings inside every wall of a (n + 2 pad) 2r box (spacing < 1
        compresses the fluid, so that the pressure solve acts)."""
        from config_builder import SimConfig
        from particle_system import ParticleSystem
        self.steps, self.checkpoint = int(steps), int(checkpoint)
        rng = np.random.default_rng(seed)
        pos = (jittered_lattice(nx, ny, nz, r * spacing, rng) + np.float32(pad * 2 * r)).astype(np.float32)
        self.pos = pos
        box = np.array([nx + 2 * pad, ny + 2 * pad, nz + 2 * pad], np.float64) * 2 * r
        centres, axes, radii, gammas = random_rings(n_rings, box, rng)
        self.gamma0 = np.asarray(gammas, np.float64)
        sigma = 4 * r
        xt = torch.as_tensor(pos, device=device)
        # per-ring unit-circulation velocity fields, FP64: U[k] (n, 3)
        U = []
        for k in range(n_rings):
            tab = ring_points(centres[k:k + 1], axes[k:k + 1], radii[k:k + 1], [1.0], K=K)
            pts = torch.as_tensor(tab, dtype=torch.float64, device=device)
            U.append(self._ring_field_f64(xt, pts, sigma))
        self.U = torch.stack(U)                              # (K, n, 3) FP64
        scene = ring_scene_config(nx + 2 * pad, ny + 2 * pad, nz + 2 * pad, r=r, dt=dt, mc_samples=mc_samples,
                                  far_field=far_field, viscosity=viscosity, seed=seed)
        ps = ParticleSystem(SimConfig(scene=scene), device=device)
        v0 = self.initial_velocity(self.gamma0)
        add_ring_particles(ps, pos, v0.cpu().numpy())
        self.ps, self.sv = ps, ps.build_solver()
        self.sv.initialize()
        self.n = ps._pn
        self.pristine = self._snapshot()
        ur = np.random.default_rng(seed + 1)
        self.u_cdf = [ur.random(int(mc_samples)) for _ in range(self.steps)]
        self.last_tapes = None

    @staticmethod
    def _ring_field_f64(x, pts, sigma):
        """FP64 ring field (the dimcv_ring_velocity sum without the f32 cast)."""
        X = x.double()
        out = torch.zeros_like(X)
        inv_s2 = 1.0 / (sigma * sigma)
        for c in range(0, pts.shape[0], 64):
            q = pts[c:c + 64]
            r = X[:, None, :] - q[None, :, 0:3]
            r2 = (r * r).sum(-1)
            core = 1.0 - torch.exp(-r2 * inv_s2)
            f = q[None, :, 6] * core / (r2 * torch.sqrt(r2))
            out += (f[..., None] * torch.cross(q[None, :, 3:6].expand_as(r), r, dim=-1)).sum(1)
        return out

    def initial_velocity(self, gamma) -> torch.Tensor:
        g = torch.as_tensor(np.asarray(gamma, np.float64), device=self.U.device)
        # sum_k g_k U_k as a broadcast sum (a GEMM call for this shape is slow)
        return (g[:, None, None] * self.U).sum(0).float()

    # ------------------------------------------------------------ state --
    def _tensors(self):
        ps, sv = self.ps, self.sv
        out = [f.t for f in vars(ps).values() if isinstance(f, Field)]
        out += [f.t for f in vars(sv).values() if isinstance(f, Field)]
        return out + [ps._counters, ps._dscal]

    def _snapshot(self):
        return {"t": [(t, t.clone()) for t in self._tensors()],
                "fields": {name: (getattr(self.ps, name).t, getattr(self.ps, name + "_buffer").t)
                           for name in N.SORTED_FIELDS},
                "pn": self.ps._pn, "nfl": self.ps._nfl}

    def _restore(self, snap):
        ps = self.ps
        for name, (a, b) in snap["fields"].items():      # undo buffer swaps
            getattr(ps, name).t, getattr(ps, name + "_buffer").t = a, b
        for t, c in snap["t"]:
            t.copy_(c)
        ps._pn, ps._nfl = snap["pn"], snap["nfl"]
        ps.invalidate_state()
        ps._nbr_valid = False

    def _set_initial(self, gamma):
        self._restore(self.pristine)
        ids = self.ps.particle_id.t[: self.n].long()
        self.ps.v.t[: self.n] = self.initial_velocity(gamma)[ids]

    def final_velocity(self) -> torch.Tensor:
        ids = self.ps.particle_id.t[: self.ps._pn].long()
        out = torch.zeros((self.n, 3), dtype=torch.float32, device=self.U.device)
        out[ids] = self.ps.v.t[: self.ps._pn]
        return out

    # ---------------------------------------------------------- rollout --
    def forward(self, gamma) -> torch.Tensor:
        """v after `steps` steps from v0(gamma), id order; keeps checkpoints."""
        self._set_initial(gamma)
        self.ckpt = {}
        for k in range(self.steps):
            if k % self.checkpoint == 0:
                self.ckpt[k] = self._snapshot()
            self.sv.step(u_cdf=self.u_cdf[k])
        if self.ps._pn != self.n:
            raise RuntimeError("particles left the box; the adjoint needs a fixed particle set")
        return self.final_velocity()

    def backward(self, v_bar: torch.Tensor) -> torch.Tensor:
        """dL/dgamma (FP64) from dL/dv_final (id order)."""
        vb = v_bar.float().contiguous().clone()
        xb = torch.zeros_like(vb)
        starts = sorted(self.ckpt)
        self.last_tapes = []
        for s0 in reversed(starts):
            s1 = min(s0 + self.checkpoint, self.steps)
            self._restore(self.ckpt[s0])
            tapes = [record_step(self.ps, self.sv, self.u_cdf[k]) for k in range(s0, s1)]
            for T in reversed(tapes):
                xb, vb = step_vjp(self.ps, self.sv, T, xb, vb)
            self.last_tapes = tapes + self.last_tapes
        self.v0_bar = vb
        return (vb.double()[None] * self.U).sum(dim=(1, 2))

    def structure(self, gamma):
        """Run the rollout from gamma recording every step's tape (the frozen
        structure the adjoint linearises about)."""
        self._set_initial(gamma)
        tapes = [record_step(self.ps, self.sv, self.u_cdf[k]) for k in range(self.steps)]
        return tapes


class _RolloutFn(torch.autograd.Function):

    @staticmethod
    def forward(ctx, gamma, model):
        ctx.model = model
        return model.forward(gamma.detach().cpu().numpy())

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.model.backward(grad_out)
        return g.to(dtype=torch.float64, device="cpu"), None


def rollout(gamma: torch.Tensor, model: RingRollout) -> torch.Tensor:
    """Differentiable rollout: final velocities (id order) as a function of
    the ring circulations."""
    return _RolloutFn.apply(gamma, model)
