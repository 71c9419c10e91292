"""Rollout adjoint against central finite differences.

The adjoint (hand-written VJP kernels, csrc/adjoint.hip, behind a
torch.autograd.Function with checkpoints) differentiates the rollout with
every step's discrete structure frozen.  The oracle's FP64 frozen rollout
(oracle/frozen.cpp, pinned pass by pass in test_frozen_oracle.py) evaluates
that same function; its central differences (eps = 1e-3 in the ring
circulations) must match the adjoint gradient to 1e-4 relative.  The scene
is compressed so that the pressure solve acts, and the rollout spans two
checkpoint segments.
"""

import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu


def host_tape(T):
    c = lambda t: t.detach().cpu().numpy()
    return {"n": T.n, "kv": T.kv, "kp": T.kp, "M": T.M, "ids": c(T.ids).astype(np.int32), "nbr": c(T.nbr),
            "nbr_count": c(T.nbr_count), "x": c(T.x), "rho": c(T.rho), "alpha_div": c(T.alpha_div),
            "alpha_p": c(T.alpha_p), "b": [c(b) for b in T.b], "da": [c(d) for d in T.da], "v_vort": c(T.v_vort),
            "v_div": c(T.v_div), "sample_idx": c(T.sample_idx)}


def frozen_params(model):
    p, sv, ps = model.ps.params, model.sv, model.ps
    f = np.float32
    return dict(dt=float(p.dt), rho0=float(sv.density_0), mV=float(ps.m_V.t[0]), m=float(ps.m.t[0]),
                h=float(p.support_radius), kernel_k=float(p.kernel_k), kernel_dk=float(p.kernel_dk),
                visc_coef=float(p.visc_coef), visc_h2=float(p.visc_h2), eps=float(p.eps), vp_factor=float(p.vp_factor),
                inv_dt=float(f(sv.inv_dt)), inv_dt2=float(f(sv.inv_dt2)), far_t=float(p.far_r2_max_rejected),
                m_eps=float(p.m_eps), rho0_f=float(sv.density_0))


def check_gradient(model, components=None, eps=1e-3, seed=3):
    import torch
    from adjoint import rollout
    from oracle.frozen import FrozenRollout
    g0 = model.gamma0.copy()
    gt = g0 + np.random.default_rng(seed).normal(0.0, 0.1, g0.shape)
    v_target = model.forward(gt).double()
    gamma = torch.tensor(g0, dtype=torch.float64, requires_grad=True)
    v = rollout(gamma, model)
    loss = ((v.double() - v_target) ** 2).sum()
    loss.backward()
    g_adj = gamma.grad.numpy()
    v32 = v.detach().double().cpu().numpy()
    fr = FrozenRollout([host_tape(T) for T in model.structure(g0)], frozen_params(model))
    U = model.U.cpu().numpy()
    x0 = model.pos.astype(np.float64)
    vt = v_target.cpu().numpy()
    run = lambda g: fr.run(x0, np.einsum("k,knd->nd", g, U))
    v64 = run(g0)
    model_err = H.rel_l2(v64, v32)
    L = lambda g: float(((run(g) - vt) ** 2).sum())
    comps = range(len(g0)) if components is None else components
    g_fd = []
    for k in comps:
        e = np.zeros_like(g0)
        e[k] = eps
        g_fd.append((L(g0 + e) - L(g0 - e)) / (2 * eps))
    g_fd = np.array(g_fd)
    ga = g_adj[list(comps)]
    return model_err, H.rel_l2(ga, g_fd), ga, g_fd


def test_rollout_adjoint_matches_central_differences():
    from adjoint import RingRollout
    model = RingRollout(14, 14, 12, n_rings=4, steps=6, checkpoint=4, mc_samples=96, spacing=0.9, pad=3, seed=7)
    model_err, err, ga, gf = check_gradient(model)
    print(f"adjoint: model {model_err:.3e}, gradient rel err {err:.3e}, adjoint {ga}, fd {gf}")
    assert model_err < 2e-5, model_err          # the frozen FP64 rollout is the forward
    tapes = model.last_tapes
    assert any(T.kp > 1 or T.da[0].max().item() > 1000.0 for T in tapes)  # the pressure solve acted
    assert err < 1e-4, (err, ga, gf)


def test_full_size_rollout_adjoint_matches_central_differences():
    """The adjoint workload at its full size (65,536 particles, 16 rings, 32
    steps, checkpoint 8, M = 1024) against central differences of the FP64
    rollout for two circulations (about 25 s on one GPU)."""
    from adjoint import RingRollout
    model = RingRollout(64, 32, 32, n_rings=16, steps=32, checkpoint=8, mc_samples=1024, seed=5, pad=2)
    model_err, err, ga, gf = check_gradient(model, components=[0, 7])
    print(f"full-size adjoint: model {model_err:.3e}, gradient rel err {err:.3e}, adjoint {ga}, fd {gf}")
    assert model_err < 1e-4 and err < 1e-4, (model_err, err, ga, gf)
