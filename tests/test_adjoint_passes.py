"""Each neighbour-pass VJP of dimcv_adj_pass against autograd.

For every pass with a neighbour sum (density, DFSPH factor, density change,
Jacobi update, non-pressure, curl, grad v, smoothing, AKDE, near field) the
test writes the pass's forward formula in torch float64 over the pair list
the GPU built, differentiates a random linear functional of it with
autograd, and compares the gradients with the hand-written VJP (f32) on the
same inputs.  The ring lattice starts at half a spacing from the walls, so
eight fluid particles sit in cell 0, which no neighbour list contains (the
reference's visiting rule): their pair terms reach their neighbours only
through the VJPs' scatter path, and the test checks the particles around
cell 0 on their own as well.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "dimcv-sph_amd"))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def scene():
    import torch
    import scenes
    import _native as N
    ps, solver, info = scenes.build_ring_system(12, 12, 10, n_rings=2, seed=3, mc_samples=128)
    ps.ensure_neighbor_lists()
    torch.cuda.synchronize()
    n, nfl = ps._pn, ps._nfl
    fl = ps.fluid_list.t[:nfl].long().cpu().numpy()
    cnt = ps.nbr_count.t[:nfl].cpu().numpy()
    nb = ps.nbr[:, :nfl].cpu().numpy().view(np.uint32)
    I, J = [], []
    for f in range(nfl):
        for k in range(min(int(cnt[f]), ps.nbr_kmax)):
            e = int(nb[k, f])
            assert e >> 31, "ring scenes hold fluid only"
            I.append(fl[f])
            J.append(e & 0x7FFFFFFF)
    gpn0 = int(ps.grid_particles_num.t[0].item())
    x = ps.x.t[:n].double().cpu().numpy()
    cell0 = np.where(np.arange(n) < gpn0)[0]
    assert len(cell0) >= 1, "the scene must hold fluid in cell 0"
    assert not np.isin(cell0, np.asarray(J)).any(), "cell-0 particles appear in no list"
    # particles within two support radii of the origin: the cell-0 particles' neighbourhood
    near0 = np.where(np.linalg.norm(x, axis=1) < 2.5 * float(ps.params.support_radius))[0]
    return dict(ps=ps, N=N, n=n, nfl=nfl, I=np.asarray(I), J=np.asarray(J), near0=near0)


def _kernels(p, torch):
    k, dk, ih = float(p.kernel_k), float(p.kernel_dk), float(p.inv_support_radius)

    def W(s):
        q = s * ih
        lo = k * (6 * q ** 3 - 6 * q ** 2 + 1)
        hi = k * 2 * (1 - q) ** 3
        return torch.where(q <= 0.5, lo, torch.where(q <= 1.0, hi, torch.zeros_like(q)))

    def gW(r):
        s = torch.sqrt((r * r).sum(-1))
        q = s * ih
        c = torch.where(q <= 0.5, dk * q * (3 * q - 2), -dk * (1 - q) ** 2)
        c = torch.where((s > 1e-5) & (q <= 1.0), c, torch.zeros_like(c))
        return (c * ih / s)[:, None] * r

    def E(r):
        q2 = (r * ih) ** 2
        e = torch.where(q2 < 1.0, 0.75 * (1 - q2), torch.zeros_like(q2))
        return e.prod(-1)

    def A(s):
        eps = float(p.eps)
        Ee = torch.exp(-s / eps)
        return (1 - Ee) / (s * (s + eps * Ee))

    return W, gW, E, A


def _run(S, pass_id, f0=None, f1=None, f2=None, g_in=None, mode=0, outs=("x_bar",), widths=None, dscalar=0.0):
    import torch
    ps, N, n = S["ps"], S["N"], S["n"]
    cap = int(ps.particle_max_num)
    dev = ps.x.t.device
    widths = widths or {}
    bufs = {}
    for o in outs:
        w = widths.get(o, 3 if o == "x_bar" else 1)
        bufs[o] = torch.zeros((cap, w), dtype=torch.float32, device=dev)

    def dev32(a):
        if a is None:
            return None
        t = torch.zeros((cap,) + a.shape[1:], dtype=torch.float32, device=dev)
        t[:n] = torch.as_tensor(a, dtype=torch.float32, device=dev)
        return t
    ins = [dev32(a) for a in (f0, f1, f2, g_in)]
    args = N.AdjArgs()
    ptr = lambda t: t.data_ptr() if t is not None else None
    args.f0, args.f1, args.f2, args.g_in = (ptr(t) for t in ins)
    for o in ("x_bar", "b0", "b1", "b2"):
        setattr(args, o, ptr(bufs.get(o)))
    args.dscalar, args.mode = dscalar, mode
    st = ps.state()
    rc = N.lib().dimcv_adj_pass(ctypes.byref(st), ctypes.byref(ps.params), pass_id, ctypes.byref(args),
                                S["nfl"], torch.cuda.current_stream().cuda_stream)
    N.check(rc, "dimcv_adj_pass")
    torch.cuda.synchronize()
    return {o: b[:n].double().cpu().numpy() for o, b in bufs.items()}


def _compare(S, name, got, ref, tol=2e-5):
    near0 = S["near0"]
    e_all = _rel(got, ref)
    e_near = _rel(got[near0], ref[near0])
    print(f"{name}: rel L2 {e_all:.2e} (around cell 0: {e_near:.2e})")
    assert e_all < tol and e_near < tol, (name, e_all, e_near)


def _setup(S, seed):
    import torch
    ps = S["ps"]
    n = S["n"]
    rng = np.random.default_rng(seed)
    x = torch.tensor(ps.x.t[:n].double().cpu().numpy(), requires_grad=True)
    I = torch.as_tensor(S["I"])
    J = torch.as_tensor(S["J"])
    mV = torch.as_tensor(ps.m_V.t[:n].double().cpu().numpy())
    m = torch.as_tensor(ps.m.t[:n].double().cpu().numpy())
    W, gW, E, A = _kernels(ps.params, torch)
    return rng, x, I, J, mV, m, W, gW, E, A


def test_density_factor_and_div_vjps(scene):
    import torch
    S, N, p = scene, scene["N"], scene["ps"].params
    n = S["n"]
    rng, x, I, J, mV, m, W, gW, E, A = _setup(S, 1)
    rho0 = float(p.density0)
    # density: rho_i = rho0 (mV_i W(0) + sum_j mV_j W(|r_ij|))
    rb = rng.normal(size=n)
    r = x[I] - x[J]
    rho = rho0 * (mV * W(torch.zeros(n, dtype=torch.float64)) +
                  torch.zeros(n, dtype=torch.float64).index_add(0, I, mV[J] * W(torch.sqrt((r * r).sum(-1)))))
    (gx,) = torch.autograd.grad((rho * torch.as_tensor(rb)).sum(), x)
    got = _run(S, N.ADJ_DENSITY, g_in=rb)
    _compare(S, "density x_bar", got["x_bar"], gx.numpy())
    # DFSPH factor: alpha_i = 1 / (sum_j |m_j gW|^2 + |sum_j m_j gW|^2)
    g = m[J][:, None] * gW(x[I] - x[J])
    s1 = torch.zeros(n, dtype=torch.float64).index_add(0, I, (g * g).sum(-1))
    G = torch.zeros((n, 3), dtype=torch.float64).index_add(0, I, g)
    den = s1 + (G * G).sum(-1)
    alpha = torch.where(den > 1e-6, 1.0 / den, torch.zeros_like(den))
    ab = rng.normal(size=n)
    (gx,) = torch.autograd.grad((alpha * torch.as_tensor(ab)).sum(), x)
    got = _run(S, N.ADJ_FACTOR, f0=alpha.detach().numpy(), g_in=ab)
    _compare(S, "factor x_bar", got["x_bar"], gx.numpy(), tol=1e-4)
    # density change (mode 0): b_i = [b_i > 0] sum_j mV_j (v_i - v_j).gW_ij
    v = torch.tensor(rng.normal(size=(n, 3)), requires_grad=True)
    gate = rng.normal(size=n)
    bb = rng.normal(size=n)
    dv = ((v[I] - v[J]) * gW(x[I] - x[J])).sum(-1) * mV[J]
    out = torch.zeros(n, dtype=torch.float64).index_add(0, I, dv) * torch.as_tensor((gate > 0).astype(np.float64))
    gx, gv = torch.autograd.grad((out * torch.as_tensor(bb)).sum(), (x, v))
    got = _run(S, N.ADJ_DIV, f0=v.detach().numpy(), f1=gate, g_in=bb, mode=0, outs=("x_bar", "b1"),
               widths={"b1": 3})
    _compare(S, "div x_bar", got["x_bar"], gx.numpy())
    _compare(S, "div v_bar", got["b1"], gv.numpy())


def test_jacobi_nonpressure_curl_gradv_vjps(scene):
    import torch
    S, N, p = scene, scene["N"], scene["ps"].params
    n = S["n"]
    rng, x, I, J, mV, m, W, gW, E, A = _setup(S, 2)
    rho0, dt = float(p.density0), float(p.dt)
    # Jacobi update (mode 1): P_i = sum_{|k_i + k_j| > eps} dt (k_i + k_j) m_j gW_ij, k = (da - rho0) alpha
    da = (rho0 + rng.normal(scale=2.0, size=n)).astype(np.float32)
    al = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
    k32 = (da - np.float32(rho0)) * al
    keep = np.abs(k32[S["I"]] + k32[S["J"]]) > np.float32(p.m_eps)
    kk = torch.tensor(k32.astype(np.float64), requires_grad=True)
    Ik, Jk = I[torch.as_tensor(keep)], J[torch.as_tensor(keep)]
    P = torch.zeros((n, 3), dtype=torch.float64).index_add(
        0, Ik, (dt * (kk[Ik] + kk[Jk]) * m[Jk])[:, None] * gW(x[Ik] - x[Jk]))
    pb = rng.normal(size=(n, 3))
    gx, gk = torch.autograd.grad((P * torch.as_tensor(pb)).sum(), (x, kk))
    got = _run(S, N.ADJ_JACOBI, f0=da, f1=al, g_in=pb, mode=1, outs=("x_bar", "b0"))
    _compare(S, "jacobi x_bar", got["x_bar"], gx.numpy())
    _compare(S, "jacobi kappa_bar", got["b0"][:, 0], gk.numpy())
    # non-pressure: a_i = sum_j c mV_j ((v_i - v_j).r) / (|r|^2 + 0.01 h^2) gW_ij
    v = torch.tensor(rng.normal(size=(n, 3)), requires_grad=True)
    r = x[I] - x[J]
    c = float(p.visc_coef)
    coef = c * mV[J] * ((v[I] - v[J]) * r).sum(-1) / ((r * r).sum(-1) + float(p.visc_h2))
    a = torch.zeros((n, 3), dtype=torch.float64).index_add(0, I, coef[:, None] * gW(r))
    abar = rng.normal(size=(n, 3))
    gx, gv = torch.autograd.grad((a * torch.as_tensor(abar)).sum(), (x, v))
    got = _run(S, N.ADJ_NONPRESSURE, f0=v.detach().numpy(), g_in=abar, outs=("x_bar", "b1"), widths={"b1": 3})
    _compare(S, "nonpressure x_bar", got["x_bar"], gx.numpy())
    _compare(S, "nonpressure v_bar", got["b1"], gv.numpy())
    # curl: C_i = sum_j mV_j (u_j - u_i) x gW_ij
    u = torch.tensor(rng.normal(size=(n, 3)), requires_grad=True)
    C = torch.zeros((n, 3), dtype=torch.float64).index_add(
        0, I, mV[J][:, None] * torch.cross(u[J] - u[I], gW(x[I] - x[J]), dim=-1))
    cb = rng.normal(size=(n, 3))
    gx, gu = torch.autograd.grad((C * torch.as_tensor(cb)).sum(), (x, u))
    got = _run(S, N.ADJ_CURL, f0=u.detach().numpy(), g_in=cb, outs=("x_bar", "b1"), widths={"b1": 3})
    _compare(S, "curl x_bar", got["x_bar"], gx.numpy())
    _compare(S, "curl u_bar", got["b1"], gu.numpy())
    # grad v: G_i[r][c] = sum_j mV_j (v_j - v_i)_r gW_ij_c
    g = gW(x[I] - x[J])
    Gm = torch.zeros((n, 3, 3), dtype=torch.float64).index_add(
        0, I, mV[J][:, None, None] * (v[J] - v[I])[:, :, None] * g[:, None, :])
    Gb = rng.normal(size=(n, 9))
    gx, gv = torch.autograd.grad((Gm.reshape(n, 9) * torch.as_tensor(Gb)).sum(), (x, v))
    got = _run(S, N.ADJ_GRAD_V, f0=v.detach().numpy(), g_in=Gb, outs=("x_bar", "b1"), widths={"b1": 3})
    _compare(S, "grad_v x_bar", got["x_bar"], gx.numpy())
    _compare(S, "grad_v v_bar", got["b1"], gv.numpy())


def test_smooth_kden_near_vjps(scene):
    import torch
    S, N, p = scene, scene["N"], scene["ps"].params
    n = S["n"]
    rng, x, I, J, mV, m, W, gW, E, A = _setup(S, 3)
    zero = torch.zeros(n, dtype=torch.float64)
    # smoothing: w_i = mV_i d_i W(0) + sum_j mV_j d_j W(|r_ij|)
    d = torch.tensor(rng.normal(size=(n, 3)), requires_grad=True)
    r = x[I] - x[J]
    w = (mV * W(zero))[:, None] * d + torch.zeros((n, 3), dtype=torch.float64).index_add(
        0, I, (mV[J] * W(torch.sqrt((r * r).sum(-1))))[:, None] * d[J])
    wb = rng.normal(size=(n, 3))
    gx, gd = torch.autograd.grad((w * torch.as_tensor(wb)).sum(), (x, d))
    got = _run(S, N.ADJ_SMOOTH, f0=d.detach().numpy(), g_in=wb, outs=("x_bar", "b1"), widths={"b1": 3})
    _compare(S, "smooth x_bar", got["x_bar"], gx.numpy())
    _compare(S, "smooth d_bar", got["b1"], gd.numpy())
    # AKDE: kden_i = kvn_i E(0) + sum_j kvn_j E(r_ij)
    kvn = torch.tensor(rng.uniform(0.0, 2.0, size=n), requires_grad=True)
    kd = kvn * E(torch.zeros((1, 3), dtype=torch.float64)) + zero.index_add(0, I, kvn[J] * E(x[I] - x[J]))
    kb = rng.normal(size=n)
    gx, gk = torch.autograd.grad((kd * torch.as_tensor(kb)).sum(), (x, kvn))
    got = _run(S, N.ADJ_KDEN, f0=kvn.detach().numpy(), g_in=kb, outs=("x_bar", "b1"))
    _compare(S, "kden x_bar", got["x_bar"], gx.numpy())
    _compare(S, "kden kvn_bar", got["b1"][:, 0], gk.numpy())
    # near field: vv_i = fac sum_j A(|r_xy|) (r_xy x w_j); S_i = K - kden_i - sum_j kden_j
    wv = torch.tensor(rng.normal(size=(n, 3)), requires_grad=True)
    kden = torch.tensor(rng.uniform(0.5, 2.0, size=n), requires_grad=True)
    r = x[I] - x[J]
    rxy = torch.cat([r[:, :2], torch.zeros_like(r[:, 2:])], dim=1)
    s = torch.sqrt((rxy * rxy).sum(-1))
    vv = float(p.vp_factor) * torch.zeros((n, 3), dtype=torch.float64).index_add(
        0, I, A(s)[:, None] * torch.cross(rxy, wv[J], dim=-1))
    Sn = -(kden + zero.index_add(0, I, kden[J]))
    ub = rng.normal(size=(n, 3))
    sb = rng.normal(size=n)
    loss = (vv * torch.as_tensor(ub)).sum() + (Sn * torch.as_tensor(sb)).sum()
    gx, gw, gk = torch.autograd.grad(loss, (x, wv, kden))
    got = _run(S, N.ADJ_NEAR, f0=wv.detach().numpy(), f1=sb, g_in=ub, outs=("x_bar", "b0", "b1"),
               widths={"b0": 3})
    _compare(S, "near x_bar", got["x_bar"], gx.numpy())
    _compare(S, "near w_bar", got["b0"], gw.numpy())
    _compare(S, "near kden_bar", got["b1"][:, 0], gk.numpy())
