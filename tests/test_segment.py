"""Segment solver Biot-Savart (segment_solver.py:378-418, 468-542): pins of
the oracle's restatement by what the mathematics fixes, and the GPU's
parity with it.

* A long straight filament cut into short collinear segments induces, at
  distance rho from its middle, the velocity of the Biot-Savart law for a
  finite straight line, Gamma / (4 pi rho) * 2L / sqrt(L^2 + rho^2), along
  t x e_rho (right-hand rule), to the midpoint rule's O((l/rho)^2) and the
  regularisation's O(R^2/rho^2).
* The endpoint form (vortex vector gamma L t) and the query form (gamma d)
  agree to f32 rounding at factor 0; u - u_inf is linear in gamma;
  inactive sources add nothing and inactive targets are not written.
* RK4 with no vorticity translates every endpoint by dt u_inf.
* A vortex ring moves along its axis (right-handed with its circulation),
  every segment with the same axial velocity, and stays round.
"""
import numpy as np
import pytest

import oracle.refsim as R


def _filament(n, L, z0=0.0):
    z = np.linspace(-L, L, n + 1)
    xm = np.stack([np.zeros(n), np.zeros(n), z[:-1] + z0], 1).astype(np.float32)
    xp = np.stack([np.zeros(n), np.zeros(n), z[1:] + z0], 1).astype(np.float32)
    return xm, xp


def test_straight_filament_closed_form():
    n, L, rho, G = 4000, 1.0, 0.25, 1.3
    xm, xp = _filament(n, L)
    # one short target segment at distance rho on the x axis, along y
    tm = np.array([[rho, -1e-4, 0.0]], np.float32)
    tp = np.array([[rho, 1e-4, 0.0]], np.float32)
    xm = np.concatenate([xm, tm])
    xp = np.concatenate([xp, tp])
    gamma = np.concatenate([np.full(n, G, np.float32), [np.float32(0.0)]])
    act = np.ones(n + 1, np.int32)
    c = R.seg_const(reg_radius=1e-4)
    um, up = R.segment_velocity(xm, xp, gamma, act, c, form=0)
    # t = +z: at (px, py, 0) the velocity is |u| (-py, px, 0) / rho', right-handed
    for u, q in ((um[n], tm[0]), (up[n], tp[0])):
        rp = float(np.hypot(q[0], q[1]))
        mag = G / (4 * np.pi * rp) * 2 * L / np.sqrt(L * L + rp * rp)
        want = mag * np.array([-q[1], q[0], 0.0]) / rp
        assert np.linalg.norm(u - want) < 2e-4 * mag, (u, want)
    # the filament's own segments see no velocity along the filament
    assert np.abs(um[:n, 2]).max() < 1e-6


def _random_segments(n, seed=0, inactive=0.1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, (n, 3))
    d = rng.normal(size=(n, 3)) * 0.02
    xm = (c - 0.5 * d).astype(np.float32)
    xp = (c + 0.5 * d).astype(np.float32)
    gamma = rng.normal(0.3, 0.2, n).astype(np.float32)
    act = (rng.random(n) >= inactive).astype(np.int32)
    return xm, xp, gamma, act


def test_forms_linearity_and_activity():
    xm, xp, gamma, act = _random_segments(500, seed=1)
    c0 = R.seg_const(reg_radius=0.01, u_inf=(0.0, 0.0, 0.0))
    um0, up0 = R.segment_velocity(xm, xp, gamma, act, c0, form=0)
    um1, up1 = R.segment_velocity(xm, xp, gamma, act, c0, form=1, factor=0.0)
    on = act == 1
    np.testing.assert_allclose(um1[on], um0[on], rtol=2e-5, atol=1e-6 * np.abs(um0).max())
    np.testing.assert_allclose(up1[on], up0[on], rtol=2e-5, atol=1e-6 * np.abs(up0).max())
    cu = R.seg_const(reg_radius=0.01, u_inf=(0.5, -1.0, 2.0))
    um2, _ = R.segment_velocity(xm, xp, 2 * gamma, act, cu, form=0)
    np.testing.assert_allclose(um2[on] - np.float32([0.5, -1.0, 2.0]), 2 * um0[on], rtol=1e-4,
                               atol=1e-5 * np.abs(um0).max())
    # inactive targets are left as they were; inactive sources add nothing
    sent = np.full((500, 3), 7.0, np.float32)
    om, _ = R.segment_velocity(xm, xp, gamma, act, c0, form=0, out_m=sent.copy(), out_p=sent.copy())
    assert np.all(om[~on] == 7.0)
    g2 = gamma.copy()
    g2[~on] = 1e6
    um3, _ = R.segment_velocity(xm, xp, g2, act, c0, form=0)
    assert np.array_equal(um3[on], um0[on])


def test_rk4_translation_and_ring_motion():
    xm, xp, gamma, act = _random_segments(200, seed=2, inactive=0.0)
    u = (0.3, -0.2, 1.1)
    c = R.seg_const(reg_radius=0.01, dt=0.01, u_inf=u)
    g0 = np.zeros_like(gamma)
    vm, vp = R.segment_velocity(xm, xp, g0, act, c, form=0)
    xm1, xp1, age = xm.copy(), xp.copy(), np.zeros(200, np.float32)
    R.segment_rk4(xm1, xp1, g0, act, age, c, vm, vp)
    np.testing.assert_allclose(xm1 - xm, np.tile(np.float32(u) * 0.01, (200, 1)), atol=2e-7)
    np.testing.assert_allclose(age, 0.01, rtol=1e-6)
    # a ring of radius 0.15 about +x, circulation 0.25 (right-handed): moves along +x
    n = 128
    ang = np.linspace(0.0, 2 * np.pi, n + 1, dtype=np.float32)[:-1]
    pts = np.stack([np.full(n, 0.5), 0.5 + 0.15 * np.cos(ang), 0.5 + 0.15 * np.sin(ang)], 1).astype(np.float32)
    rm, rp = pts, np.roll(pts, -1, axis=0)
    g = np.full(n, 0.25, np.float32)
    a = np.ones(n, np.int32)
    cr = R.seg_const(reg_radius=0.01, dt=0.002)
    vm, vp = R.segment_velocity(rm, rp, g, a, cr, form=0)
    assert np.all(vm[:, 0] > 0) and np.ptp(vm[:, 0]) < 1e-3 * vm[:, 0].mean()
    rm1, rp1, age = rm.copy(), rp.copy(), np.zeros(n, np.float32)
    R.segment_rk4(rm1, rp1, g, a, age, cr, vm, vp)
    rad1 = np.linalg.norm(rm1[:, 1:] - 0.5, axis=1)
    assert np.ptp(rad1) < 1e-5 * rad1.mean()     # the ring stays round
    assert np.all(rm1[:, 0] > rm[:, 0])


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1000, 4097])
def test_gpu_segment_velocity_and_rk4_bitexact(n):
    import torch
    from segment_solver import SegmentConfig, SegmentSolver, SegmentSystem
    xm, xp, gamma, act = _random_segments(n, seed=n)
    cfg = SegmentConfig({"SegmentConfiguration": {"segmentMaxNum": n + 7, "timeStepSize": 0.002,
                                                  "regularizationRadiusR": 0.01, "gammaDecay": 0.999,
                                                  "backgroundVelocity": [0.2, 0.0, -0.1]}})
    ss = SegmentSystem(cfg)
    sv = SegmentSolver(ss)
    sv.add_segments(xm, xp, gamma)
    ss.active.t[:n] = torch.as_tensor(act, device=ss.device)
    c = R.seg_const(reg_radius=0.01, dt=0.002, gamma_decay=0.999, u_inf=(0.2, 0.0, -0.1))
    for step in range(2):
        sv.compute_endpoint_velocity()
        vm, vp = sv.v_minus.to_numpy()[:n], sv.v_plus.to_numpy()[:n]
        rm, rp = R.segment_velocity(xm, xp, gamma, act, c, form=0, out_m=vm.copy() * 0, out_p=vp.copy() * 0)
        on = act == 1
        assert np.array_equal(vm[on], rm[on]) and np.array_equal(vp[on], rp[on]), step
        sv.advect_segments_rk4()
        age = np.zeros(n, np.float32) if step == 0 else age
        R.segment_rk4(xm, xp, gamma, act, age, c, rm, rp)
        for name, ref in (("x_minus", xm), ("x_plus", xp), ("gamma", gamma), ("age", age)):
            assert np.array_equal(getattr(ss, name).to_numpy()[:n], ref), (step, name)
