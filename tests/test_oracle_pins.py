"""Pins of the oracle's neighbour passes, per-particle maps and comparators.

Each test fixes one oracle function (oracle/refsim.cpp) by something the
mathematics or the physics determines independently of the code: a closed
form on a perfect lattice (shell sums of the cubic spline and its exact
derivative, evaluated in FP64 from the kernel's definition), an exact
conservation law, a special case that reduces to another pinned function,
or a two- or three-particle configuration whose answer follows from the
geometry.  Each is chosen so that a dropped term, a flipped sign, a wrong
index (i for j) or a transposed operand fails at least one of them.

Lattice conventions: spacing d = 2r, support h = 4r = 2d, m_V = 0.8 d^3,
m = m_V rho (rho = 1000 at insertion).  The moment c of the lattice,

    sum_j m_V r_j (x) gradW(r_j) = -c I,   c = -(m_V/3) sum_j |r_j| W'(|r_j|),

(r_j = x_i - x_j, sum over the 26 stencil points inside the support) is the
factor every linear-field closed form below carries; c = 0.81604 here.
"""
import math

import numpy as np
import pytest

import oracle.refsim as R

L = R.lib()

RAD = 0.01
D = 2 * RAD
H = 4 * RAD
KW = 2.5465
MV = 0.8 * D ** 3
RHO0 = 1000.0
DT = 1e-3


# ------------------------------------------------------------ FP64 maths --
def W64(s):
    """Cubic spline W(s) (sph_base.py:23-44), FP64."""
    q = np.asarray(s, np.float64) / H
    k = KW / H ** 3
    return np.where(q <= 0.5, k * (6 * q ** 3 - 6 * q ** 2 + 1),
                    np.where(q <= 1.0, 2 * k * (1 - q) ** 3, 0.0))


def dW64(s):
    """dW/ds, the exact derivative of W64 (gradW = W'(|r|) r/|r|)."""
    q = np.asarray(s, np.float64) / H
    k = KW / H ** 3
    return np.where(q <= 0.5, k / H * (18 * q ** 2 - 12 * q),
                    np.where(q <= 1.0, -6 * k / H * (1 - q) ** 2, 0.0))


def stencil():
    """Offsets r = x_i - x_j of a perfect lattice inside the support."""
    o = np.array([(a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)
                  if (a, b, c) != (0, 0, 0)], np.float64) * D
    s = np.linalg.norm(o, axis=1)
    keep = s < H - 1e-12
    return o[keep], s[keep]


OFFS, DIST = stencil()
C_MOM = -(MV / 3.0) * float(np.sum(DIST * dW64(DIST)))


def test_moment_constant():
    # sanity of the FP64 helpers: W integrates to 2.5465/(8/pi) and W' is its derivative
    s = np.linspace(0, H, 200001)
    assert abs(4 * np.pi * np.trapezoid(s ** 2 * W64(s), s) - KW / (8 / np.pi)) < 1e-6
    e = 1e-7
    mid = np.linspace(0.001, 0.039, 77)
    np.testing.assert_allclose((W64(mid + e) - W64(mid - e)) / (2 * e), dW64(mid), rtol=1e-5, atol=1e-3)
    assert len(OFFS) == 26
    assert C_MOM == pytest.approx(0.81604, abs=1e-5)


# ------------------------------------------------------------- scenes -----
def _cfg(**kw):
    c = {"domainStart": [0, 0, 0], "domainEnd": [1.0, 1.0, 1.0], "particleRadius": RAD,
         "density0": RHO0, "timeStepSize": DT, "gravitation": [0, 0, 0], "volumeScale": 0.2,
         "GenProbPara": 5, "DelProbPara": 3.5}
    c.update(kw)
    return c


def make(pos, vel=None, density=None, material=None, dynamic=None, extra=0, **cfg):
    """RefSim holding `pos` (f32), initialised (activity, partition, sort,
    boundary volumes).  Returns (sim, slot_of) with slot_of[k] the sorted slot
    of input particle k."""
    pos = np.ascontiguousarray(pos, np.float32)
    m = pos.shape[0]
    c = _cfg(**cfg)
    c["reservedCapacity"] = m + extra
    s = R.RefSim({"Configuration": c})
    vel = np.zeros_like(pos) if vel is None else np.ascontiguousarray(vel, np.float32)
    dens = np.full(m, RHO0, np.float32) if density is None else np.asarray(density, np.float32)
    mat = np.ones(m, np.int32) if material is None else np.asarray(material, np.int32)
    dyn = np.ones(m, np.int32) if dynamic is None else np.asarray(dynamic, np.int32)
    s.add_particles(0, m, pos, vel, dens, np.zeros(m, np.float32), mat, dyn, np.zeros((m, 3), np.int32))
    s.i32("solid_particle_num")[0] = int((mat == 0).sum())
    # tag each insertion index in pressure (moved by the sort, read by no pass here)
    s.field("pressure")[:m] = np.arange(m, dtype=np.float32)
    s.initialize()
    n = s.particle_num
    slot_of = np.full(m, -1, np.int64)
    slot_of[s.field("pressure")[:n].astype(np.int64)] = np.arange(n)
    return s, slot_of


def lattice(n, offset=3.5):
    g = (np.arange(n) + offset) * D
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)


def interior(x, n=10, offset=3.5, layers=2):
    lo = (offset + layers) * D - 1e-6
    hi = (offset + n - 1 - layers) * D + 1e-6
    return np.all((x >= lo) & (x <= hi), axis=1)


def lattice_sim(n=10, vel=None, **kw):
    pos = lattice(n)
    v = None if vel is None else vel(pos)
    s, _ = make(pos, v, **kw)
    x = s.field("x")[: s.particle_num].astype(np.float64)
    return s, x, interior(x, n)


def periodic_box(n=12, vel=None, density=None, seed=0, jitter=0.1 * RAD):
    """Fluid-only periodic box of n^3 lattice sites, jittered, with the
    particles of cell 0 removed: no list holds cell 0 (particle_system.py
    :576-578), so without it every pair is seen from both sides."""
    rng = np.random.default_rng(seed)
    box = n * D
    pos = lattice(n, offset=0.5) + rng.uniform(-jitter, jitter, (n ** 3, 3))
    pos = pos[~np.all(pos < H, axis=1)].astype(np.float32)
    m = pos.shape[0]
    v = None if vel is None else vel(pos, rng)
    dens = None if density is None else density(m, rng)
    s, _ = make(pos, v, dens, domainEnd=[box, box, box], periodic=True)
    assert s.particle_num == m
    return s, rng


def f32v(s, name):
    return s.field(name)[: s.particle_num].astype(np.float64)


# ------------------------------------------------- A24, A23, A25 ----------
def test_grad_v_linear_field_closed_form():
    # v = A x: G_i = sum m_V (v_j - v_i) (x) gradW = -A sum m_V r (x) gradW = c A
    rng = np.random.default_rng(2)
    A = rng.normal(size=(3, 3))
    s, x, inner = lattice_sim(vel=lambda p: p @ A.T)
    L.or_compute_grad_v(s.sim)
    G = f32v(s, "grad_v").reshape(-1, 3, 3)[inner]
    np.testing.assert_allclose(G, np.broadcast_to(C_MOM * A, G.shape), rtol=0, atol=2e-5 * np.abs(A).max())


def test_vorticity_rigid_rotation_closed_form():
    # v = W x x: "vorticity" = sum m_V (v_j - v_i) x gradW = -2 c W (= c times -curl v)
    Om = np.array([0.7, -1.3, 2.1])
    s, x, inner = lattice_sim(vel=lambda p: np.cross(Om, p))
    L.or_compute_vorticity(s.sim)
    om = f32v(s, "vorticity")[inner]
    np.testing.assert_allclose(om, np.broadcast_to(-2 * C_MOM * Om, om.shape), rtol=0, atol=2e-5)
    # v = A x with A symmetric (pure strain): no vorticity
    A = np.array([[1.0, 0.5, -0.2], [0.5, -0.3, 0.8], [-0.2, 0.8, 0.4]])
    s, x, inner = lattice_sim(vel=lambda p: p @ A.T)
    L.or_compute_vorticity(s.sim)
    assert np.abs(f32v(s, "vorticity")[inner]).max() < 2e-5


def test_d_vorticity_pins():
    # Delta omega = omega(v) - curl(v*), v* = v + a dt (compute_v_star)
    Om = np.array([0.4, 1.1, -0.6])       # velocity: rigid rotation
    Oa = np.array([-2.0, 0.5, 3.0])       # acceleration: another rigid rotation
    s, x, inner = lattice_sim(vel=lambda p: np.cross(Om, p))
    L.or_compute_vorticity(s.sim)
    # (1) a = 0: v* = v, so Delta omega vanishes exactly
    s.field("acceleration")[:] = 0.0
    L.or_compute_v_star(s.sim)
    L.or_compute_d_vorticity(s.sim)
    assert np.all(s.field("d_vort")[: s.particle_num] == 0.0)
    # (2) uniform acceleration: v*_j - v*_i = v_j - v_i up to rounding
    s.field("acceleration")[: s.particle_num] = np.array([3.0, -7.0, 5.0], np.float32)
    L.or_compute_v_star(s.sim)
    L.or_compute_d_vorticity(s.sim)
    assert np.abs(f32v(s, "d_vort")).max() < 1e-4
    # (3) a = Oa x x: Delta omega = -dt omega(a) = 2 c dt Oa
    s.field("acceleration")[: s.particle_num] = np.cross(Oa, x).astype(np.float32)
    L.or_compute_v_star(s.sim)
    L.or_compute_d_vorticity(s.sim)
    dv = f32v(s, "d_vort")[inner]
    np.testing.assert_allclose(dv, np.broadcast_to(2 * C_MOM * DT * Oa, dv.shape), rtol=0, atol=3e-5)


# ---------------------------------------------------------- A14, A15 -------
def test_dfsph_factor_lattice_closed_form():
    # interior: sum m gradW = 0 by symmetry, alpha = 1 / (m^2 sum_j W'(|r_j|)^2)
    s, x, inner = lattice_sim()
    L.or_compute_DFSPH_factor(s.sim)
    m = MV * RHO0
    want = 1.0 / (m * m * float(np.sum(dW64(DIST) ** 2)))
    np.testing.assert_allclose(f32v(s, "dfsph_factor")[inner], want, rtol=2e-5)


def test_dfsph_factor_few_particle_cases():
    m = MV * RHO0
    a = 0.3 * H
    g = float(dW64(a))
    # two fluid particles: both sums see the one term, alpha = 1/(2 m^2 W'^2)
    s, slot = make([[0.3, 0.3, 0.3], [0.3 + a, 0.3, 0.3]])
    L.or_compute_DFSPH_factor(s.sim)
    np.testing.assert_allclose(s.field("dfsph_factor")[slot], 1.0 / (2 * m * m * g * g), rtol=2e-5)
    # fluid between a fluid neighbour (+x) and a solid (-x) at equal distance:
    # the solid enters only the squared vector sum, with weight m_V,s rho0
    pos = [[0.3, 0.3, 0.3], [0.3 + a, 0.3, 0.3], [0.3 - a, 0.3, 0.3]]
    s, slot = make(pos, material=[1, 1, 0], dynamic=[1, 1, 0])
    mvs = float(s.field("m_V")[slot[2]])
    np.testing.assert_allclose(mvs, 3.0 / float(W64(0.0)), rtol=1e-6)   # lone solid
    L.or_compute_DFSPH_factor(s.sim)
    want = 1.0 / (m * m * g * g + (m - mvs * RHO0) ** 2 * g * g)
    np.testing.assert_allclose(s.field("dfsph_factor")[slot[0]], want, rtol=2e-5)
    # an isolated particle: the sum is below 1e-6, alpha = 0
    s, slot = make([[0.5, 0.5, 0.5]])
    L.or_compute_DFSPH_factor(s.sim)
    assert s.field("dfsph_factor")[slot[0]] == 0.0


# -------------------------------------------------------- A16, A33 --------
def test_density_change_linear_field_closed_form():
    # v = A x: rho_dot = sum m_V (v_i - v_j).gradW = sum_ab A_ab sum m_V r_b gradW_a = -c tr A
    A = np.array([[-30.0, 4.0, 1.0], [2.0, -20.0, -3.0], [5.0, 1.0, -10.0]])
    s, x, inner = lattice_sim(vel=lambda p: p @ A.T)
    L.or_compute_density_change(s.sim)
    np.testing.assert_allclose(f32v(s, "density_adv")[inner], -C_MOM * np.trace(A), rtol=1e-4)
    # expansion: rho_dot clamps to 0
    s, x, inner = lattice_sim(vel=lambda p: -(p @ A.T))
    L.or_compute_density_change(s.sim)
    assert np.all(f32v(s, "density_adv") == 0.0)


def test_density_change_needs_twenty_neighbours():
    # a particle with n neighbours on a sphere, all approaching it: rho_dot > 0
    # for n = 20 and exactly 0 for n = 19 (dimcv_sph.py:351-356); solids count
    rng = np.random.default_rng(3)
    dirs = rng.normal(size=(20, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    c0 = np.array([0.5, 0.5, 0.5])
    for nn in (19, 20):
        pos = np.concatenate([c0[None], c0 + 0.6 * H * dirs[:nn]])
        vel = np.concatenate([np.zeros((1, 3)), -2.0 * dirs[:nn]])
        mat = np.ones(nn + 1, np.int32)
        mat[1:6] = 0                                    # five of them solid
        s, slot = make(pos, vel, material=mat, dynamic=(mat == 1).astype(np.int32))
        assert len(s.neighbors(int(slot[0]))) == nn
        L.or_compute_density_change(s.sim)
        got = float(s.field("density_adv")[slot[0]])
        if nn == 19:
            assert got == 0.0
        else:
            # v_i - v_j = 2 dir_j, gradW(x_i - x_j) = W'(0.6h) (-dir_j): each term 2 m_V,j |W'|
            mv = s.field("m_V")[slot[1:]].astype(np.float64)
            want = float(np.sum(mv * 2.0 * -dW64(0.6 * H)))
            assert got == pytest.approx(want, rel=1e-5)


def test_density_adv_closed_forms():
    # rho* = max(rho + dt sum m_j (v_i - v_j).gradW, rho0) = max(rho - dt rho0 c tr A, rho0)
    A = -200.0 * np.eye(3) + np.array([[0, 3.0, 0], [0, 0, -2.0], [1.0, 0, 0]])
    s, x, inner = lattice_sim(vel=lambda p: p @ A.T)
    L.or_compute_densities(s.sim)
    L.or_compute_density_adv(s.sim)
    rho = f32v(s, "density")
    want = rho - DT * RHO0 * C_MOM * np.trace(A)
    assert np.all(want[inner] > RHO0)
    np.testing.assert_allclose(f32v(s, "density_adv")[inner], want[inner], rtol=2e-6)
    # weak compression: rho + dt(...) < rho0 clamps to rho0
    s, x, inner = lattice_sim(vel=lambda p: 0.01 * (p @ A.T))
    L.or_compute_densities(s.sim)
    L.or_compute_density_adv(s.sim)
    assert np.all(f32v(s, "density_adv") == RHO0)
    # a fluid particle moving into a solid: the solid enters with m_V,s rho0
    a = 0.4 * H
    pos = [[0.5, 0.5, 0.5], [0.5 - a, 0.5, 0.5]]
    s, slot = make(pos, [[-50.0, 0, 0], [0, 0, 0]], material=[1, 0], dynamic=[1, 0])
    s.field("density")[slot[0]] = 900.0
    L.or_compute_density_adv(s.sim)
    mvs = float(s.field("m_V")[slot[1]])
    want = 900.0 + DT * mvs * RHO0 * 50.0 * -float(dW64(a))
    assert want > RHO0
    assert float(s.field("density_adv")[slot[0]]) == pytest.approx(want, rel=2e-6)


def test_densities_see_solids_with_their_volume():
    # rho_i = rho0 (m_V,i W(0) + sum_j m_V,j W_ij) over fluid and solid j
    a = 0.45 * H
    pos = [[0.5, 0.5, 0.5], [0.5 + a, 0.5, 0.5], [0.5, 0.5 - a, 0.5]]
    s, slot = make(pos, material=[1, 1, 0], dynamic=[1, 1, 0])
    L.or_compute_densities(s.sim)
    mvs = float(s.field("m_V")[slot[2]])
    want = RHO0 * (MV * W64(0.0) + MV * W64(a) + mvs * W64(a))
    assert float(s.field("density")[slot[0]]) == pytest.approx(float(want), rel=2e-6)


# ------------------------------------------------- A19, A34, A20, A35 -----
def _set_kappa(s, kappa, which, alpha=None, rho=None):
    """Fields whose product is kappa: alpha (dfsph_factor, default 1) and, in
    the divergence update, rho (default 1)."""
    n = s.particle_num
    alpha = np.ones(n) if alpha is None else np.asarray(alpha, np.float64)
    rho = np.ones(n) if rho is None else np.asarray(rho, np.float64)
    s.field("dfsph_factor")[:n] = alpha.astype(np.float32)
    if which == "div":          # kappa = rho_dot alpha rho
        s.field("density")[:n] = rho.astype(np.float32)
        s.field("density_adv")[:n] = (kappa / (alpha * rho)).astype(np.float32)
    else:                       # kappa = (rho* - rho0) alpha
        s.field("density_adv")[:n] = (RHO0 + kappa / alpha).astype(np.float32)
    s.field("v")[:n] = 0.0


@pytest.mark.parametrize("which", ["div", "pressure"])
def test_jacobi_update_linear_kappa_closed_form(which):
    # kappa = k0 + g.x: sum_j (kappa_i + kappa_j) m gradW = (m c / m_V) g, so
    # Delta v = -dt rho0 c g: the velocity change points down the kappa gradient
    g = np.array([500.0, -300.0, 200.0])
    s, x, inner = lattice_sim()
    _set_kappa(s, 1000.0 + x @ g, which)
    if which == "div":
        L.or_divergence_solver_iteration_kernel(s.sim)
    else:
        L.or_pressure_solve_iteration_kernel(s.sim)
    dv = f32v(s, "v")[inner]
    np.testing.assert_allclose(dv, np.broadcast_to(-DT * RHO0 * C_MOM * g, dv.shape), rtol=0,
                               atol=2e-4 * np.abs(g).max())


@pytest.mark.parametrize("which", ["div", "pressure"])
def test_jacobi_update_pairs_push_apart(which):
    a = 0.35 * H
    m = MV * RHO0
    gabs = -float(dW64(a))
    kap = 4.0
    # two fluid particles, both compressed: each is pushed away from the other
    # with weight kappa_i + kappa_j; kappa_j = 2.5 kappa_i comes from alpha_j = 2,
    # rho_j = 1.25 (divergence) or alpha_j = 2 (pressure), so each factor is read at j
    s, slot = make([[0.5, 0.5, 0.5], [0.5 + a, 0.5, 0.5]])
    kv, al, rh = np.empty(2), np.empty(2), np.empty(2)
    kv[slot], al[slot], rh[slot] = [kap, 2.5 * kap], [1.0, 2.0], [1.0, 1.25]
    _set_kappa(s, kv, which, al, rh)
    (L.or_divergence_solver_iteration_kernel if which == "div" else L.or_pressure_solve_iteration_kernel)(s.sim)
    v = s.field("v")[slot].astype(np.float64)
    np.testing.assert_allclose(v[0], [-DT * 3.5 * kap * m * gabs, 0, 0], rtol=2e-5, atol=1e-12)
    np.testing.assert_allclose(v[1], -v[0], rtol=2e-6)
    # fluid next to a solid: pushed away with weight kappa_i m_V,s rho0 (solid kappa unused)
    s, slot = make([[0.5, 0.5, 0.5], [0.5 - a, 0.5, 0.5]], material=[1, 0], dynamic=[1, 0])
    _set_kappa(s, np.array([kap, 0.0]) if slot[0] == 0 else np.array([0.0, kap]), which)
    (L.or_divergence_solver_iteration_kernel if which == "div" else L.or_pressure_solve_iteration_kernel)(s.sim)
    mvs = float(s.field("m_V")[slot[1]])
    np.testing.assert_allclose(s.field("v")[slot[0]], [DT * kap * mvs * RHO0 * gabs, 0, 0], rtol=2e-5, atol=1e-12)
    assert np.all(s.field("v")[slot[1]] == 0.0)
    # |kappa_i| <= 1e-5: no solid term
    _set_kappa(s, np.full(2, 5e-6), which)
    (L.or_divergence_solver_iteration_kernel if which == "div" else L.or_pressure_solve_iteration_kernel)(s.sim)
    assert np.all(s.field("v")[slot[0]] == 0.0)


@pytest.mark.parametrize("which", ["div", "pressure"])
def test_jacobi_update_transit_weight(which):
    # v_i += w(x_i) Delta v, w = clamp((x - a)/(b - a), 0, 1)  (dimcv_sph.py:401-405)
    a = 0.35 * H
    kap = 4.0
    full = None
    for xc, w in ((0.45, None), (0.55, 0.5), (0.65, 1.0), (0.75, 1.0)):
        pos = [[xc, 0.5, 0.5], [xc, 0.5 + a, 0.5]]
        s, slot = make(pos, divergenceTransitArea=[0.45, 0.65])
        _set_kappa(s, np.full(2, kap), which)
        (L.or_divergence_solver_iteration_kernel if which == "div" else L.or_pressure_solve_iteration_kernel)(s.sim)
        v = s.field("v")[slot[0]].astype(np.float64)
        if w is None:
            assert np.all(v == 0.0)
        elif w == 1.0:
            full = v if full is None else full
            np.testing.assert_allclose(v, full, rtol=1e-6)
        else:
            half = v
    np.testing.assert_allclose(half, 0.5 * full, rtol=1e-5)


@pytest.mark.parametrize("which", ["div", "pressure"])
def test_jacobi_update_conserves_momentum(which):
    # Delta v_i = -dt sum_j [|k_i + k_j| > eps](k_i + k_j) m_j gradW_ij: the pair
    # terms m_i m_j (k_i + k_j) gradW_ij are antisymmetric, so sum_i m_i Delta v_i
    # = 0 for any masses; unequal masses catch an i/j index mix-up
    s, rng = periodic_box(density=lambda m, rng: rng.uniform(500, 1500, m).astype(np.float32))
    n = s.particle_num
    kappa = rng.normal(0.0, 3.0, n)
    kappa[::7] = 2e-6                 # some pairs fall under the 1e-5 gate
    _set_kappa(s, kappa, which, rng.uniform(0.5, 2.0, n), rng.uniform(0.5, 2.0, n))
    (L.or_divergence_solver_iteration_kernel if which == "div" else L.or_pressure_solve_iteration_kernel)(s.sim)
    m = f32v(s, "m")
    dv = f32v(s, "v")
    p = (m[:, None] * dv).sum(0)
    scale = (m[:, None] * np.abs(dv)).sum(0)
    assert np.all(np.abs(p) < 1e-5 * scale), (p, scale)


def test_jacobi_loops_stop_at_first_converged_iteration():
    # a compressed lattice: the pressure solve iterates; the average residual
    # after the last iteration is <= eta and after the one before it is not
    # (dimcv_sph.py:597-614; eta = 0.05 * 0.01 * rho0)
    pos = lattice(10) * 0.9
    for solve in ("pressure", "divergence"):
        vel = None if solve == "pressure" else -8000.0 * (pos - pos.mean(0))
        s, _ = make(pos, vel)
        L.or_compute_densities(s.sim)
        L.or_compute_DFSPH_factor(s.sim)
        if solve == "pressure":
            s.pressure_solve()
            k = s.last_kp
            eta, offset = 0.05 * 0.01 * RHO0, RHO0
            it_fn, dens_fn = L.or_pressure_solve_iteration_kernel, L.or_compute_density_adv
        else:
            s.divergence_solve()
            k = s.last_kv
            eta, offset = (1.0 / DT) * 0.1 * 0.01 * RHO0, 0.0
            it_fn, dens_fn = L.or_divergence_solver_iteration_kernel, L.or_compute_density_change
        assert 1 < k < 100, (solve, k)
        # replay the same solve pass by pass, recording every residual
        s2, _ = make(pos, vel)
        L.or_compute_densities(s2.sim)
        L.or_compute_DFSPH_factor(s2.sim)
        dens_fn(s2.sim)
        L.or_multiply_time_step(s2.sim, 0, np.float32(1.0 / DT ** 2 if solve == "pressure" else 1.0 / DT))
        res = []
        for _ in range(k):
            it_fn(s2.sim)
            dens_fn(s2.sim)
            t = s2.field("density_adv")[: s2.particle_num].astype(np.float64) - offset
            res.append(math.fsum(t) / s2.fluid_particle_num)
        assert res[-1] <= eta < min(res[:-1]), (solve, res, eta)


def test_density_error_is_the_weighted_sum():
    # A18: sum over fluid of w(x)(rho_adv - offset), w = clamp((x-a)/(b-a), 0, 1)
    rng = np.random.default_rng(11)
    pos = rng.uniform(0.05, 0.95, (3000, 3))
    mat = (rng.random(3000) < 0.8).astype(np.int32)
    for transit in (None, [0.3, 0.6]):
        kw = {} if transit is None else {"divergenceTransitArea": transit}
        s, slot = make(pos, material=mat, dynamic=mat, **kw)
        da = rng.normal(1000.0, 30.0, s.particle_num).astype(np.float32)
        s.field("density_adv")[: s.particle_num] = da
        got = L.or_compute_density_error(s.sim, np.float32(RHO0))
        x = f32v(s, "x")[:, 0]
        fl = s.field("material")[: s.particle_num] == 1
        w = np.ones_like(x) if transit is None else np.clip((x - 0.3) / 0.3, 0.0, 1.0)
        want = math.fsum((w * (da.astype(np.float64) - RHO0))[fl])
        assert got == pytest.approx(want, rel=1e-5, abs=1e-2)


# ------------------------------------------------------------- A21 --------
def _visc_sim(**kw):
    s, rng = periodic_box(**kw)
    L.or_compute_non_pressure_forces(s.sim)
    return s, rng


def test_viscosity_conserves_momentum_and_dissipates():
    # pair terms m_V,i m_V,j (v_ij.r_ij) gradW_ij / (|r|^2 + 0.01 h^2): antisymmetric,
    # so sum_i m_V,i a_i = 0; and sum_i m_V,i v_i.a_i = 1/2 sum_ij ... (v_ij.r)^2 W'/|r| <= 0
    s, rng = _visc_sim(vel=lambda p, rng: rng.normal(size=p.shape).astype(np.float32),
                       density=lambda m, rng: rng.uniform(500, 1500, m).astype(np.float32))
    a = f32v(s, "acceleration")
    v = f32v(s, "v")
    mv = f32v(s, "m_V")
    p = (mv[:, None] * a).sum(0)
    assert np.all(np.abs(p) < 1e-5 * (mv[:, None] * np.abs(a)).sum(0))
    assert (mv * np.einsum("ij,ij->i", v, a)).sum() < 0


def test_viscosity_rigid_motion_and_shear_wave_closed_form():
    box = 12 * D
    # rigid translation and rotation: v_ij.r_ij = 0 for every pair, a = g
    for vel in (lambda p, rng: np.tile([1.0, -2.0, 0.5], (p.shape[0], 1)),
                lambda p, rng: np.cross([0.3, 0.2, -1.0], p - box / 2)):
        s, _ = periodic_box(vel=vel, jitter=0.0)
        L.or_compute_non_pressure_forces(s.sim)
        x = f32v(s, "x")
        inside = np.all((x > H) & (x < box - H), axis=1)   # rotation is not periodic
        assert np.abs(f32v(s, "acceleration")[inside]).max() < 1e-4
    # shear wave v = U sin(k y) e_x on the perfect periodic lattice: for every
    # particle away from cell 0, a_x = -U sin(k y_i) 2(d+2) nu sum_j m_V
    # (1 - cos(k r_y)) r_x^2 |W'| / (|r| (|r|^2 + 0.01 h^2)), a_y = a_z = 0
    # (the sin(k r_y) part cancels by r_x -> -r_x).  Its Taylor limit is
    # nu d2v/dy2 times 5 sum m_V r_x^2 r_y^2 |W'| / (|r|(|r|^2+0.01h^2)) (-> 0.8 for a
    # fine kernel): the viscous force of the Laplacian, with its sign.
    U, nu = 1.3, 1e-3
    k = 2 * np.pi / box
    s, _ = periodic_box(vel=lambda p, rng: np.stack([U * np.sin(k * p[:, 1]), 0 * p[:, 0], 0 * p[:, 0]], 1),
                        jitter=0.0)
    L.or_compute_non_pressure_forces(s.sim)
    x, a = f32v(s, "x"), f32v(s, "acceleration")
    cell = (x / H).astype(int)
    far0 = ~np.all((cell <= 1) | (cell >= 5), axis=1)
    rx, ry, sn = OFFS[:, 0], OFFS[:, 1], DIST
    coef = 2 * 5 * nu * float(np.sum(MV * (1 - np.cos(k * ry)) * rx ** 2 * -dW64(sn) / (sn * (sn ** 2 + 0.01 * H * H))))
    want = -U * np.sin(k * x[:, 1]) * coef
    np.testing.assert_allclose(a[far0, 0], want[far0], rtol=0, atol=2e-4 * np.abs(want).max())
    assert np.abs(a[far0, 1:]).max() < 2e-4 * np.abs(want).max()
    taylor = 5 * float(np.sum(MV * rx ** 2 * ry ** 2 * -dW64(sn) / (sn * (sn ** 2 + 0.01 * H * H))))
    assert coef / (nu * k * k) == pytest.approx(taylor, rel=0.05)   # O((k d)^2) from the limit


def test_gravity_and_static_solids():
    pos = [[0.5, 0.5, 0.5], [0.5 + 0.3 * H, 0.5, 0.5]]
    s, slot = make(pos, material=[1, 0], dynamic=[1, 0], gravitation=[0.0, -9.81, 0.0])
    s.field("acceleration")[:] = 7.0
    L.or_compute_non_pressure_forces(s.sim)
    np.testing.assert_array_equal(s.field("acceleration")[slot[0]], np.float32([0, -9.81, 0]))
    np.testing.assert_array_equal(s.field("acceleration")[slot[1]], [0, 0, 0])


# --------------------------------------------------------- A26, A36 --------
def test_smoothing_closed_forms():
    # smoothed_i = m_V,i dv_i W(0) + sum_fluid m_V,j dv_j W_ij: a constant dv gives
    # dv rho_i / rho0 with rho from compute_densities (pinned: 799.98 in the
    # interior); a linear dv keeps its slope (first moment vanishes)
    s, x, inner = lattice_sim()
    L.or_compute_densities(s.sim)
    rho = f32v(s, "density")
    D0 = np.array([1.5, -2.0, 0.25])
    B = np.array([[3.0, 0.0, 1.0], [0.0, -2.0, 0.5], [4.0, 1.0, 0.0]])
    n = s.particle_num
    s.field("d_vort")[:n] = D0.astype(np.float32)
    L.or_compute_all_d_vort_smoothed(s.sim)
    np.testing.assert_allclose(f32v(s, "d_vort_smoothed")[inner], (rho[:, None] / RHO0 * D0)[inner], rtol=2e-6)
    s.field("d_vort")[:n] = (D0 + x @ B.T).astype(np.float32)
    L.or_compute_all_d_vort_smoothed(s.sim)
    np.testing.assert_allclose(f32v(s, "d_vort_smoothed")[inner], (rho[:, None] / RHO0 * (D0 + x @ B.T))[inner],
                               rtol=1e-5)
    # two particles with different dv: each adds the other's dv (not its own)
    # weighted by W at their distance
    rr = 0.7 * H
    s, slot = make([[0.5, 0.5, 0.5], [0.5 + rr, 0.5, 0.5]])
    d0, d1 = np.array([1.0, -2.0, 0.5]), np.array([-3.0, 0.25, 4.0])
    s.field("d_vort")[slot[0]] = d0
    s.field("d_vort")[slot[1]] = d1
    L.or_compute_all_d_vort_smoothed(s.sim)
    r01 = float(np.float32(0.5 + rr) - np.float32(0.5))
    np.testing.assert_allclose(s.field("d_vort_smoothed")[slot[0]], MV * (W64(0.0) * d0 + W64(r01) * d1), rtol=2e-6)
    np.testing.assert_allclose(s.field("d_vort_smoothed")[slot[1]], MV * (W64(0.0) * d1 + W64(r01) * d0), rtol=2e-6)
    # an isolated particle keeps only its own term
    s, slot = make([[0.5, 0.5, 0.5]])
    s.field("d_vort")[slot[0]] = [2.0, 0.0, -1.0]
    L.or_compute_all_d_vort_smoothed(s.sim)
    np.testing.assert_allclose(s.field("d_vort_smoothed")[slot[0]], MV * float(W64(0.0)) * np.array([2.0, 0, -1.0]),
                               rtol=1e-6)


def test_smoothing_skips_particles_outside_the_vortex_domain():
    pos = lattice(6)
    s, _ = make(pos, vortexEnforcingDomainStart=[0.0, 0.0, 0.0], vortexEnforcingDomainEnd=[0.15, 1.0, 1.0])
    n = s.particle_num
    s.field("d_vort")[:n] = 1.0
    s.field("d_vort_smoothed")[:n] = -7.0
    L.or_compute_all_d_vort_smoothed(s.sim)
    x = f32v(s, "x")
    out = x[:, 0] > np.float32(0.15)
    assert out.any() and (~out).any()
    assert np.all(s.field("d_vort_smoothed")[:n][out] == -7.0)
    assert np.all(s.field("d_vort_smoothed")[:n][~out] > 0.0)


def test_vorticity_vis_closed_forms():
    # Shepard average over all neighbours, self included:
    # vis_i = sum m_V W omega / (sum m_V W + 1e-6)
    s, x, inner = lattice_sim()
    L.or_compute_densities(s.sim)
    S = f32v(s, "density") / RHO0
    om = np.array([2.0, -1.0, 4.0])
    s.field("vorticity")[: s.particle_num] = om.astype(np.float32)
    L.or_compute_vorticity_vis(s.sim)
    np.testing.assert_allclose(f32v(s, "vorticity_vis"), (S / (S + 1e-6))[:, None] * om, rtol=2e-6)
    # solids count with their m_V and their vorticity (a solid with omega = 0 dilutes)
    a = 0.4 * H
    s, slot = make([[0.5, 0.5, 0.5], [0.5 + a, 0.5, 0.5]], material=[1, 0], dynamic=[1, 0])
    s.field("vorticity")[slot[0]] = om
    s.field("vorticity")[slot[1]] = 0.0
    L.or_compute_vorticity_vis(s.sim)
    wi, wj = MV * float(W64(0.0)), float(s.field("m_V")[slot[1]]) * float(W64(a))
    np.testing.assert_allclose(s.field("vorticity_vis")[slot[0]], wi * om / (wi + wj + 1e-6), rtol=5e-6)


# ------------------------------------------------------------- A27 --------
def test_sample_birth_and_death_probabilities():
    # p_prod = 1 - (KVN + 1)^-(0.018 exp(-a clamp(t,0,1)) + 0.004), p_del = exp(-b KVN)
    pos = np.array([[0.1 + 0.1 * i, 0.5, 0.5] for i in range(6)])
    s, slot = make(pos)
    n = s.particle_num
    G = s.field("grad_v").reshape(-1, 3, 3)
    rot = np.array([[0, -1.0, 0], [1.0, 0, 0], [0, 0, 0]])
    shear = np.array([[0, 1.0, 0], [0, 0, 0], [0, 0, 0]])
    G[:n] = 0.0                                      # KVN = 0
    u_lo = np.zeros(n, np.float32)
    u_hi = np.full(n, np.float32(np.nextafter(np.float32(1), np.float32(0))))
    # p_prod(0) = 0: no birth even for u1 = 0
    s.field("is_sample")[:n] = 0
    L.or_compute_kinematic_vorticity_num_sample(s.sim, R._ptr(u_lo), R._ptr(u_hi), 1)
    assert np.all(s.field("is_sample")[:n] == 0)
    # p_del(0) = 1: a sample with KVN = 0 dies for any u2 < 1
    s.field("is_sample")[:n] = 1
    L.or_compute_kinematic_vorticity_num_sample(s.sim, R._ptr(u_hi), R._ptr(u_hi), 1)
    assert np.all(s.field("is_sample")[:n] == 0)
    # KVN >= 2 never dies (rotation: KVN = |W|/1e-8); KVN = 1 (shear) dies iff u2 < e^-3.5
    G[:n] = rot
    s.field("is_sample")[:n] = 1
    L.or_compute_kinematic_vorticity_num_sample(s.sim, R._ptr(u_hi), R._ptr(u_lo), 1)
    assert np.all(s.field("is_sample")[:n] == 1)
    G[:n] = shear
    pdel = math.exp(-3.5)
    u2 = np.array([pdel * 0.999, pdel * 1.001] * 3, np.float32)
    L.or_compute_kinematic_vorticity_num_sample(s.sim, R._ptr(u_hi), R._ptr(u2), 1)
    np.testing.assert_array_equal(s.field("is_sample")[:n], (u2 >= pdel).astype(np.uint8))
    # birth of a shear particle (KVN = 1) at life times -3, 0, 0.5, 1, 4: the
    # exponent clamps t to [0, 1]
    for t in (-3.0, 0.0, 0.5, 1.0, 4.0):
        r = 0.018 * math.exp(-5.0 * min(max(t, 0.0), 1.0)) + 0.004
        p = 1.0 - 2.0 ** (-r)
        s.field("is_sample")[:n] = 0
        s.field("life_time")[:n] = t
        u1 = np.array([p * 0.995, p * 1.005] * 3, np.float32)
        L.or_compute_kinematic_vorticity_num_sample(s.sim, R._ptr(u1), R._ptr(u_hi), 1)
        np.testing.assert_array_equal(s.field("is_sample")[:n], (u1 < p).astype(np.uint8))


# ------------------------------------------------------------- A29 --------
def test_kernel_density_pair_closed_form():
    # kden_i = KVN_i E(0) + KVN_j prod_d 0.75 (1 - (|r_d|/h)^2), E(0) = 0.421875
    r = np.array([0.2, -0.35, 0.1]) * H
    s, slot = make([[0.5, 0.5, 0.5], [0.5 + r[0], 0.5 + r[1], 0.5 + r[2]]])
    s.field("kvn")[slot] = [3.0, 5.0]
    L.or_compute_kernel_density(s.sim)
    E = float(np.prod(0.75 * (1 - (r / H) ** 2)))
    np.testing.assert_allclose(s.field("kernel_density")[slot], [3.0 * 0.421875 + 5.0 * E, 5.0 * 0.421875 + 3.0 * E],
                               rtol=2e-6)
    assert s.kden_sum == pytest.approx(8.0 * (0.421875 + E), rel=2e-6)


# ------------------------------------------------------------- A12 --------
def test_boundary_volume_closed_forms():
    # m_V = 3 / (W(0) + sum_solid W): interior of a solid lattice 3 / (k S / h^3)
    # with S the lattice shell sum; a lone solid 3 / W(0); fluid does not count
    n = 8
    pos = lattice(n)
    s, _ = make(pos, material=np.zeros(len(pos), np.int32), dynamic=np.zeros(len(pos), np.int32), extra=10)
    x = f32v(s, "x")
    inner = interior(x, n)
    S = float(W64(0.0) + np.sum(W64(DIST)))
    np.testing.assert_allclose(f32v(s, "m_V")[inner], 3.0 / S, rtol=2e-6)
    a = 0.3 * H
    s, slot = make([[0.5, 0.5, 0.5], [0.5 + a, 0.5, 0.5], [0.5 - a, 0.5, 0.5]], material=[0, 1, 1],
                   dynamic=[0, 1, 1])
    assert float(s.field("m_V")[slot[0]]) == pytest.approx(3.0 / float(W64(0.0)), rel=1e-6)
    s, slot = make([[0.5, 0.5, 0.5], [0.5 + a, 0.5, 0.5]], material=[0, 0], dynamic=[0, 0])
    np.testing.assert_allclose(s.field("m_V")[slot], 3.0 / float(W64(0.0) + W64(a)), rtol=1e-6)


# ------------------------------------------------------------- A13 --------
def test_init_cylinder_carves_object_two():
    rng = np.random.default_rng(12)
    pos = rng.uniform(0.4, 0.9, (2000, 3)).astype(np.float32)
    c = _cfg()
    c["reservedCapacity"] = 2000
    s = R.RefSim({"Configuration": c})
    s.add_particles(2, 1500, pos[:1500], np.zeros((1500, 3), np.float32), np.full(1500, 1000, np.float32),
                    np.zeros(1500, np.float32), np.zeros(1500, np.int32), np.zeros(1500, np.int32),
                    np.zeros((1500, 3), np.int32))
    s.add_particles(1, 500, pos[1500:], np.zeros((500, 3), np.float32), np.full(500, 1000, np.float32),
                    np.zeros(500, np.float32), np.ones(500, np.int32), np.ones(500, np.int32),
                    np.zeros((500, 3), np.int32))
    # RefSim.__init__ already carved nothing (no particles then); carve now
    L.or_init_cylinder(s.sim, 2, np.float32(0.65), np.float32(0.5), np.float32(0.1))
    x = s.field("x")[:2000]
    d = np.sqrt((pos[:, 0] - np.float32(0.65)) ** 2 + (pos[:, 1] - np.float32(0.5)) ** 2)
    cut = np.zeros(2000, bool)
    cut[:1500] = d[:1500] > np.float32(0.1)
    assert cut.sum() > 100 and (~cut[:1500]).sum() > 100
    assert np.all(x[cut] == 0.0) and np.all(s.field("is_active")[:2000][cut] == 0)
    np.testing.assert_array_equal(x[~cut], pos[~cut])
    assert np.all(s.field("is_active")[:2000][~cut] == 1)


# ------------------------------------------------- per-particle maps ------
def test_per_particle_maps():
    # A17 multiply_time_step, A22 v*, A31 life time, A32 predict, A36 x_temp,
    # A37 advect: each is one line of the reference; checked on a mixed set
    rng = np.random.default_rng(13)
    pos = rng.uniform(0.05, 0.95, (400, 3))
    mat = (rng.random(400) < 0.7).astype(np.int32)
    dyn = mat.copy()
    dyn[:20] = 0                                      # a few static fluid particles
    s, slot = make(pos, rng.normal(size=(400, 3)), material=mat, dynamic=dyn,
                   vortexEnforcingDomainStart=[0.0, 0.0, 0.0], vortexEnforcingDomainEnd=[0.5, 1.0, 1.0])
    n = s.particle_num
    f = s.field("material")[:n] == 1
    dynm = s.field("is_dynamic")[:n] == 1
    acc = rng.normal(size=(n, 3)).astype(np.float32)
    s.field("acceleration")[:n] = acc
    alpha = rng.random(n).astype(np.float32)
    s.field("dfsph_factor")[:n] = alpha
    L.or_multiply_time_step(s.sim, 0, np.float32(4.0))
    np.testing.assert_array_equal(s.field("dfsph_factor")[:n], np.where(f, alpha * np.float32(4.0), alpha))
    v0 = s.field("v")[:n].copy()
    dt = np.float32(DT)
    L.or_compute_v_star(s.sim)
    np.testing.assert_array_equal(s.field("v_star")[:n], np.where(f[:, None], v0 + acc * dt, 0.0))
    lt = rng.random(n).astype(np.float32)
    s.field("life_time")[:n] = lt
    L.or_update_particle_life_time(s.sim)
    dom = f & (s.field("x")[:n, 0] <= np.float32(0.5))
    np.testing.assert_array_equal(s.field("life_time")[:n], np.where(dom, lt + dt, lt))
    L.or_predict_velocity(s.sim)
    v1 = np.where((f & dynm)[:, None], v0 + dt * acc, v0)
    np.testing.assert_array_equal(s.field("v")[:n], v1)
    x0 = s.field("x")[:n].copy()
    L.or_copy_x_temp(s.sim)
    np.testing.assert_array_equal(s.field("x_temp")[:n], x0)
    L.or_advect(s.sim)
    np.testing.assert_array_equal(s.field("x")[:n], np.where(dynm[:, None], x0 + dt * v1, x0))


# ------------------------------------------------------------- A38 --------
def test_bounce_back_clamps_and_reflects():
    # sph_base.py:118-123,149-179: clamp to [pad, size - pad] with pad = h, and
    # v -= (1 + 0.5)(v.n) n, n the normalised sum of the face normals
    pad = H
    hi = 1.0 - H
    pos = np.array([[0.5, 0.5, 0.5],            # inside: untouched
                    [0.99, 0.5, 0.5],           # beyond the upper x face
                    [0.01, 0.02, 0.5],          # lower x and y faces (corner)
                    [0.5, 0.5, pad],            # exactly on the lower pad: counts (<=)
                    [0.5, 0.5, hi],             # exactly on the upper bound: does not (>)
                    [0.99, 0.5, 0.97],          # upper x and z
                    [0.02, 0.5, 0.5],           # a solid beyond the face: untouched
                    [0.01, 0.99, 0.5]])         # lower x and upper y: normal (-1, 1, 0)
    vel = np.array([[1.0, 2.0, 3.0], [2.0, -1.0, 0.5], [-1.0, -3.0, 2.0], [0.3, 0.1, -2.0],
                    [0.0, 0.0, 4.0], [1.0, 0.0, 1.0], [-5.0, 0.0, 0.0], [-1.0, 2.0, 0.3]])
    mat = np.array([1, 1, 1, 1, 1, 1, 0, 1])
    s, slot = make(pos, vel, material=mat, dynamic=mat)
    L.or_enforce_boundary_3d(s.sim)
    x = s.field("x")[slot].astype(np.float64)
    v = s.field("v")[slot].astype(np.float64)

    def refl(vv, nrm):
        nrm = np.asarray(nrm, float) / np.linalg.norm(nrm)
        return vv - 1.5 * np.dot(vv, nrm) * nrm

    f32 = lambda a: np.float32(a).astype(np.float64)
    np.testing.assert_array_equal(x[0], pos[0].astype(np.float32))
    np.testing.assert_array_equal(v[0], vel[0])
    np.testing.assert_allclose(x[1], [f32(hi), 0.5, 0.5], rtol=1e-7)
    np.testing.assert_allclose(v[1], [-1.0, -1.0, 0.5], rtol=1e-6)         # normal part * -0.5
    np.testing.assert_allclose(x[2], [f32(pad), f32(pad), 0.5], rtol=1e-7)
    np.testing.assert_allclose(v[2], refl(vel[2], [-1, -1, 0]), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(v[3], [0.3, 0.1, 1.0], rtol=1e-6)
    np.testing.assert_array_equal(v[4], vel[4])
    np.testing.assert_allclose(v[5], refl(vel[5], [1, 0, 1]), rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(x[6], pos[6].astype(np.float32))
    np.testing.assert_array_equal(v[6], vel[6])
    np.testing.assert_allclose(x[7], [f32(pad), f32(hi), 0.5], rtol=1e-7)
    np.testing.assert_allclose(v[7], refl(vel[7], [-1, 1, 0]), rtol=1e-6, atol=1e-6)


# ------------------------------------------------------------- B8 ---------
def test_ring_velocity_on_axis_closed_form():
    # on the axis every discretisation point is at the same distance, so the
    # K-point sum equals the thin-ring law exactly: u = G R^2 / (2 (R^2 + z^2)^1.5)
    # along +n (right-hand rule), times the core factor 1 - exp(-(R^2 + z^2)/sigma^2)
    c = np.array([[0.3, 0.4, 0.5]])
    nvec = np.array([[1.0, 2.0, -2.0]]) / 3.0
    Rr, G, sig = np.array([0.12]), np.array([1.7]), 0.04
    z = np.linspace(-0.3, 0.3, 13)
    x = (c + z[:, None] * nvec).astype(np.float32)
    zz = np.einsum("ij,j->i", x.astype(np.float64) - c, nvec[0])
    u = R.ring_velocity(x, c, nvec, Rr, G, sig, K=64).astype(np.float64)
    d2 = Rr[0] ** 2 + zz ** 2
    mag = G[0] * Rr[0] ** 2 / (2 * d2 ** 1.5) * (1 - np.exp(-d2 / sig ** 2))
    np.testing.assert_allclose(u, mag[:, None] * nvec, rtol=1e-5, atol=1e-6 * mag.max())


# --------------------------------------------------- A30, B1, B2 ----------
def test_near_field_point_vortex_closed_forms():
    # both particles within h: near part of vort2vel and of the FP64 direct sum
    # = (0.1592 V) (r_xy x w_j) / |r_xy|^2 with the 2-D norm (the z offset is
    # ignored); exp(-|r|/1e-6) underflows, so K = 1/|r_xy|^2 exactly
    r = np.array([0.3, -0.4, 0.5]) * H
    s, slot = make([[0.5, 0.5, 0.5], [0.5 + r[0], 0.5 + r[1], 0.5 + r[2]]])
    w = np.array([[0.5, -1.0, 2.0], [-3.0, 0.25, 1.0]])
    s.field("d_vort_smoothed")[slot] = w
    s.field("kernel_density")[slot] = 1.0
    s.field("v")[:2] = 1e3                                 # the |v_vort| < |v| gate passes
    fac = 0.1592 * 0.2 * MV
    out = np.zeros((s.particle_max_num, 3), np.float32)
    s.i32("num_samples")[0] = 0
    L.or_vort2vel(s.sim, R._ptr(out))
    d64 = s.direct_bs_f64()
    x = s.field("x")[slot].astype(np.float64)
    for i, j in ((0, 1), (1, 0)):
        rij = x[i] - x[j]
        rxy = np.array([rij[0], rij[1], 0.0])
        want = fac * np.cross(rxy, w[j]) / np.dot(rxy, rxy)
        np.testing.assert_allclose(out[slot[i]], want, rtol=2e-6)
        np.testing.assert_allclose(d64[slot[i]], want, rtol=1e-9)


def test_far_field_point_vortex_uses_the_3d_norm():
    # far part of the FP64 direct sum: (0.1592 V) (r_xy x w) / |r|^2, 3-D norm
    r = np.array([2.1, -1.3, 1.7]) * H
    s, slot = make([[0.4, 0.5, 0.4], [0.4 + r[0], 0.5 + r[1], 0.4 + r[2]]])
    w = np.array([[0.0, 0.0, 0.0], [1.0, -2.0, 3.0]])
    s.field("d_vort_smoothed")[slot] = w
    fac = 0.1592 * 0.2 * MV
    d64 = s.direct_bs_f64()
    x = s.field("x")[slot].astype(np.float64)
    rij = x[0] - x[1]
    want = fac * np.cross([rij[0], rij[1], 0.0], w[1]) / np.dot(rij, rij)
    np.testing.assert_allclose(d64[slot[0]], want, rtol=1e-12)


def test_vort2vel_equals_direct_sum_for_constant_kden_and_all_samples():
    # With kden = c everywhere and sample_idx = every in-domain fluid slot, the
    # MC weight S_i / (kden_s n_i) is 1 for every accepted sample:
    #   S_i = kden_sum - kden_i - sum_nbr kden_j = c (F - 1 - |N(i)|),
    #   n_i = #{s : |r_is| > h} = F - 1 - |N(i)|  (no pair at exactly h),
    # so vort2vel's v_vort is the direct Biot-Savart sum (pinned above by the
    # point-vortex closed forms), to f32 rounding.  No fluid in cell 0, whose
    # particles no neighbour list holds.
    rng = np.random.default_rng(21)
    n = 9
    pos = lattice(n) + rng.uniform(-0.1 * RAD, 0.1 * RAD, (n ** 3, 3))
    s, slot = make(pos)
    F = s.particle_num
    assert F == n ** 3 and s.grid_particles_num()[0] == 0
    # a coherent z-vorticity plus noise, so that the far part does not cancel
    s.field("d_vort_smoothed")[:F] = (np.array([0.0, 0.0, 1.0]) + 0.3 * rng.normal(size=(F, 3))).astype(np.float32)
    s.field("kernel_density")[:F] = 2.5
    lib = R.lib()
    import ctypes
    ctypes.cast(lib.or_ptr(s.sim, b"kden_sum"), ctypes.POINTER(ctypes.c_double))[0] = 2.5 * F
    s.field("sample_idx")[:F] = np.arange(F, dtype=np.int32)
    s.i32("num_samples")[0] = F
    out = np.zeros((s.particle_max_num, 3), np.float32)
    L.or_vort2vel(s.sim, R._ptr(out))
    d64 = s.direct_bs_f64()
    err = np.linalg.norm(out[:F] - d64[:F], axis=1) / np.linalg.norm(d64[:F], axis=1)
    assert err.max() < 2e-5, err.max()
    # the far part carries real weight here: dropping it, or flipping its sign,
    # moves v_vort far beyond the tolerance above
    s.i32("num_samples")[0] = 0
    near = np.zeros_like(out)
    L.or_vort2vel(s.sim, R._ptr(near))
    far = d64[:F] - near[:F]
    assert np.linalg.norm(far) > 0.3 * np.linalg.norm(d64[:F])


def test_rigid_reaction_is_the_opposite_impulse():
    # dimcv_sph.py:546-549: the reaction on a dynamic rigid neighbour j is the
    # fluid particle's velocity change from j, reversed and scaled by rho_i /
    # rho_j (equal and opposite momentum for m = m_V0 rho): m_j dt a_j = -m_i dv_i
    a = 0.35 * H
    pos = [[0.5, 0.5, 0.5], [0.5 - a, 0.5, 0.5]]
    s, slot = make(pos, material=[1, 0], dynamic=[1, 1], density=[1000.0, 600.0])
    n = s.particle_num
    s.field("density")[slot[0]] = 950.0
    s.field("dfsph_factor")[:n] = 1.0
    s.field("density_adv")[slot[0]] = RHO0 + 4.0      # kappa_i = 4
    s.field("v")[:n] = 0.0
    s.field("acceleration")[:n] = 0.0
    L.or_pressure_solve_iteration_kernel(s.sim)
    dv_i = s.field("v")[slot[0]].astype(np.float64)
    acc_j = s.field("acceleration")[slot[1]].astype(np.float64)
    assert dv_i[0] > 0                                   # pushed away from the rigid particle (+x)
    np.testing.assert_allclose(acc_j * DT * 600.0, -dv_i * 950.0, rtol=1e-5)
    # a static solid receives nothing
    s2, slot2 = make(pos, material=[1, 0], dynamic=[1, 0])
    s2.field("dfsph_factor")[:2] = 1.0
    s2.field("density_adv")[slot2[0]] = RHO0 + 4.0
    s2.field("acceleration")[:2] = 0.0
    L.or_pressure_solve_iteration_kernel(s2.sim)
    assert np.all(s2.field("acceleration")[slot2[1]] == 0.0)
