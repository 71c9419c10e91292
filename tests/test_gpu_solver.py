"""GPU parity of the far-field modes, the Jacobi loops and bounce-back.

* The fast and Kahan far fields (the modes the 1M workload runs) are
  compared on the far part itself: with identical state, samples and
  neighbour lists, the exact mode's v_vort is bitwise the oracle's, and the
  near part is computed bit-identically in every mode, so
  v_vort(fast) - v_vort(exact) is the fast far part's deviation from the
  oracle's far part.  It is held to 1e-5 relative (the FP32 tolerance) at
  1,440 targets (16 source splits), 131,072 (8 splits) and 1,048,576 (two
  halves), and the same comparison against the oracle with its far part
  sign-flipped must fail, so the test can fail.
* Kahan accumulation (B4): its error against the FP64 direct sum is no
  larger than naive FP32's.
* A tightened divergence tolerance makes the divergence solve iterate
  (k_v from 2 up to the 100-iteration cap); every field stays bit-exact.
* bounceBackBoundary: the clamp-and-reflect pass after each step, bit-exact.
"""
import copy

import numpy as np
import pytest

import harness as H
import _native as N

pytestmark = pytest.mark.gpu


def _far_setup(nx, ny, nz, m, seed):
    """Ring lattice on both sides with a shared vorticity field, KVN, AKDE
    and inverse-CDF samples (all bitwise equal between GPU and oracle)."""
    import oracle.refsim as R
    ps, solver, ref, info = H.make_ring_pair(nx, ny, nz, n_rings=4, seed=seed, mc_samples=m,
                                             far_field="exact")
    rng = np.random.default_rng(seed)
    n = ps.particle_max_num
    pn = ps.particle_num[None]
    # coherent z-vorticity plus noise, so the far part does not cancel
    w = (np.array([0.0, 0.0, 30.0]) + 10.0 * rng.normal(size=(n, 3))).astype(np.float32)
    kvn = (0.05 + rng.random(n)).astype(np.float32)
    solver.d_vort_smoothed.from_numpy(w)
    ref.field("d_vort_smoothed")[:] = w
    solver.kinematic_vort_num.from_numpy(kvn)
    ref.field("kvn")[:] = kvn
    ps.ensure_neighbor_lists()
    solver.compute_kernel_density()
    R.lib().or_compute_kernel_density(ref.sim)
    assert np.array_equal(solver.kernel_density.to_numpy()[:pn], ref.field("kernel_density")[:pn])
    assert solver.kden_sum[None] == ref.kden_sum
    u = rng.random(m)
    solver.importance_sample(u, m)
    ref.cdf_sample(u, m)
    assert np.array_equal(solver.sample_idx.to_numpy()[:m], ref.field("sample_idx")[:m])
    return ps, solver, ref, rng


def _v_vort(ps, solver, mode):
    ps.params.far_mode = mode
    solver.v_vort.fill(0.0)
    solver.vort2vel()
    return solver.v_vort.to_numpy()


@pytest.mark.parametrize("shape,m,n_targets", [((12, 12, 10), 256, None),
                                                ((64, 64, 32), 1024, 3000),
                                                ((128, 128, 64), 4096, 2000)],
                         ids=["1440-16splits", "131072-8splits", "1M-2halves"])
def test_fast_and_kahan_far_field_vs_oracle(shape, m, n_targets):
    ps, solver, ref, rng = _far_setup(*shape, m, seed=31)
    pn = ps.particle_num[None]
    splits = N.lib().dimcv_far_split_count(pn)
    assert splits == {1440: 16, 131072: 8, 1048576: 2}[pn]
    slots = np.arange(pn) if n_targets is None else np.sort(rng.choice(pn, n_targets, replace=False))
    vv_ref, far_ref, added = ref.vort2vel_targets(slots)
    assert np.all(added > 0)
    far_share = np.linalg.norm(far_ref) / np.linalg.norm(vv_ref)
    assert far_share > 0.05, far_share
    exact = _v_vort(ps, solver, N.FAR_EXACT)[slots]
    np.testing.assert_array_equal(exact, vv_ref)            # exact mode: bitwise the oracle
    for mode in (N.FAR_FAST, N.FAR_KAHAN):
        got = _v_vort(ps, solver, mode)[slots].astype(np.float64)
        dev = got - exact                                     # deviation of the far part
        err = np.linalg.norm(dev) / np.linalg.norm(far_ref)
        assert err < 1e-5, (mode, err)
        per = np.linalg.norm(dev, axis=1) / np.maximum(np.linalg.norm(far_ref, axis=1), 1e-30)
        assert np.quantile(per, 0.99) < 1e-4, (mode, np.quantile(per, 0.99))
        # planted error on the oracle side: a sign-flipped far part must fail
        flipped = vv_ref.astype(np.float64) - 2 * far_ref
        assert np.linalg.norm(got - flipped) / np.linalg.norm(far_ref) > 1.0


def test_kahan_error_no_larger_than_naive():
    """B4: against the FP64 direct Biot-Savart sum, Kahan-compensated f32
    accumulation errs no more than naive f32 (65,536 particles)."""
    ps, solver, ref, info = H.make_ring_pair(64, 64, 16, n_rings=4, seed=17, mc_samples=256,
                                             far_field="exact")
    rng = np.random.default_rng(17)
    n = ps.particle_max_num
    w = (np.array([0.0, 0.0, 30.0]) + 10.0 * rng.normal(size=(n, 3))).astype(np.float32)
    solver.d_vort_smoothed.from_numpy(w)
    pn = ps.particle_num[None]
    d64 = solver.direct_biot_savart("f64")[:pn]
    d32 = solver.direct_biot_savart("f32")[:pn].astype(np.float64)
    dk = solver.direct_biot_savart("kahan")[:pn].astype(np.float64)
    e32 = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    ek = np.linalg.norm(dk - d64) / np.linalg.norm(d64)
    assert ek <= e32, (ek, e32)
    assert e32 < 1e-5
    # per particle the compensated sum is never worse by more than its own rounding
    pe32 = np.linalg.norm(d32 - d64, axis=1)
    pek = np.linalg.norm(dk - d64, axis=1)
    assert np.mean(pek <= pe32 + 1e-7 * np.linalg.norm(d64, axis=1)) > 0.99


@pytest.mark.parametrize("max_error_v,steps", [(1e-6, 14), (1e-8, 4)], ids=["kv-iterates", "kv-cap-100"])
def test_divergence_solve_iterates_bitexact(max_error_v, steps):
    """A20: with the divergence tolerance tightened (the reference's
    max_error_V, dimcv_sph.py:24,389) the solve iterates; fields, iteration
    counts and sample sets stay bit-exact against the oracle every step."""
    scene = H.load_scene("mini_karman.json")
    ps, solver, ref = H.make_karman_pair(scene)
    solver.max_error_V = max_error_v
    ref.max_error_V = max_error_v
    rng = np.random.default_rng(7)
    kvs = []
    for k in range(steps):
        u1 = rng.random(ps.particle_max_num, dtype=np.float32)
        u2 = rng.random(ps.particle_max_num, dtype=np.float32)
        solver.step(u1, u2)
        ref.step(u1, u2)
        bad = H.mismatches(ps, solver, ref)
        assert not bad, f"step {k}: {bad}"
        assert solver.last_iterations_v == ref.last_kv, (k, solver.last_iterations_v, ref.last_kv)
        assert solver.last_iterations_p == ref.last_kp
        kvs.append(ref.last_kv)
    assert max(kvs) > 1, kvs
    if max_error_v == 1e-8:
        assert 100 in kvs, kvs


@pytest.mark.parametrize("chunk", [0, 2, None], ids=["host-adaptive", "host-every-2", "device-loop"])
def test_divergence_solve_iterates_fixed_m_step(chunk):
    """The same for the fixed-M one-call step (dimcv_step_begin/finish) on a
    ring lattice: the first host convergence check after the previous step's
    iteration count (chunk 0), every 2 iterations, or no host check at all
    (chunk unset = -1: the remaining iterations in one persistent launch per
    solve, the counters read once at the end of the step)."""
    ps, solver, ref, info = H.make_ring_pair(12, 12, 10, n_rings=2, seed=3, mc_samples=128, far_field="exact")
    if chunk is not None:
        ps.cfg.config["Configuration"]["jacobiChunk"] = chunk
    solver.max_error_V = 1e-7
    ref.max_error_V = 1e-7
    rng = np.random.default_rng(9)
    kvs = []
    for k in range(5):
        u = rng.random(128)
        solver.step(u_cdf=u)
        ref.step(u_cdf=u)
        bad = H.mismatches(ps, solver, ref)
        assert not bad, f"step {k}: {bad}"
        assert solver.last_iterations_v == ref.last_kv
        assert solver.last_iterations_p == ref.last_kp
        kvs.append(ref.last_kv)
    assert max(kvs) > 1, kvs


def _expanding_ring_pair():
    """A 12 x 12 x 10 ring lattice two spacings inside its box (clear of the
    bounce padding h = 2 spacings), expanding at 150/s on top of the ring
    velocity, so that its outer layers cross the padding within a few steps."""
    import scenes
    from config_builder import SimConfig
    from oracle.refsim import RefSim
    from particle_system import ParticleSystem
    info = scenes.make_ring_scene(12, 12, 10, n_rings=2, seed=5, mc_samples=128, far_field="exact")
    d = 0.02
    pos = (info["pos"] + np.float32(2 * d)).astype(np.float32)
    c = pos.mean(0)
    vel = (info["vel"] + 150.0 * (pos - c)).astype(np.float32)
    cfg = info["scene"]["Configuration"]
    cfg["domainEnd"] = [16 * d, 16 * d, 14 * d]
    cfg["bounceBackBoundary"] = True
    ps = ParticleSystem(SimConfig(scene=copy.deepcopy(info["scene"])))
    scenes.add_ring_particles(ps, pos, vel)
    solver = ps.build_solver()
    solver.initialize()
    ref = RefSim(copy.deepcopy(info["scene"]))
    n = pos.shape[0]
    ref.add_particles(0, n, pos, vel, np.full(n, 1000.0, np.float32), np.zeros(n, np.float32),
                      np.ones(n, np.int32), np.ones(n, np.int32),
                      np.tile(np.array([[50, 100, 200]], np.int32), (n, 1)))
    ref.initialize()
    return ps, solver, ref


@pytest.mark.parametrize("kind", ["karman", "ring"])
def test_bounce_back_bitexact(kind):
    """A38: enforce_boundary_3D after every step (karman_vortex.py:226-230,
    sph_base.py:149-179), bit-exact against the oracle's restatement; in the
    expanding ring lattice the outer layers cross the padding, are clamped
    onto it and reflected."""
    if kind == "karman":
        scene = H.load_scene("mini_karman.json")
        scene = copy.deepcopy(scene)
        scene["Configuration"]["bounceBackBoundary"] = True
        ps, solver, ref = H.make_karman_pair(scene)
        steps = 16
    else:
        ps, solver, ref = _expanding_ring_pair()
        steps = 5
    rng = np.random.default_rng(3)
    x0 = ref.field("x")[: ref.particle_num].copy()
    for k in range(steps):
        u1 = rng.random(ps.particle_max_num, dtype=np.float32)
        u2 = rng.random(ps.particle_max_num, dtype=np.float32)
        u = rng.random(128)
        if kind == "karman":
            solver.step(u1, u2)
            ref.step(u1, u2)
        else:
            solver.step(u_cdf=u)
            ref.step(u_cdf=u)
        bad = H.mismatches(ps, solver, ref)
        assert not bad, f"step {k}: {bad}"
    if kind == "ring":
        # the pass acted: the outer layers crossed the padding and were
        # clamped onto it; every fluid particle lies within [h, size - h]
        h = np.float32(0.04)
        x = ref.field("x")[: ref.particle_num]
        hi = np.array(ps.params.bounce_hi, np.float32)
        assert np.isfinite(x).all()
        assert np.all(x >= h) and np.all(x <= hi)
        assert np.all(x0 > h)
        assert np.sum(np.any((x == h) | (x == hi), axis=1)) > 100


@pytest.mark.parametrize("kind", ["fixed-m-step", "fixed-m-step-async", "karman-substep"])
def test_list_width_overflow_fails_before_writing_state(kind):
    """A list wider than neighborListWidth raises RuntimeError before any
    writer of x, v, is_sample or life_time runs: the positions and
    velocities after the failed step are those before it (the grid build,
    the reference's initialize_particle_system, may only reorder slots).
    With async_steps the failed step reports at the next flush()."""
    from config_builder import SimConfig
    from particle_system import ParticleSystem
    import scenes
    if kind.startswith("fixed-m-step"):
        ps, solver, info = scenes.build_ring_system(12, 12, 10, n_rings=2, seed=3, mc_samples=128,
                                                    far_field="fast")
        solver.async_steps = kind.endswith("async")
        solver.step(u_cdf=np.random.default_rng(1).random(128))
        solver.flush()
        # lists of the allocated width, read only up to 16 entries: the
        # lattice has ~26 neighbours per particle
        ps.nbr_kmax = 16
        ps.params.nbr_kmax = 16
    else:
        scene = copy.deepcopy(H.load_scene("mini_karman.json"))
        scene["Configuration"]["neighborListWidth"] = 8
        ps = ParticleSystem(SimConfig(scene=scene))
        solver = ps.build_solver()
        solver.initialize()
    rng = np.random.default_rng(2)

    def snapshot():
        pn = ps.particle_num[None]
        x, v = ps.x.to_numpy()[:pn].copy(), ps.v.to_numpy()[:pn].copy()
        lt, smp = ps.life_time.to_numpy()[:pn].copy(), ps.is_sample.to_numpy()[:pn].copy()
        o = np.lexsort(x.T)
        return x[o], v[o], lt[o], smp[o]

    for attempt in range(2):
        if kind == "karman-substep" and attempt == 0:
            # let fluid in first (8 neighbours suffice for the first emitter layers)
            for _ in range(30):
                try:
                    solver.step(rng.random(ps.particle_max_num, dtype=np.float32),
                                rng.random(ps.particle_max_num, dtype=np.float32))
                except RuntimeError:
                    break
        before = snapshot()
        with pytest.raises(RuntimeError, match="neighbour list width"):
            if kind.startswith("fixed-m-step"):
                solver.step(u_cdf=rng.random(128))
                solver.flush()
            else:
                solver.cnt = 1     # no emission in this step
                solver.step(rng.random(ps.particle_max_num, dtype=np.float32),
                            rng.random(ps.particle_max_num, dtype=np.float32))
        after = snapshot()
        for a, b, name in zip(before, after, ("x", "v", "life_time", "is_sample")):
            assert np.array_equal(a, b), (attempt, name)


def test_async_steps_equal_synchronous_steps():
    """Pipelined steps (solver.async_steps: each step() returns once enqueued,
    its counts arrive at the next step() or flush()) against the oracle, bit
    for bit, on a ring lattice whose pressure solve iterates (compressed to
    0.92 spacing) and whose divergence solve iterates (maxErrorV 1e-7): the
    same kernels as the synchronous step, with the uniforms of consecutive
    steps in separate buffers."""
    import scenes
    from config_builder import SimConfig
    from oracle.refsim import RefSim
    from particle_system import ParticleSystem
    info = scenes.make_ring_scene(12, 12, 10, n_rings=2, seed=3, mc_samples=128, far_field="exact")
    pos = (info["pos"] * np.float32(0.92)).astype(np.float32)
    ps = ParticleSystem(SimConfig(scene=copy.deepcopy(info["scene"])))
    scenes.add_ring_particles(ps, pos, info["vel"])
    solver = ps.build_solver()
    solver.initialize()
    ref = RefSim(copy.deepcopy(info["scene"]))
    n = pos.shape[0]
    ref.add_particles(0, n, pos, info["vel"], np.full(n, 1000.0, np.float32), np.zeros(n, np.float32),
                      np.ones(n, np.int32), np.ones(n, np.int32),
                      np.tile(np.array([[50, 100, 200]], np.int32), (n, 1)))
    ref.initialize()
    solver.max_error_V = 1e-7
    ref.max_error_V = 1e-7
    solver.async_steps = True
    rng = np.random.default_rng(4)
    kv, kp = [], []
    for k in range(6):
        u = rng.random(128)
        solver.step(u_cdf=u)
        ref.step(u_cdf=u)
        kv.append(ref.last_kv)
        kp.append(ref.last_kp)
        if k >= 1:  # the previous step's counts arrived with this step
            assert (solver.last_iterations_v, solver.last_iterations_p) == (kv[-2], kp[-2]), k
    solver.flush()
    assert (solver.last_iterations_v, solver.last_iterations_p) == (kv[-1], kp[-1])
    bad = H.mismatches(ps, solver, ref)
    assert not bad, bad
    assert max(kv) > 1 and max(kp) > 1, (kv, kp)
