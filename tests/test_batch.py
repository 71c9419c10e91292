"""Batched scenes: every scene of a batch is bitwise its solo run.

The solo run of a scene is ``scenes.build_ring_system`` + the native
``KarmanVortexSolver.step`` (itself bit-exact against the oracle in the exact
far-field mode, test_gpu_parity.py).  The batch holds scenes of uneven sizes
and viscosities, one of them in a trimmed box whose outer lattice layer is
deactivated by the first grid build (so the scenes' slot ranges move).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PS_COMPARE = ["x", "v", "density", "pressure", "dfsph_factor", "density_adv", "vorticity", "vorticity_vis",
              "life_time", "material", "object_id", "grid_ids"]
SOLVER_COMPARE = ["kernel_density", "kinematic_vort_num", "d_vort_smoothed", "d_vort", "v_star", "grad_v",
                  "x_temp", "v_vort"]


def _specs():
    from batch import RingSpec
    return [RingSpec(12, 10, 8, n_rings=2, seed=11, viscosity=3e-4),
            RingSpec(20, 14, 12, n_rings=3, seed=12, viscosity=5e-3, trim=0.6),
            RingSpec(16, 16, 16, n_rings=4, seed=13, viscosity=1e-3, mc_samples=300)]


def _grid_ids_local(batch, s, a):
    return a - batch.cell_base[s]


def _compare_scene(batch, s, ps, sv):
    import _native as N
    sc = batch.scene_counters_host()[s]
    k = int(sc[N.SC_SLOTS])
    assert k == ps.particle_num[None], (s, k, ps.particle_num[None])
    assert int(sc[N.SC_FLUID]) == ps.fluid_particle_num[None]
    bad = []
    for name in PS_COMPARE:
        b = batch.scene_field(name, s)
        r = getattr(ps, name).to_numpy()[:k]
        if name == "grid_ids":
            b = _grid_ids_local(batch, s, b)
        if not np.array_equal(b.view(np.uint8), np.ascontiguousarray(r).view(np.uint8)):
            bad.append(name)
    for name in SOLVER_COMPARE:
        b = batch.scene_field(name, s)
        r = getattr(sv, name).to_numpy()[:k]
        if not np.array_equal(b.view(np.uint8), np.ascontiguousarray(r).view(np.uint8)):
            bad.append(name)
    assert not bad, (s, bad)
    m = int(sc[N.SC_NUM_SAMPLES])
    assert m == sv.num_samples[None]
    assert np.array_equal(batch.scene_samples(s), sv.sample_idx.to_numpy()[:m])
    assert tuple(batch.last_iterations[s]) == (sv.last_iterations_v, sv.last_iterations_p)


@pytest.mark.parametrize("far", ["fast", "kahan", "exact"])
def test_batch_scenes_equal_solo_runs(far):
    from batch import SceneBatch, solo_system
    specs = _specs()
    batch = SceneBatch(specs, far_field=far)
    solos = [solo_system(sp, far_field=far) for sp in specs]
    for s, (ps, sv) in enumerate(solos):
        assert batch.scene_counters_host()[s][3] == ps.fluid_particle_num[None]
    for step in range(3):
        batch.step()
        for ps, sv in solos:
            sv.step()
        for s, (ps, sv) in enumerate(solos):
            _compare_scene(batch, s, ps, sv)
    # the trimmed scene lost its outer layer; the others kept every particle
    assert batch.scene_counters_host()[1][6] == 19 * 14 * 12
    assert batch.scene_counters_host()[0][6] == 12 * 10 * 8


@pytest.mark.parametrize("max_error_v", [1e-6, 1e-10], ids=["own-counts", "iteration-cap"])
def test_iterating_solves_equal_solo_runs(max_error_v):
    """A tighter divergence tolerance (max_error_V, dimcv_sph.py:24,389) makes
    the scenes' divergence solves iterate, each to its own count (or all to
    the 100-iteration cap, 300 grid barriers per solve), inside the one
    persistent launch per solve; fields and both iteration counts stay
    bitwise those of the solo runs, eagerly and from the captured graph."""
    from batch import SceneBatch, solo_system
    specs = _specs()
    eager = SceneBatch(specs, far_field="fast")
    graph = SceneBatch(specs, far_field="fast")
    solos = [solo_system(sp, far_field="fast") for sp in specs]
    for sv in [eager.solver, graph.solver] + [sv for _, sv in solos]:
        sv.max_error_V = max_error_v
    graph.capture()
    kv = set()
    for step in range(3):
        u = eager.uniforms()
        graph.uniforms()   # same seeded draws: keep the generators in step
        eager.step(u)
        graph.step_graph(u)
        for ps, sv in solos:
            sv.step()
        for s, (ps, sv) in enumerate(solos):
            _compare_scene(eager, s, ps, sv)
            _compare_scene(graph, s, ps, sv)
            kv.add(int(eager.last_iterations[s][0]))
    graph.release_graph()
    if max_error_v > 1e-8:
        assert max(kv) > 1 and len(kv) > 1, kv
    else:
        assert kv == {eager.solver.m_max_iterations_v}, kv


def test_graph_steps_equal_eager_steps_and_autotune_keeps_state():
    from batch import SceneBatch, autotune
    specs = _specs()
    eager = SceneBatch(specs)
    graph = SceneBatch(specs)
    cfg = autotune(graph, far_workgroups=(256, 2048), nbr_blocks=(64, 256), steps=1)
    assert cfg["far_workgroups"] in (256, 2048) and cfg["nbr_block"] in (64, 256)
    graph.capture()
    for step in range(4):   # both parities of the buffer swap, twice
        eager.step()
        graph.step_graph()
        for s in range(len(specs)):
            for name in ["x", "v", "density", "density_adv", "dfsph_factor"]:
                a, b = eager.scene_field(name, s), graph.scene_field(name, s)
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (step, s, name)
            assert np.array_equal(eager.scene_samples(s), graph.scene_samples(s))
        assert np.array_equal(eager.last_iterations, graph.last_iterations)
    graph.release_graph()


def test_batched_particle_on_upper_face_keys_like_solo():
    """A particle exactly on a scene's upper z face keys into the next row's
    first cell as in a solo run (the reference's flat index), and one on the
    upper x face (flat index past the scene's table) leaves the scene; all
    fields stay bitwise equal to the solo runs."""
    import _native as N
    from batch import SceneBatch, solo_system
    specs = _specs()
    batch = SceneBatch(specs, far_field="exact")
    solos = [solo_system(sp, far_field="exact") for sp in specs]
    bx = batch.ps.x.to_numpy()
    lost = {}
    for s, axis in ((0, 2), (2, 0)):
        ps, sv = solos[s]
        k = ps.particle_num[None]
        x = ps.x.to_numpy()
        start = int(batch.scene_counters_host()[s][N.SC_SLOT_START])
        assert np.array_equal(bx[start:start + k], x[:k])      # same slots, same order
        end = np.float32(ps.params.domain_end[axis])
        if axis == 2:   # a particle of the first x column: its next-row cell is in the table
            cand = np.nonzero(x[:k, 0] < np.float32(0.04))[0]
            j = int(cand[np.argmax(x[cand, 2])])
        else:           # the last x layer: flat index past the table
            j = int(np.argmax(x[:k, 0]))
        x[j, axis] = end
        ps.x.from_numpy(x)
        bx[start + j, axis] = end
        lost[s] = 1 if axis == 0 else 0
    batch.ps.x.from_numpy(bx)
    n0 = [ps.particle_num[None] for ps, _ in solos]
    for step in range(2):
        batch.step()
        for ps, sv in solos:
            sv.step()
        for s, (ps, sv) in enumerate(solos):
            _compare_scene(batch, s, ps, sv)
        if step == 0:   # the first build kept the z-face particle, dropped the x-face one
            for s, d in lost.items():
                assert solos[s][0].particle_num[None] == n0[s] - d


@pytest.mark.parametrize("path", ["eager", "graph"])
def test_list_width_overflow_raises_and_keeps_the_scenes(path):
    """A neighbour list wider than nbr_kmax: the eager step raises after the
    list build, the graph step (one launch, device-side solves) after its
    status readback with the state writers skipped on the device; either
    way every scene keeps its state."""
    from batch import SceneBatch
    specs = _specs()
    batch = SceneBatch(specs, far_field="fast")
    batch.step()
    # lists of the allocated width, read only up to 16 entries (~26 neighbours)
    batch.ps.nbr_kmax = 16
    batch.ps.params.nbr_kmax = 16
    if path == "graph":
        batch.capture()
    def snapshot(s):
        # the eager build may have re-sorted the slots: compare in x order
        x, v, lt = (batch.scene_field(n, s).copy() for n in ("x", "v", "life_time"))
        o = np.lexsort(x.T)
        return x[o], v[o], lt[o]

    before = [snapshot(s) for s in range(len(specs))]
    with pytest.raises(RuntimeError, match="neighbour list width"):
        if path == "graph":
            batch.step_graph()
        else:
            batch.step()
    for s in range(len(specs)):
        for a, b, n in zip(before[s], snapshot(s), ("x", "v", "life_time")):
            assert np.array_equal(a, b), (s, n)
    if path == "graph":
        batch.release_graph()
