"""F2: the exporters read the device buffers.  export_ply formats the rows
on the device; its file is compared byte for byte with the reference's
writer restated on the oracle's (bit-identical) state, including planted
special values (NaN, infinities, signed zeros, subnormals, 1e30) and
inactive slots.  export_png selects its rows on the device; its image is
compared byte for byte with the same figure drawn from the oracle's state."""
import pathlib

import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu


def _stepped_pair(steps=30):
    ps, solver, ref = H.make_karman_pair(H.load_scene("mini_karman.json"))
    rng = np.random.default_rng(5)
    for _ in range(steps):
        u1 = rng.random(ps.particle_max_num, dtype=np.float32)
        u2 = rng.random(ps.particle_max_num, dtype=np.float32)
        solver.step(u1, u2)
        ref.step(u1, u2)
    assert not H.mismatches(ps, solver, ref)
    return ps, solver, ref


def test_export_ply_bytes_match_reference_writer(tmp_path):
    from oracle.refsim import export_ply
    ps, solver, ref = _stepped_pair()
    pn = ps.particle_num[None]
    # planted values and inactive slots, the same on both sides
    w = ps.vorticity.to_numpy()
    specials = np.float32([np.nan, np.inf, -np.inf, -0.0, 0.0, 1e-45, -3e-42, 1e30, -2.5e-7, 0.49999997e-7])
    w[:len(specials), 0] = specials
    w[len(specials):2 * len(specials), 2] = specials[::-1]
    ps.vorticity.from_numpy(w)
    ref.field("vorticity")[:] = w
    act = ps.is_active.to_numpy().astype(np.uint8)
    act[3:pn:97] = 0
    ps.is_active.from_numpy(act.astype(bool))
    ref.field("is_active")[:] = act
    a, b = tmp_path / "gpu", tmp_path / "ref"
    solver.export_ply(7, a)
    export_ply(ref, 7, b)
    ga, gb = (a / "frame_0007.ply").read_bytes(), (b / "frame_0007.ply").read_bytes()
    assert ga == gb
    assert ga.count(b"\n") == 15 + int(act[:pn].sum())
    assert b"nan" in ga and b"-inf" in ga and b"-0.0000000" in ga


def test_export_png_matches_reference_figure(tmp_path):
    pytest.importorskip("matplotlib")
    from oracle.refsim import export_png
    ps, solver, ref = _stepped_pair(12)
    a, b = tmp_path / "gpu", tmp_path / "ref"
    a.mkdir()
    b.mkdir()
    solver.export_png(3, a)
    export_png(ref, 3, b)
    assert (a / "vorticity_0003.png").read_bytes() == (b / "vorticity_0003.png").read_bytes()
