"""Benchmark: DIMCV-SPH particle-steps/s at 1M particles on one MI355X.

Workload (the 1M-particle configuration): 1,048,576 fluid particles on a
128 x 128 x 64 jittered cubic lattice (spacing 2r, r = 0.01, jitter +-0.1 r,
seed 2), initial velocity from 8 seeded Lamb-Oseen vortex rings (core sigma =
h = 4r), cubic spline h = 4r, rho0 = 1000, dt = 2e-4, nu = 1e-3, FP32,
M = 4096 inverse-CDF importance samples per step (MC-VVPE, "fast" far field).
A step is one full KarmanVortexSolver.step(): grid build + neighbour lists,
DFSPH divergence and pressure Jacobi solves, vorticity passes, KVN/AKDE,
importance sampling, velocity recovery, advection.

Also measured in the same run:
  * the dominant kernel (MC-VVPE far field) with HIP events on the launch
    stream -> roofline vs the FP32 vector peak;
  * the direct O(N^2) Biot-Savart comparator at N = 131,072 (64 x 64 x 32
    lattice, same generator) in FP32, against which the FP32-compensated
    MC-VVPE recovery is compared at its relative L2 error of the recovered
    field against the FP64 direct sum, M swept up to N;
  * the CPU oracle (a C++ restatement of the reference step, OpenMP) on a
    bounded sample of the same 1M workload.

  * batched scenes: 64 uneven ring scenes stepped by one batched, graph-
    replayed step against the same scenes stepped one after another;
  * differentiation: forward and adjoint times of a 32-step, 65,536-particle
    ring rollout with 16 circulations as parameters.

  * Jacobi iterations at scale: ms per divergence and per pressure
    iteration at 1M particles, from steps whose solves iterate.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                       [--no-direct] [--no-batching] [--batch-steps K] [--no-adjoint]
                       [--no-jacobi]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "dimcv-sph_amd")
for _p in (PKG, REPO):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "particle-steps/s at 1M, MC-VVPE vs direct Biot-Savart at fixed L2 vel err"
FP32_PEAK_TFLOPS = 157.3
# Issue ceiling of the far-field kernel as built: 14.75 vector instructions per
# (target, sample) pair (59 per four-sample trip in the gfx950 ISA of the
# loop), at most two vector
# instructions per SIMD quad-cycle (dual issue), 1024 SIMDs at 2.4 GHz.
VALU_PER_PAIR = 14.75
FP32_ISSUE_TFLOPS = round(1024 * 2.4e9 / 4 * 2 * 64 / VALU_PER_PAIR * 24 / 1e12, 1)
HBM_PEAK_GBS = 8000.0
FLOPS_PER_PAIR_MC = 24
FLOPS_PER_PAIR_DIRECT = 22


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_gpu(args):
    import numpy as np
    import torch
    import scenes

    t0 = time.time()
    ps, solver, info = scenes.build_ring_system(128, 128, 64, n_rings=8, seed=2, mc_samples=4096,
                                                far_field=args.far_field)
    torch.cuda.synchronize()
    n = ps.particle_num[None]
    log(f"1M scene ready in {time.time() - t0:.1f}s: {n} particles, grid {list(ps.grid_num)}")
    for _ in range(args.warmup):
        solver.step()
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} steps done (k_v={solver.last_iterations_v}, k_p={solver.last_iterations_p})")

    # HIP events around the far-field kernel, recorded by dimcv_substep on the
    # launch stream inside every timed step
    solver.time_far_field = True
    # timed: pipelined steps (async_steps): each step() returns once the step
    # is enqueued, and its counts arrive with the next step(), so the device
    # runs the next grid build while the host prepares it
    solver.async_steps = not args.sync_steps
    kv, kp, far = [], [], []
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        solver.step()
        kv.append(solver.last_iterations_v)
        kp.append(solver.last_iterations_p)
        far.append(solver.last_far_ms)
    solver.flush()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    if solver.async_steps:
        # each step() brought the previous step's counts: drop the first
        # (the last untimed step's) and add the last timed step's
        kv, kp, far = kv[1:], kp[1:], far[1:]
        kv.append(solver.last_iterations_v)
        kp.append(solver.last_iterations_p)
        far.append(solver.last_far_ms)
    # for reference: the synchronous step (each step() waits for the device)
    # on the next steps (the flow evolves, so later steps run a little slower)
    pipelined = solver.async_steps
    async_after_ms = None
    solver.async_steps = False
    n_sync = min(args.steps, 20)
    t0s = time.perf_counter()
    for _ in range(n_sync):
        solver.step()
    torch.cuda.synchronize()
    sync_ms = (time.perf_counter() - t0s) * 1e3 / n_sync
    if pipelined:
        solver.async_steps = True
        t0a = time.perf_counter()
        for _ in range(n_sync):
            solver.step()
        solver.flush()
        torch.cuda.synchronize()
        async_after_ms = (time.perf_counter() - t0a) * 1e3 / n_sync
        solver.async_steps = False
        log(f"after the timed steps: {n_sync} synchronous steps {sync_ms:.3f} ms/step, then {n_sync} "
            f"pipelined {async_after_ms:.3f} ms/step")
    n_fluid = ps.fluid_particle_num[None]
    sec = t2 - t1
    far_ms = float(np.mean(far))
    m = solver.num_samples[None]
    pairs = n_fluid * m
    achieved = FLOPS_PER_PAIR_MC * pairs / (far_ms * 1e-3) / 1e12
    log(f"timed {args.steps} steps: {sec * 1e3 / args.steps:.3f} ms/step ({sync_ms:.3f} synchronous), "
        f"far field {far_ms:.3f} ms ({achieved:.1f} TFLOP/s), k_v mean {np.mean(kv):.2f}, "
        f"k_p mean {np.mean(kp):.2f}")
    res = {
        "n_fluid": n_fluid, "sec": sec, "far_ms": far_ms, "achieved_tflops": achieved, "sync_ms": sync_ms,
        "async_after_ms": async_after_ms,
        "pairs": pairs, "kv": float(np.mean(kv)), "kp": float(np.mean(kp)), "m": m,
        "n_cells": int(ps.grid_num[0]) * int(ps.grid_num[1]) * int(ps.grid_num[2]),
    }
    if args.traffic_probe:
        # a plain copy of known size (64 Mi f32 read, 64 Mi f32 written): the
        # FETCH_SIZE / WRITE_SIZE counters of this kernel scale the far field's
        a = torch.ones(64 * 2 ** 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        log("traffic probe: 3 x elementwise copy of 256 MiB done")
    del solver, ps
    torch.cuda.empty_cache()
    return res, info


def _recovery_errors(ps, solver, mode, m, u, v0, ref64, fluid):
    """Run importance sampling (M = m) + vort2vel in far-field `mode` on the
    current state; return (ms, rel L2 error of the recovered field, rel L2
    error of the updated velocity) against the FP64 direct recovery."""
    import numpy as np
    import torch
    import _native as N
    ps.params.far_mode = {"fast": N.FAR_FAST, "kahan": N.FAR_KAHAN}[mode]
    ps.v.t.copy_(v0)
    solver.importance_sample(u[:m], m)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    solver.vort2vel()
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1])
    vv = solver.v_vort.to_numpy()[fluid].astype(np.float64)
    v_new = ps.v.to_numpy()[fluid].astype(np.float64)
    v_old = v0.cpu().numpy()[fluid].astype(np.float64)
    r = ref64[fluid]
    gate = np.linalg.norm(r, axis=1) < np.linalg.norm(v_old, axis=1)
    v_ref = v_old + gate[:, None] * r
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return ms, rel(vv, r), rel(v_new, v_ref)


def run_recovery_comparison(nx, ny, nz, ms_list, direct_reps=2):
    """MC-VVPE vs direct O(N^2) Biot-Savart on one state of an N = nx*ny*nz
    ring lattice: FP64 direct as the oracle, FP32 naive and Kahan direct
    timed, MC (FP32 naive and Kahan) for each M."""
    import numpy as np
    import torch
    import scenes

    ps, solver, info = scenes.build_ring_system(nx, ny, nz, n_rings=8, seed=2, mc_samples=4096,
                                                far_field="kahan")
    solver.step()  # populate every field the recovery reads
    ps.ensure_neighbor_lists()
    torch.cuda.synchronize()
    pn = ps.particle_num[None]
    fluid = np.zeros(ps.particle_max_num, bool)
    fluid[ps.fluid_list.to_numpy()[: ps._nfl]] = True
    v0 = ps.v.t.clone()
    t0 = time.perf_counter()
    ref64 = solver.direct_biot_savart("f64")
    f64_ms = (time.perf_counter() - t0) * 1e3
    out = {"n": pn, "direct_f64_ms": round(f64_ms, 2)}
    for prec in ("f32", "kahan"):
        solver.direct_biot_savart(prec)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(direct_reps):
            d = solver.direct_biot_savart(prec)
        ev[1].record()
        torch.cuda.synchronize()
        dms = ev[0].elapsed_time(ev[1]) / direct_reps
        err = float(np.linalg.norm(d[fluid] - ref64[fluid]) / np.linalg.norm(ref64[fluid]))
        out[f"direct_{prec}_ms"] = round(dms, 4)
        out[f"direct_{prec}_rel_l2_err"] = err
        out[f"direct_{prec}_tflops"] = round(FLOPS_PER_PAIR_DIRECT * pn * (pn - 1) / (dms * 1e-3) / 1e12, 2)
    u = np.random.default_rng(7).random(max(ms_list))
    sweep = []
    for m in ms_list:
        row = {"m": m}
        for mode in ("fast", "kahan"):
            ms, e_field, e_vel = _recovery_errors(ps, solver, mode, m, u, v0, ref64, fluid)
            row[f"{mode}_ms"] = round(ms, 4)
            row[f"{mode}_rel_l2_err_recovered"] = round(e_field, 6)
            row[f"{mode}_rel_l2_err_velocity"] = round(e_vel, 8)
        row["speedup_vs_direct_f32"] = round(out["direct_f32_ms"] / row["kahan_ms"], 2)
        sweep.append(row)
    out["mc"] = sweep
    # the comparison point: the recovered field v_vort (what the MC estimator
    # computes) against the FP64 direct sum; MC noise falls as M^-1/2
    ok = [r for r in sweep if r["kahan_rel_l2_err_recovered"] <= 1e-2]
    if ok:
        best = max(ok, key=lambda r: r["speedup_vs_direct_f32"])
        out["at_1e-2_recovered_error"] = {"reached": True, "m": best["m"], "kahan_ms": best["kahan_ms"],
                                          "speedup_vs_direct_f32": best["speedup_vs_direct_f32"],
                                          "rel_l2_err_recovered": best["kahan_rel_l2_err_recovered"]}
    else:
        # fit err = a M^-1/2 through the sweep; the M it predicts for 1e-2 and
        # the MC time there (time linear in M) against the direct sum
        ms_ = np.array([r["m"] for r in sweep], float)
        es = np.array([r["kahan_rel_l2_err_recovered"] for r in sweep], float)
        a = float(np.exp(np.mean(np.log(es) + 0.5 * np.log(ms_))))
        m_need = (a / 1e-2) ** 2
        t_per_m = float(np.mean([r["kahan_ms"] / r["m"] for r in sweep[-2:]]))
        out["at_1e-2_recovered_error"] = {
            "reached": False, "max_m_swept": int(ms_.max()),
            "rel_l2_err_recovered_at_max_m": float(es[np.argmax(ms_)]),
            "fit": f"err = {a:.3f} M^-1/2", "m_predicted_for_1e-2": int(round(m_need)),
            "predicted_speedup_vs_direct_f32": round(out["direct_f32_ms"] / (t_per_m * m_need), 4),
            "note": "the importance-sampled estimator does not reach 1e-2 at any M <= N on this state; "
                    "the north-star target of >= 20x the direct sum at <= 1e-2 is not met"}
    out["error_definitions"] = (
        "recovered: relative L2 error over fluid particles of the recovered field v_vort (near + MC far "
        "part) against the FP64 direct Biot-Savart field -- the comparison point; velocity: the same for "
        "the velocity after the gated update, dominated by the unchanged velocity, reported for reference "
        "only")
    ps.v.t.copy_(v0)
    log(f"recovery comparison at N={pn}: {json.dumps(out)}")
    del solver, ps
    torch.cuda.empty_cache()
    return out


def run_batching(n_scenes=64, steps=10, warmup=1, seed=7):
    """Batched scenes (64 ring scenes, N log-uniform in [8192, 131072], nu
    log-uniform in [1e-4, 1e-2], 2-6 rings, M = 2% of N): graph-replayed
    batched steps against the same scenes stepped one after another (each its
    own ParticleSystem + native substep), with every scene's final x and v
    compared bitwise between the two."""
    import numpy as np
    import torch
    from batch import SceneBatch, autotune, random_specs, solo_system

    specs = random_specs(n_scenes, seed)
    n_total = sum(sp.n for sp in specs)
    t0 = time.time()
    batch = SceneBatch(specs)
    log(f"batch of {n_scenes} scenes ({n_total} particles) ready in {time.time() - t0:.1f}s")
    for _ in range(warmup):
        batch.step()
    tune = autotune(batch)
    log(f"autotune: {tune}")
    batch.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.step_graph()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    log(f"batched: {steps} graph steps in {t_batch * 1e3:.1f} ms")
    t_seq, equal = 0.0, 0
    for s, sp in enumerate(specs):
        ps, sv = solo_system(sp)
        for _ in range(warmup):
            sv.step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            sv.step()
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t1
        k = ps.particle_num[None]
        same = all(np.array_equal(batch.scene_field(f, s).view(np.uint8),
                                  getattr(ps, f).to_numpy()[:k].view(np.uint8)) for f in ("x", "v"))
        equal += int(same)
        del ps, sv
        if (s + 1) % 16 == 0:
            log(f"sequential: {s + 1}/{n_scenes} scenes, {t_seq * 1e3:.1f} ms so far, {equal} bitwise equal")
    batch.release_graph()
    out = {"workload": f"{n_scenes} ring scenes, N log-uniform [8192, 131072], M = 2% N",
           "n_scenes": n_scenes, "n_particles": n_total, "steps": steps,
           "batched_ms_per_step": round(t_batch * 1e3 / steps, 4),
           "sequential_ms_per_step": round(t_seq * 1e3 / steps, 4),
           "batched_particle_steps_per_s": round(n_total * steps / t_batch, 1),
           "sequential_particle_steps_per_s": round(n_total * steps / t_seq, 1),
           "speedup": round(t_seq / t_batch, 3),
           "scenes_bitwise_equal": f"{equal}/{n_scenes}",
           "autotune": {k: tune[k] for k in ("far_workgroups", "nbr_block", "trials")}}
    log(f"batching: {json.dumps(out)}")
    return out


def run_adjoint(steps=32, checkpoint=8):
    """Differentiation workload: 65,536 particles (64 x 32 x 32 lattice, two
    spacings clear of every wall), 16
    rings whose circulations are the parameters, a 32-step FP32 rollout with
    checkpoints every 8 steps, L2 loss against the rollout of seeded
    perturbed circulations; forward and adjoint (backward) times."""
    import numpy as np
    import torch
    from adjoint import RingRollout, rollout

    t0 = time.time()
    model = RingRollout(64, 32, 32, n_rings=16, steps=steps, checkpoint=checkpoint, mc_samples=1024, seed=5, pad=2)
    g0 = model.gamma0.copy()
    gt = g0 + np.random.default_rng(3).normal(0.0, 0.1, g0.shape)
    v_target = model.forward(gt).double()
    log(f"adjoint scene ready in {time.time() - t0:.1f}s ({model.n} particles)")
    # one untimed rollout + backward first (the caching allocator grows to
    # hold the tapes); then the timed one
    gw = torch.tensor(g0, dtype=torch.float64, requires_grad=True)
    torch.cuda.synchronize()
    tc = time.perf_counter()
    ((rollout(gw, model).double() - v_target) ** 2).sum().backward()
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - tc) * 1e3
    gamma = torch.tensor(g0, dtype=torch.float64, requires_grad=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    v = rollout(gamma, model)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    loss = ((v.double() - v_target) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    g = gamma.grad.numpy()
    out = {"workload": "65,536 particles, 16 rings (circulations as parameters), 32 steps, checkpoint 8",
           "n_particles": model.n, "steps": steps, "checkpoint": checkpoint, "mc_samples": 1024,
           "forward_ms": round((t2 - t1) * 1e3, 2), "backward_ms": round((t3 - t2) * 1e3, 2),
           "backward_over_forward": round((t3 - t2) / (t2 - t1), 2),
           "first_forward_plus_backward_ms": round(cold_ms, 2), "loss": loss.item(),
           "grad_norm": float(np.linalg.norm(g)), "grad": [round(float(x), 6) for x in g]}
    log(f"adjoint: {json.dumps(out)}")
    return out


def _timed_steps(solver, n, before=None):
    """n steps, each timed alone (the step returns after the device finishes);
    before(k) runs untimed ahead of step k.  Returns [(ms, k_v, k_p)]."""
    import torch
    out = []
    for k in range(n):
        if before is not None:
            before(k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.step()
        out.append(((time.perf_counter() - t0) * 1e3, solver.last_iterations_v, solver.last_iterations_p))
    return out


def _slope(rows, col):
    """Least-squares ms per extra iteration of column col (1: k_v, 2: k_p)."""
    import numpy as np
    k = np.array([r[col] for r in rows], float)
    t = np.array([r[0] for r in rows], float)
    if np.ptp(k) == 0:
        return None
    a = np.vstack([k, np.ones_like(k)]).T
    return float(np.linalg.lstsq(a, t, rcond=None)[0][0])


def run_jacobi_scale(far_field="fast"):
    """Cost of one Jacobi iteration at 1M particles (the default workload
    converges both solves in one iteration).  Pressure: the 1M ring lattice
    compressed to 0.92 of its spacing (rho about 1.28 x 800 > rho0), whose
    pressure solve iterates from the 100-iteration cap down to 1 as the
    compression relaxes.  Divergence: the uncompressed lattice with a
    converging flow -a (x - c) added before each step, a swept so that the
    divergence solve iterates a varying number of times (rho stays below rho0,
    so the pressure solve converges at once).  The step time against the
    iteration count gives ms per iteration (least squares; iterations past the
    first run inside the solve's persistent launch), and the solve's
    algorithmic bytes per iteration (95 and 99 B per particle) over it an HBM
    rate."""
    import numpy as np
    import torch
    import scenes
    from config_builder import SimConfig
    from particle_system import ParticleSystem

    info = scenes.make_ring_scene(128, 128, 64, n_rings=8, seed=2, mc_samples=4096, far_field=far_field)
    out = {"workload": "1M ring lattice: pressure solve on the lattice compressed to 0.92 spacing; divergence "
                       "solve (maxErrorV 1e-3) with a converging flow -a (x - c) added before each step, a swept"}
    # pressure iterations
    ps = ParticleSystem(SimConfig(scene=info["scene"]))
    scenes.add_ring_particles(ps, (info["pos"] * np.float32(0.92)).astype(np.float32), info["vel"])
    solver = ps.build_solver()
    solver.initialize()
    n = ps.fluid_particle_num[None]
    rows_p = _timed_steps(solver, 10)
    del solver, ps
    torch.cuda.empty_cache()
    # divergence iterations (tolerance maxErrorV 1e-3 %, so that a moderate
    # converging flow takes several iterations to remove)
    ps, solver, _ = scenes.build_ring_system(128, 128, 64, n_rings=8, seed=2, mc_samples=4096, far_field=far_field)
    solver.max_error_V = 1e-3
    solver.step()
    rates = [0.0, 1.0, 3.0, 10.0, 0.0, 3.0, 10.0, 1.0, 10.0, 3.0]

    def converge(k):
        if rates[k] > 0.0:
            x = ps.x.t[: ps.particle_num[None]]
            ps.v.t[: x.shape[0]] -= rates[k] * (x - x.mean(0))

    rows_v = _timed_steps(solver, len(rates), converge)
    del solver, ps
    torch.cuda.empty_cache()
    ms_p, ms_v = _slope(rows_p, 2), _slope(rows_v, 1)
    out.update({
        "n_particles": n,
        "pressure_steps": [{"ms": round(t, 3), "k_v": kv, "k_p": kp} for t, kv, kp in rows_p],
        "divergence_steps": [{"ms": round(t, 3), "k_v": kv, "k_p": kp, "converging_rate": a}
                             for (t, kv, kp), a in zip(rows_v, rates)],
        "ms_per_pressure_iteration": None if ms_p is None else round(ms_p, 5),
        "ms_per_divergence_iteration": None if ms_v is None else round(ms_v, 5),
        "pressure_iteration_gbs": None if not ms_p else round(99 * n / (ms_p * 1e-3) / 1e9, 1),
        "divergence_iteration_gbs": None if not ms_v else round(95 * n / (ms_v * 1e-3) / 1e9, 1),
        "note": "iteration bytes: 95 B (divergence: density change + velocity update) and 99 B (pressure: "
                "density_adv + velocity update) per particle, read and written once"})
    log(f"jacobi at scale: {json.dumps(out)}")
    return out


def valu_issue_probe():
    """Measured vector-ALU issue rate and shader clock of this GPU under dense
    independent FP32 FMA streams (all-VGPR operands, then one scalar operand)."""
    import ctypes
    import torch
    import _native as N
    lib = N.lib()
    need = lib.dimcv_valu_issue_probe_bytes()
    if need < 0:
        raise RuntimeError("dimcv_valu_issue_probe_bytes: no device")
    buf = torch.empty(int(need), dtype=torch.uint8, device="cuda")
    out = (ctypes.c_double * 4)()
    N.check(lib.dimcv_valu_issue_probe(ctypes.c_void_p(buf.data_ptr()), int(need), out,
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
            "dimcv_valu_issue_probe")
    r = {"clock_ghz_vgpr_fma": round(out[0], 3), "issue_per_quad_cycle_vgpr_fma": round(out[1], 3),
         "clock_ghz_sgpr_fma": round(out[2], 3), "issue_per_quad_cycle_sgpr_fma": round(out[3], 3)}
    log(f"valu issue probe: {json.dumps(r)}")
    return r


def grid_sph_rate(res, ms_per_step):
    """Algorithmic HBM bytes of the grid and SPH passes per step (each
    particle's fields read and written once per pass):
    B = N (833 + 95 k_v + 99 k_p) + 8 N_c (12 + 2 (k_v + k_p)), over the
    step's time outside the far-field kernel."""
    n, nc, kv, kp = res["n_fluid"], res["n_cells"], res["kv"], res["kp"]
    b = n * (833 + 95 * kv + 99 * kp) + 8 * nc * (12 + 2 * (kv + kp))
    ms = ms_per_step - res["far_ms"]
    gbs = b / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "algorithmic_bytes_per_step": int(round(b)), "ms_per_step": round(ms, 4),
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "passes": "the step outside the far-field kernel: grid build and sort, neighbour lists, density, "
                      "DFSPH factor and both Jacobi solves, non-pressure forces, vorticity, grad v, d_vort, "
                      "smoothing, KVN, AKDE, CDF sampling, near field, life time, predict, advect",
            "note": "bytes counted once per particle per pass (k_v, k_p the mean iteration counts); the "
                    "neighbour passes re-read neighbour fields through L2 per pair and are vector-ALU bound "
                    "(about a third of their instructions are the bit-exact IEEE divisions and square root "
                    "of grad W), far below the HBM bound"}


def run_cpu_baseline(info, steps=1):
    """The oracle on a bounded sample of the 1M workload: one step."""
    import copy
    import numpy as np
    from oracle.refsim import RefSim
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ref = RefSim(copy.deepcopy(info["scene"]))
    n = info["pos"].shape[0]
    ref.add_particles(0, n, info["pos"], info["vel"], np.full(n, 1000.0, np.float32),
                      np.zeros(n, np.float32), np.ones(n, np.int32), np.ones(n, np.int32),
                      np.tile(np.array([[50, 100, 200]], np.int32), (n, 1)))
    ref.initialize()
    rng = np.random.default_rng(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        ref.step(u_cdf=rng.random(4096))
    sec = time.perf_counter() - t0
    log(f"cpu oracle: {steps} step(s) at {n} particles in {sec:.1f}s with {threads} threads")
    return {"value": n * steps / sec, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the 1,048,576-particle workload after initialisation "
                      f"(oracle: C++ restatement of the reference step, OpenMP, exact far field)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-jacobi", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="after the timed steps, run a 256 MiB elementwise copy (PMC scale reference)")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--sync-steps", action="store_true",
                    help="time synchronous steps (each step() waits for the device) instead of pipelined ones")
    ap.add_argument("--far-field", default="fast", choices=["fast", "kahan"],
                    help="far-field evaluation of the timed 1M step")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} (WORLD_SIZE={world}) is not "
              f"supported by this build (the multi-GPU slab decomposition is design-only)",
              file=sys.stderr)
        sys.exit(2)

    res, info = run_gpu(args)
    probe = valu_issue_probe()  # right after the timed steps, on a warm GPU
    direct = None
    if not args.no_direct:
        direct = {"comparator_131072": run_recovery_comparison(64, 64, 32, [256, 1024, 4096, 16384, 131072]),
                  "precision_262144": run_recovery_comparison(64, 64, 64, [1024, 4096, 16384, 65536, 262144])}
    batching = None if args.no_batching else run_batching(steps=args.batch_steps)
    adjoint = None if args.no_adjoint else run_adjoint()
    jacobi = None if args.no_jacobi else run_jacobi_scale(args.far_field)
    cpu = None if args.no_cpu_baseline else run_cpu_baseline(info, steps=5)
    value = res["n_fluid"] * args.steps / res["sec"]
    traffic = None
    prof = os.path.join(REPO, "profiles", "mc_far_field_1M_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["sec"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded jittered lattice + Lamb-Oseen vortex rings",
        "config": {"workload": "1M-particle vortex-ring lattice, MC-VVPE M=4096",
                   "n_particles": res["n_fluid"], "lattice": [128, 128, 64], "rings": 8,
                   "mc_samples": res["m"], "dt": 2e-4, "far_field": args.far_field,
                   "mean_k_v": res["kv"], "mean_k_p": res["kp"], "parallelism": "single GPU",
                   "steps_pipelined": not args.sync_steps,
                   "ms_per_step_synchronous_next_20": round(res["sync_ms"], 4),
                   "ms_per_step_pipelined_20_after_those": None if res["async_after_ms"] is None
                   else round(res["async_after_ms"], 4)},
        "roofline": {"bound": "valu", "kernel": "MC-VVPE far field: k_far_split over two sample halves + k_far_combine (k_vv_far_fast when unsplit)",
                     "achieved": round(res["achieved_tflops"], 3), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(res["achieved_tflops"] / FP32_PEAK_TFLOPS, 4),
                     "traffic": traffic, "avg_launch_ms": round(res["far_ms"], 5),
                     "work_per_launch": f"{res['pairs']} (target, sample) pairs x {FLOPS_PER_PAIR_MC} flop",
                     "issue_ceiling": FP32_ISSUE_TFLOPS,
                     "frac_of_issue_ceiling": round(res["achieved_tflops"] / FP32_ISSUE_TFLOPS, 4),
                     "measured_issue": probe,
                     "note": "FP32 vector-ALU bound (no matrix-core work). peak = MI355X FP32 vector 157.3 TFLOP/s. Built without packed "
                             "FP32 (-packed-fp32-ops); the loop issues 59 vector instructions per four samples (14.75 per pair) and "
                             "the SIMDs dual-issue part of them (SQ_ACTIVE_INST_VALU2), so the issue ceiling is 2 "
                             "instructions per SIMD quad-cycle: issue_ceiling TFLOP/s at 24 flop per pair and 2.4 GHz. "
                             "measured_issue: the clock and issue rate measured on this GPU right after the timed steps "
                             "by dense independent FMA streams (all-vector operands pair up to ~1.9 per quad-cycle at "
                             "~2.0 GHz; a scalar operand in every instruction stops pairing). The far field itself runs "
                             "at 2.3-2.4 GHz (its own wave timed with s_memtime in a separate build; GRBM_GUI_ACTIVE "
                             "over its duration agrees) and issues ~1.4 instructions per SIMD quad-cycle there: 28 of "
                             "its 59 instructions per four samples read a scalar register and 5 issue alone. A "
                             "matrix-core form (S = W Q on v_mfma_f32_16x16x4_f32) measured slower: no VALU/MFMA "
                             "co-execution (SQ_VALU_MFMA_COEXEC_CYCLES = 0) and 7 of 16 columns used"},
        "grid_sph_passes": grid_sph_rate(res, res["sec"] * 1e3 / args.steps),
        "cpu_baseline": cpu,
        "mc_vs_direct": direct,
        "batching": batching,
        "adjoint": adjoint,
        "jacobi_at_scale": jacobi,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
