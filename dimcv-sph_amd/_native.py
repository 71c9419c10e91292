"""ctypes binding of libdimcv.so (the gfx950 HIP library, C ABI in include/dimcv.h).

The product has exactly one compute path: this library.  If the shared object
is missing or fails to load, :func:`lib` raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libdimcv.so")
SOURCES = ["csrc/util.hip", "csrc/grid.hip", "csrc/sph.hip", "csrc/vvpe.hip", "csrc/step.hip",
           "csrc/batch.hip", "csrc/adjoint.hip", "csrc/export.hip",
           "csrc/segment.hip", "csrc/probe.hip"]
HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17",
    # one IEEE operation per source operation: the SPH passes reproduce the
    # serial reference arithmetic bit for bit
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

_LIB = None

# dimcv_counter / dimcv_dscalar / dimcv_far_mode (include/dimcv.h)
C_PARTICLE_NUM, C_FLUID_NUM, C_SOLID_NUM, C_NUM_SAMPLES = 0, 1, 2, 3
C_FLUID_LIST, C_NBR_MAX, C_KEY_ERROR, C_PREV_NUM, C_SAMPLE_SLOTS = 4, 5, 6, 7, 8
C_JACOBI_DONE, C_JACOBI_ITERS = 9, 10
C_COUNT = 16
D_KDEN_SUM, D_RESIDUAL, D_CDF_TOTAL, D_COUNT = 0, 1, 2, 8
FAR_EXACT, FAR_FAST, FAR_KAHAN = 0, 1, 2
# dimcv_scene_counter / dimcv_scene_dscalar
SC_DONE, SC_ITERS, SC_NUM_SAMPLES, SC_FLUID, SC_FLUID_START, SC_SLOT_START, SC_SLOTS = 0, 1, 2, 3, 4, 5, 6
SC_TILE_START, SC_SPLITS, SC_PART_BASE, SC_ITEM_START, SC_ITERS_V, SC_ITERS_P, SC_FLUID_SRC = 7, 8, 9, 10, 11, 12, 13
SC_COUNT = 16
SD_KDEN_SUM, SD_RESIDUAL, SD_CDF_TOTAL, SD_COUNT = 0, 1, 2, 4


def build(verbose: bool = False) -> str:
    """Compile every HIP source into libdimcv.so for gfx950 (in-tree)."""
    tmp = LIB_PATH + ".tmp"
    cmd = ["hipcc", *HIPCC_FLAGS, "-o", tmp, *[os.path.join(HERE, s) for s in SOURCES]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


F3 = ctypes.c_float * 3
I3 = ctypes.c_int32 * 3


class Params(ctypes.Structure):
    """dimcv_params (include/dimcv.h)."""
    _fields_ = [
        ("grid_num", I3), ("n_cells", ctypes.c_int32), ("n_max", ctypes.c_int64),
        ("nbr_kmax", ctypes.c_int32), ("periodic", ctypes.c_int32),
        ("grid_size", ctypes.c_float), ("support_radius", ctypes.c_float),
        ("inv_support_radius", ctypes.c_float), ("kernel_k", ctypes.c_float),
        ("kernel_dk", ctypes.c_float), ("density0", ctypes.c_float), ("dt", ctypes.c_float),
        ("gravity", F3), ("visc_coef", ctypes.c_float), ("visc_h2", ctypes.c_float),
        ("m_eps", ctypes.c_float), ("domain_start", F3), ("domain_end", F3), ("box_size", F3),
        ("has_vortex_domain", ctypes.c_int32), ("vortex_start", F3), ("vortex_end", F3),
        ("has_transit", ctypes.c_int32), ("transit_a", ctypes.c_float), ("transit_ba", ctypes.c_float),
        ("vp_factor", ctypes.c_float), ("eps", ctypes.c_float), ("inv_eps", ctypes.c_float),
        ("neg_gen_a", ctypes.c_float), ("neg_del_b", ctypes.c_float),
        ("far_r2_max_rejected", ctypes.c_float), ("bounce_padding", ctypes.c_float),
        ("bounce_hi", F3), ("far_mode", ctypes.c_int32), ("m_V0", ctypes.c_float),
        ("nbr_r2_limit", ctypes.c_float),
        ("vp_factor_f64", ctypes.c_double),
        ("dyn_rigid", ctypes.c_int32), ("inv_dt", ctypes.c_float),
    ]


STATE_POINTERS = [
    "object_id", "x", "x_0", "v", "acceleration", "m_V", "m", "density", "pressure", "material",
    "color", "is_dynamic", "vorticity", "vorticity_vis", "dfsph_factor", "density_adv", "is_active",
    "is_sample", "life_time", "grid_ids",
    "object_id_buffer", "x_buffer", "x_0_buffer", "v_buffer", "m_V_buffer", "m_buffer",
    "density_buffer", "pressure_buffer", "material_buffer", "color_buffer", "is_dynamic_buffer",
    "dfsph_factor_buffer", "density_adv_buffer", "is_active_buffer", "is_sample_buffer",
    "life_time_buffer", "grid_ids_buffer",
    "grid_particles_num", "grid_particles_num_temp",
    "d_vort", "d_vort_smoothed", "v_star", "grad_v", "x_temp", "sample_idx", "kinematic_vort_num",
    "kernel_density", "num_particles_each_emitter",
]
STATE_POINTERS_2 = ["fluid_list", "nbr_count", "nbr", "v_vort", "counters", "dscal", "workspace"]

# Fields the grid build permutes into their *_buffer twin (dimcv_state [sorted]).
SORTED_FIELDS = ["object_id", "x", "x_0", "v", "m_V", "m", "density", "pressure", "material", "color",
                 "is_dynamic", "dfsph_factor", "density_adv", "is_active", "is_sample", "life_time",
                 "grid_ids", "particle_id"]


class StepOpts(ctypes.Structure):
    """dimcv_step_opts (include/dimcv.h)."""
    _fields_ = [
        ("u1", ctypes.c_void_p), ("u2", ctypes.c_void_p), ("u_cdf", ctypes.c_void_p),
        ("m", ctypes.c_int32), ("max_iter_v", ctypes.c_int32), ("max_iter_p", ctypes.c_int32),
        ("chunk", ctypes.c_int32), ("eta_v", ctypes.c_double), ("eta_p", ctypes.c_double),
        ("inv_dt", ctypes.c_float), ("dt", ctypes.c_float), ("inv_dt2", ctypes.c_float),
        ("density0", ctypes.c_float), ("fluid_count", ctypes.c_int32), ("bounce_back", ctypes.c_int32),
        ("out_kv", ctypes.c_int32), ("out_kp", ctypes.c_int32),
        ("time_far", ctypes.c_int32), ("out_far_ms", ctypes.c_float),
    ]


# dimcv_adj_pass_id
(ADJ_DENSITY, ADJ_FACTOR, ADJ_DIV, ADJ_JACOBI, ADJ_NONPRESSURE, ADJ_CURL, ADJ_GRAD_V, ADJ_SMOOTH, ADJ_KVN,
 ADJ_KDEN, ADJ_GATE, ADJ_NEAR, ADJ_FAR_TARGETS, ADJ_FAR_SOURCES) = range(14)


class AdjArgs(ctypes.Structure):
    """dimcv_adj_args (include/dimcv.h)."""
    _fields_ = [("f0", ctypes.c_void_p), ("f1", ctypes.c_void_p), ("f2", ctypes.c_void_p),
                ("g_in", ctypes.c_void_p), ("x_bar", ctypes.c_void_p), ("b0", ctypes.c_void_p),
                ("b1", ctypes.c_void_p), ("b2", ctypes.c_void_p), ("dscalar", ctypes.c_double),
                ("scalar", ctypes.c_float), ("mode", ctypes.c_int32)]


class Emitter(ctypes.Structure):
    """dimcv_emitter (include/dimcv.h)."""
    _fields_ = [("pos", ctypes.c_void_p), ("n", ctypes.c_int32), ("object_id", ctypes.c_int32),
                ("velocity", F3), ("density", ctypes.c_float), ("color", I3)]


class Segments(ctypes.Structure):
    """dimcv_segments (include/dimcv.h)."""
    _fields_ = [("x_minus", ctypes.c_void_p), ("x_plus", ctypes.c_void_p), ("gamma", ctypes.c_void_p),
                ("active", ctypes.c_void_p), ("age", ctypes.c_void_p), ("sources", ctypes.c_void_p),
                ("n", ctypes.c_int32), ("reg_radius", ctypes.c_float), ("dt", ctypes.c_float),
                ("dt6", ctypes.c_float), ("inv4pi", ctypes.c_float), ("gamma_decay", ctypes.c_float),
                ("u_inf", F3)]


class Scene(ctypes.Structure):
    """dimcv_scene (include/dimcv.h)."""
    _fields_ = [("grid_num", I3), ("cell_base", ctypes.c_int32), ("n_cells", ctypes.c_int32),
                ("m", ctypes.c_int32), ("sample_base", ctypes.c_int32), ("pad0", ctypes.c_int32),
                ("domain_start", F3), ("domain_end", F3), ("visc_coef", ctypes.c_float),
                ("pad1", ctypes.c_float)]


class State(ctypes.Structure):
    """dimcv_state (include/dimcv.h)."""
    _fields_ = ([(n, ctypes.c_void_p) for n in STATE_POINTERS]
                + [("n_emitters", ctypes.c_int32), ("pad0", ctypes.c_int32)]
                + [(n, ctypes.c_void_p) for n in STATE_POINTERS_2]
                + [("workspace_bytes", ctypes.c_int64)]
                + [("scene_id", ctypes.c_void_p), ("scene_id_buffer", ctypes.c_void_p),
                   ("scenes", ctypes.c_void_p), ("n_scenes", ctypes.c_int32), ("pad1", ctypes.c_int32),
                   ("scene_counters", ctypes.c_void_p), ("scene_dscal", ctypes.c_void_p),
                   ("block_scene", ctypes.c_void_p), ("batch_workspace", ctypes.c_void_p),
                   ("batch_workspace_bytes", ctypes.c_int64),
                   ("particle_id", ctypes.c_void_p), ("particle_id_buffer", ctypes.c_void_p)])


def lib():
    """Load libdimcv.so; raises if it is absent (no fallback path exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(hipcc --offload-arch=gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    P, I32, I64, F32, F64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double
    sig = {
        "dimcv_abi_version": (I32, []),
        "dimcv_last_error": (ctypes.c_char_p, []),
        "dimcv_workspace_bytes": (I64, [I64, I64, I64]),
        "dimcv_valu_issue_probe_bytes": (I64, []),
        "dimcv_valu_issue_probe": (I32, [P, I64, P, P]),
        "dimcv_add_particles": (I32, [P, P, I32, I32, P, P, P, P, P, P, P, P, P]),
        "dimcv_init_cylinder": (I32, [P, P, I32, F32, F32, F32, P]),
        "dimcv_emit": (I32, [P, P, P, I32, P]),
        "dimcv_segment_velocity": (I32, [P, I32, F32, P, P, P, P, P]),
        "dimcv_segment_rk4": (I32, [P, P, P, P, P]),
        "dimcv_export_ply_size": (I32, [P, P, I32, P, I32, P, P, P]),
        "dimcv_format_f7": (I32, [F32, ctypes.c_char_p]),
        "dimcv_export_ply_rows": (I32, [P, P, I32, P, I32, P, I64, P]),
        "dimcv_grid_build": (I32, [P, P, I32, P, P]),
        "dimcv_neighbor_build": (I32, [P, P, I32, P]),
        "dimcv_compute_static_boundary_volume": (I32, [P, P, I32, P]),
        "dimcv_compute_moving_boundary_volume": (I32, [P, P, I32, P]),
        "dimcv_rigid_moments": (I32, [P, P, I32, I32, P, P, I32, P, P]),
        "dimcv_rigid_apply": (I32, [P, P, I32, P, I32, P]),
        "dimcv_multiply_time_step": (I32, [P, P, P, F32, I32, P]),
        "dimcv_compute_density_error": (I32, [P, P, F32, I32, P, P]),
        "dimcv_compute_non_pressure_forces": (I32, [P, P, I32, I32, P]),
        "dimcv_compute_kvn_sample": (I32, [P, P, P, P, I32, I32, P]),
        "dimcv_cdf_sample": (I32, [P, P, P, I32, I32, P]),
        "dimcv_vort2vel": (I32, [P, P, I32, I32, P]),
        "dimcv_vort2vel_near": (I32, [P, P, I32, P]),
        "dimcv_vort2vel_far": (I32, [P, P, I32, I32, P]),
        "dimcv_far_split_count": (I32, [I64]),
        "dimcv_cdf_sample_noalloc": (I32, [P, P, P, I32, I32, P]),
        "dimcv_direct_bs": (I32, [P, P, I32, P, I32, P]),
        "dimcv_ring_velocity": (I32, [I64, P, F64, I64, P, P, P]),
        "dimcv_substep": (I32, [P, P, P, I32, I32, P]),
        "dimcv_step": (I32, [P, P, P, P, I32, P, P]),
        "dimcv_step_begin": (I32, [P, P, P, P, I32, P]),
        "dimcv_step_finish": (I32, [P, P, P, P, I32, P, P]),
        "dimcv_step_finish_async": (I32, [P, P, P, P, I32, P]),
        "dimcv_step_wait": (I32, [P, P, P]),
        "dimcv_batch_workspace_bytes": (I64, [I32, P, I64]),
        "dimcv_batch_grid_build": (I32, [P, P, P]),
        "dimcv_batch_neighbor_build": (I32, [P, P, P]),
        "dimcv_batch_status": (I32, [P, P, P, P]),
        "dimcv_batch_substep": (I32, [P, P, P, P, P]),
        "dimcv_batch_set_launch": (I32, [I32, I32]),
        "dimcv_batch_graph_create": (I32, [P, P, P, P, P, P]),
        "dimcv_batch_graph_step": (I32, [P, I32, P, P]),
        "dimcv_batch_graph_destroy": (None, [P]),
        "dimcv_adj_pass": (I32, [P, P, I32, P, I32, P]),
    }
    for name in ["dimcv_compute_densities", "dimcv_compute_DFSPH_factor",
                 "dimcv_compute_density_change", "dimcv_divergence_solver_iteration",
                 "dimcv_predict_velocity", "dimcv_advect", "dimcv_compute_density_adv",
                 "dimcv_pressure_solve_iteration", "dimcv_copy_x_temp",
                 "dimcv_compute_vorticity_vis", "dimcv_enforce_boundary_3d", "dimcv_compute_v_star",
                 "dimcv_compute_vorticity", "dimcv_compute_grad_v", "dimcv_dimcv_init",
                 "dimcv_compute_d_vorticity", "dimcv_compute_all_d_vort_smoothed",
                 "dimcv_record_sample_idx", "dimcv_compute_kernel_density",
                 "dimcv_update_particle_life_time"]:
        sig[name] = (I32, [P, P, I32, P])
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.dimcv_abi_version() != 1:
        raise RuntimeError("libdimcv ABI version mismatch")
    _LIB = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().dimcv_last_error().decode(errors="replace")
        raise RuntimeError(f"{what or 'libdimcv'} failed ({rc}): {msg}")
