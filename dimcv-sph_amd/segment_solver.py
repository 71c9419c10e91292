"""Segment solver: its O(N^2) Biot-Savart and RK4 advection on the GPU.

Drop-in for the velocity part of the reference's segment solver
(segment_config.py, segment_system.py, segment_solver.py): ``SegmentConfig``,
``SegmentSystem`` (the segment pool, segment_system.py:7-49) and
``SegmentSolver`` with ``compute_endpoint_velocity`` (:362-418),
``_compute_velocity_at_factor`` (:468-521), ``advect_segments_rk4``
(:420-447, 458-542) and the ring seeding of ``seed_initial_segments``
(:45-128).  The sums run in libdimcv (``dimcv_segment_velocity``,
``dimcv_segment_rk4``): one lane per target segment, sources walked in
ascending order with wave-uniform loads, bitwise the serial reference
arithmetic.

The segment solver's topology operations (split, merge, delete, cull), its
boundary handler and the other seeding types stay with the reference; they
raise NotImplementedError here.
"""
from __future__ import annotations

import ctypes
import json

import numpy as np
import torch

import _native as N
from _field import Field


class SegmentConfig:
    """segment_config.py:4-23 (a dict may be passed instead of a path)."""

    def __init__(self, scene_file_path):
        if isinstance(scene_file_path, dict):
            self.config = scene_file_path
        else:
            with open(scene_file_path, "r", encoding="utf-8") as f:
                self.config = json.load(f)

    def get_cfg(self, name: str, default=None):
        return self.config.get("SegmentConfiguration", {}).get(name, default)

    def get_domain_start(self):
        return self.get_cfg("domainStart", [0.0, 0.0, 0.0])

    def get_domain_end(self):
        return self.get_cfg("domainEnd", [1.0, 1.0, 1.0])

    def get_obstacles(self):
        return self.config.get("RigidBlocks", [])

    def get_inflow(self):
        return self.config.get("FluidEmitters", [])


class SegmentSystem:
    """The segment pool (segment_system.py:7-49), device resident."""

    def __init__(self, config: SegmentConfig, device: str = "cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("SegmentSystem needs a ROCm GPU (libdimcv has no CPU path)")
        self.cfg = config
        self.dim = 3
        self.device = torch.device(device)
        self.domain_start = np.array(self.cfg.get_domain_start(), dtype=np.float32)
        self.domain_end = np.array(self.cfg.get_domain_end(), dtype=np.float32)
        self.segment_max_num = int(self.cfg.get_cfg("segmentMaxNum", 200000))
        n, dev = self.segment_max_num, self.device
        self.segment_num = _Count()
        self.x_minus = Field(n, 3, torch.float32, dev)
        self.x_plus = Field(n, 3, torch.float32, dev)
        self.gamma = Field(n, 0, torch.float32, dev)
        self.active = Field(n, 0, torch.int32, dev)
        self.age = Field(n, 0, torch.float32, dev)
        self.seg_type = Field(n, 0, torch.int32, dev)
        self.center = Field(n, 3, torch.float32, dev)
        self.tangent = Field(n, 3, torch.float32, dev)
        self.length = Field(n, 0, torch.float32, dev)

    def clear(self):
        self.segment_num[None] = 0
        for f in (self.active, self.age, self.gamma, self.length, self.seg_type):
            f.fill(0)

    def update_segment_geometry(self):
        """segment_system.py:41-49: centre, unit tangent and length of each
        active segment (elementwise, f32)."""
        n = int(self.segment_num[None])
        if n == 0:
            return
        act = self.active.t[:n] == 1
        xm, xp = self.x_minus.t[:n], self.x_plus.t[:n]
        d = xp - xm
        L = torch.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) + np.float32(1e-8)
        self.center.t[:n] = torch.where(act[:, None], np.float32(0.5) * (xp + xm), self.center.t[:n])
        self.tangent.t[:n] = torch.where(act[:, None], d / L[:, None], self.tangent.t[:n])
        self.length.t[:n] = torch.where(act, L, self.length.t[:n])


class SegmentSolver:
    """SegmentSolver (segment_solver.py:8-39): constants and endpoint-velocity
    fields; the Biot-Savart sums and RK4 stages run in libdimcv."""

    def __init__(self, segment_system: SegmentSystem):
        self.ss = segment_system
        cfg = self.ss.cfg
        self.dt = float(cfg.get_cfg("timeStepSize", 0.002))
        self.reg_radius = float(cfg.get_cfg("regularizationRadiusR", 0.01))
        self.gamma_decay = float(cfg.get_cfg("gammaDecay", 1.0))
        self.split_len_threshold = float(cfg.get_cfg("splitLengthThreshold", 0.08))
        self.delete_gamma_threshold = float(cfg.get_cfg("deleteGammaThreshold", 1e-4))
        self.background_velocity = np.array(cfg.get_cfg("backgroundVelocity", [0.0, 0.0, 0.0]), dtype=np.float32)
        if cfg.get_cfg("enableBoundaryInjection", False):
            raise NotImplementedError("the segment boundary handler stays with the reference")
        n, dev = self.ss.segment_max_num, self.ss.device
        self.v_minus = Field(n, 3, torch.float32, dev)
        self.v_plus = Field(n, 3, torch.float32, dev)
        self._k = torch.zeros(24 * n, dtype=torch.float32, device=dev)
        self._sources = torch.zeros(8 * n, dtype=torch.float32, device=dev)
        self.k1_minus, self.k2_minus, self.k3_minus, self.k4_minus = (
            _View(self._k, 3 * n * s, n) for s in range(4))
        self.k1_plus, self.k2_plus, self.k3_plus, self.k4_plus = (
            _View(self._k, 3 * n * s, n) for s in range(4, 8))

    # ------------------------------------------------------------ ABI --
    def _seg(self) -> N.Segments:
        ss = self.ss
        s = N.Segments()
        s.x_minus, s.x_plus = ss.x_minus.t.data_ptr(), ss.x_plus.t.data_ptr()
        s.gamma, s.active, s.age = ss.gamma.t.data_ptr(), ss.active.t.data_ptr(), ss.age.t.data_ptr()
        s.sources = self._sources.data_ptr()
        s.n = int(ss.segment_num[None])
        s.reg_radius = np.float32(self.reg_radius)
        s.dt = np.float32(self.dt)
        s.dt6 = np.float32(self.dt / 6.0)
        s.inv4pi = np.float32(1.0 / (4.0 * np.pi))
        s.gamma_decay = np.float32(self.gamma_decay)
        for d in range(3):
            s.u_inf[d] = np.float32(self.background_velocity[d])
        return s

    def _stream(self):
        return torch._C._cuda_getCurrentRawStream(self.ss.device.index or 0)

    # ------------------------------------------------------- reference API --
    def initialize(self):
        self.ss.clear()
        self.seed_initial_segments()
        self.ss.update_segment_geometry()

    def seed_initial_segments(self):
        """segment_solver.py:45-128, ring (and none) only."""
        cfg = self.ss.cfg
        init_type = str(cfg.get_cfg("initType", "ring") or "ring").lower()
        if init_type == "none":
            return
        if init_type != "ring":
            raise NotImplementedError(f"initType={init_type} stays with the reference (ring and none here)")
        n_seg = max(3, int(cfg.get_cfg("initSegmentNum", 64)))
        ring_radius = float(cfg.get_cfg("initRingRadius", 0.15))
        gamma0 = float(cfg.get_cfg("initGamma", 0.25))
        seg_type = int(cfg.get_cfg("initSegmentTypeId", 0))
        center_cfg = cfg.get_cfg("initCenter", None)
        c = 0.5 * (self.ss.domain_start + self.ss.domain_end) if center_cfg is None \
            else np.array(center_cfg, dtype=np.float32)
        axis_cfg = cfg.get_cfg("initRingAxis", None)
        if axis_cfg is None:
            axis = self.background_velocity.copy()
            if float(np.linalg.norm(axis)) < 1e-6:
                axis = np.array([1.0, 0.0, 0.0], dtype=np.float32)
        else:
            axis = np.array(axis_cfg, dtype=np.float32)
        axis = axis / (np.linalg.norm(axis) + 1e-8)
        ref = np.array([1.0, 0.0, 0.0], dtype=np.float32)
        if abs(float(np.dot(axis, ref))) > 0.9:
            ref = np.array([0.0, 1.0, 0.0], dtype=np.float32)
        e1 = np.cross(axis, ref)
        e1 = e1 / (np.linalg.norm(e1) + 1e-8)
        e2 = np.cross(axis, e1)
        e2 = e2 / (np.linalg.norm(e2) + 1e-8)
        angles = np.linspace(0.0, 2.0 * np.pi, n_seg + 1, dtype=np.float32)[:-1]
        pts = (c[None, :] + ring_radius * (np.cos(angles)[:, None] * e1[None, :] +
                                           np.sin(angles)[:, None] * e2[None, :])).astype(np.float32)
        self.add_segments(pts, np.roll(pts, shift=-1, axis=0), gamma0 * np.ones((n_seg,), dtype=np.float32),
                          seg_type)

    def add_segments(self, x_minus, x_plus, gamma, seg_type: int = 0) -> int:
        """_seed_segments_kernel (segment_solver.py:343-360) after the pool's
        current segments; returns the number added (capacity-limited)."""
        offset = int(self.ss.segment_num[None])
        n_new = min(len(gamma), self.ss.segment_max_num - offset)
        if n_new <= 0:
            return 0
        dev = self.ss.device
        sl = slice(offset, offset + n_new)
        self.ss.x_minus.t[sl] = torch.as_tensor(np.asarray(x_minus, np.float32)[:n_new], device=dev)
        self.ss.x_plus.t[sl] = torch.as_tensor(np.asarray(x_plus, np.float32)[:n_new], device=dev)
        self.ss.gamma.t[sl] = torch.as_tensor(np.asarray(gamma, np.float32)[:n_new], device=dev)
        self.ss.active.t[sl] = 1
        self.ss.age.t[sl] = 0.0
        self.ss.seg_type.t[sl] = int(seg_type)
        self.ss.segment_num[None] = offset + n_new
        return n_new

    def compute_endpoint_velocity(self):
        s = self._seg()
        N.check(N.lib().dimcv_segment_velocity(ctypes.byref(s), 0, 0.0, None, None, self.v_minus.t.data_ptr(),
                                               self.v_plus.t.data_ptr(), self._stream()), "dimcv_segment_velocity")

    def _compute_velocity_at_factor(self, factor, k_minus_in, k_plus_in, out_minus, out_plus):
        s = self._seg()
        N.check(N.lib().dimcv_segment_velocity(ctypes.byref(s), 1, float(factor), _ptr(k_minus_in),
                                               _ptr(k_plus_in), _ptr(out_minus), _ptr(out_plus), self._stream()),
                "dimcv_segment_velocity")

    def advect_segments_rk4(self):
        s = self._seg()
        N.check(N.lib().dimcv_segment_rk4(ctypes.byref(s), self.v_minus.t.data_ptr(), self.v_plus.t.data_ptr(),
                                          self._k.data_ptr(), self._stream()), "dimcv_segment_rk4")

    def delete_weak_segments(self):
        raise NotImplementedError("segment topology operations stay with the reference")

    split_segments = merge_segments = cull_segments = delete_weak_segments

    def step(self):
        raise NotImplementedError("the segment step's topology operations stay with the reference; "
                                  "compute_endpoint_velocity and advect_segments_rk4 run here")


class _Count:
    """segment_num: a 0-d int field held on the host ([None] access)."""

    shape = ()

    def __init__(self):
        self._v = 0

    def __getitem__(self, idx):
        assert idx is None
        return self._v

    def __setitem__(self, idx, value):
        assert idx is None
        self._v = int(value)


class _View:
    """One RK4 stage array (n x 3) inside the stage scratch."""

    def __init__(self, buf: torch.Tensor, offset: int, n: int):
        self.t = buf[offset:offset + 3 * n].view(n, 3)

    def to_numpy(self):
        return self.t.cpu().numpy()


def _ptr(f) -> int:
    t = f.t if hasattr(f, "t") else f
    return t.data_ptr()
