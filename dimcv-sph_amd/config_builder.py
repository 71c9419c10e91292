"""Scene configuration (drop-in for config_builder.py:3-42 of the reference).

``SimConfig(path)`` loads a scene JSON; ``get_cfg(name)`` returns the value of
``Configuration[name]`` or ``None`` when absent (asserting instead when
``enforce_exist`` is set); the block getters return ``[]`` when a block is
missing.  A dict may be passed instead of a path.

Keys this build adds (all optional; absent keys keep reference behaviour):
  ``mcSamples``  fixed importance-sample count M (inverse-CDF sampling);
  ``farField``   "exact" | "fast" | "kahan" far-field evaluation;
  ``viscosity``  kinematic viscosity (the reference hard-codes 1e-3);
  ``rngSeed``    seed of the host uniform generator (replaces ti.random);
  ``neighborListWidth``  neighbour-list capacity per particle;
  ``periodic``   periodic box (minimum image, positions wrapped; the box must
                 start at the origin and span at least 3 cells per axis).
"""
from __future__ import annotations

import json


class SimConfig:
    def __init__(self, scene_file_path=None, scene: dict | None = None, verbose: bool = False) -> None:
        if scene is not None:
            self.config = scene
        elif isinstance(scene_file_path, dict):
            self.config = scene_file_path
        else:
            with open(scene_file_path, "r") as f:
                self.config = json.load(f)
        if verbose:
            print(self.config)

    def get_cfg(self, name, enforce_exist=False):
        if enforce_exist:
            assert name in self.config["Configuration"]
        if name not in self.config["Configuration"]:
            return None
        return self.config["Configuration"][name]

    def get_rigid_bodies(self):
        return self.config.get("RigidBodies", [])

    def get_rigid_blocks(self):
        return self.config.get("RigidBlocks", [])

    def get_fluid_blocks(self):
        return self.config.get("FluidBlocks", [])

    def get_fluid_emitters(self):
        return self.config.get("FluidEmitters", [])
