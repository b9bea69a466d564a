"""MI355X-native per-step SPH particle pipeline (WCSPH with adaptive
split/merge) behind the reference's simulation interface.

The compute path is the HIP library adaptivesph_amd/lib/libasph.so (C ABI in
include/asph*.h); this package holds its Python bindings, the scene
generators, the batched / served front ends and the launch autotuner.
"""
from . import abi  # noqa: F401
from .scene import Scene, cfg1, cfg2, cfg3, cfg4_scenes, dam_break, random_scene, sloshing_tank  # noqa: F401

__all__ = ["abi", "Scene", "cfg1", "cfg2", "cfg3", "cfg4_scenes", "dam_break", "random_scene", "sloshing_tank"]
