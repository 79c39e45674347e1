"""RenderOptions / RaySettings / AccelMode (renderer.rs:31-63, 128-155)."""
from __future__ import annotations

import enum
from dataclasses import dataclass
from typing import Optional

from . import abi


class AccelMode(enum.IntEnum):
    """AccelMode (renderer.rs:58-63)."""
    Naive = abi.ACCEL_NAIVE
    Bvh = abi.ACCEL_BVH
    Bricks = abi.ACCEL_BRICKS

    @classmethod
    def parse(cls, name: str) -> "AccelMode":
        return {"naive": cls.Naive, "bvh": cls.Bvh, "bricks": cls.Bricks}[name.lower()]


@dataclass
class RenderOptions:
    """RenderOptions with the reference defaults (renderer.rs:31-56)."""
    width: int = 800
    height: int = 800
    max_steps: int = 128
    max_dist: float = 40.0
    epsilon: float = 0.001
    step_scale: float = 0.7
    camera_z: float = 6.0
    fov_y_degrees: float = 45.0

    def to_c(self) -> abi.RenderOptionsC:
        return abi.RenderOptionsC(int(self.width), int(self.height), int(self.max_steps),
                                  float(self.max_dist), float(self.epsilon), float(self.step_scale),
                                  float(self.camera_z), float(self.fov_y_degrees))


@dataclass
class RaySettings:
    """RaySettings with the reference defaults (renderer.rs:128-144)."""
    max_depth: int = 8
    tile_size: int = 64
    aa_samples: int = 1
    debug_aov: Optional[str] = None

    def to_c(self) -> abi.RaySettingsC:
        return abi.RaySettingsC(int(self.max_depth), int(self.tile_size), int(self.aa_samples),
                                abi.AOV[self.debug_aov])
