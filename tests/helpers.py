"""Shared helpers for the test-suite (scene loading, image metrics)."""
import os

import numpy as np

from forgedthoughts_amd.forge.evaluator import load_scene_source
from forgedthoughts_amd.forge.imports import load_scene_file
from forgedthoughts_amd.options import RaySettings, RenderOptions
from forgedthoughts_amd.scene import compile_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REGRESSION = os.path.join(ROOT, "tests", "golden", "render_regression")

# The reference regression cases (crates/forgedthoughts/tests/render_regression.rs:34-107).
REGRESSION_CASES = [
    ("preview_lambert", "preview", RenderOptions(96, 96, 320, 40.0, 0.0002, 0.7, 6.0, 35.0), None),
    ("ray_metal", "ray", RenderOptions(96, 96, 360, 50.0, 0.0002, 0.7, 6.0, 35.0), RaySettings(4, 96, 1)),
    ("ray_boolean_union_round", "ray", RenderOptions(96, 96, 520, 60.0, 0.0002, 0.7, 6.0, 25.0), RaySettings(4, 96, 1)),
    ("preview_material_hook", "preview", RenderOptions(96, 96, 320, 40.0, 0.0002, 0.7, 6.0, 35.0), None),
]


def load_ft(path, library_dirs=None):
    """A scene file with its imports (file imports relative to it, library imports from
    `library_dirs` or FT_LIBRARY_PATH)."""
    return load_scene_file(path, library_dirs)


def compile_ft(path):
    return compile_scene(load_ft(path))


def compile_src(src):
    return compile_scene(load_scene_source(src))


def read_png(path):
    from PIL import Image
    return np.asarray(Image.open(path).convert("RGB"))


def diff_metrics(a, b):
    """max |d| and mean |d| over all channels (render_regression.rs:196-219)."""
    d = np.abs(a.astype(np.int32) - b.astype(np.int32))
    return int(d.max()), float(d.mean())


def with_camera(scene, origin, target, fov):
    """The scene seen through an explicit camera (what a batch view means)."""
    import copy
    import numpy as np
    c = copy.copy(scene)
    c.camera = {"present": True, "origin": np.asarray(origin, np.float32),
                "target": np.asarray(target, np.float32), "fov": np.float32(fov)}
    c._ir = None
    return c
