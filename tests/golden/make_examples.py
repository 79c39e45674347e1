"""Write the scene IR of the reference's example scenes as test fixtures.

Usage: python tests/golden/make_examples.py [EXAMPLES_DIR] [LIBRARY_DIR]
(defaults: the examples/ folder and the built-in object library of a ForgeThoughts
checkout at /root/reference).  Each example that this front end compiles becomes
tests/golden/examples/<name>.npz: the ft_scene_ir arrays (CompiledScene.save) plus
`options` = [max_steps, max_dist, epsilon, step_scale, camera_z, fov_y] from the scene's
RenderSettings (RenderOptions defaults where it sets none) and `accel`.  The fixtures hold
compiled data only, so the GPU parity tests can render the examples where no checkout of
the reference exists.
"""
import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from forgedthoughts_amd.forge.imports import load_scene_file  # noqa: E402
from forgedthoughts_amd.options import RenderOptions  # noqa: E402
from forgedthoughts_amd.scene import compile_scene, extract_scene_render_settings  # noqa: E402

OUT = os.path.join(HERE, "examples")


def main():
    examples = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples"
    library = sys.argv[2] if len(sys.argv) > 2 else "/root/reference/crates/forgedthoughts/library"
    os.makedirs(OUT, exist_ok=True)
    d = RenderOptions()
    written, skipped = [], []
    for path in sorted(glob.glob(os.path.join(examples, "*.ft"))):
        name = os.path.basename(path)[:-3]
        try:
            state = load_scene_file(path, [library])
            scene = compile_scene(state)
        except Exception as e:  # noqa: BLE001 - report what the front end cannot compile
            skipped.append(f"{name}: {type(e).__name__}: {e}")
            continue
        s = extract_scene_render_settings(state)
        opts = np.array([s.max_steps or d.max_steps, s.max_dist or d.max_dist, s.epsilon or d.epsilon,
                         s.step_scale or d.step_scale, s.camera_z or d.camera_z,
                         s.fov_y_degrees or d.fov_y_degrees], dtype=np.float32)
        scene.save(os.path.join(OUT, name + ".npz"), options=opts,
                   accel=np.int32(s.accel if s.accel is not None else 1))
        written.append(name)
    print(f"wrote {len(written)} example IRs to {OUT}")
    for line in skipped:
        print("skipped", line)


if __name__ == "__main__":
    main()
