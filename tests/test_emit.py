"""Scene specialisation on the host: the emitter's output for the reference fixtures and the
synthetic workloads compiles with hiprtc for gfx950, carries every IR constant exactly, and
is used for the scenes the default policy selects.  (Its numerics are compared with the
oracle on the GPU in test_gpu_parity.py, FT_EMIT=1.)"""
import os
import re
import struct

import numpy as np
import pytest

from forgedthoughts_amd import render, synthetic
from forgedthoughts_amd.options import AccelMode

from helpers import REGRESSION, REGRESSION_CASES, compile_ft, compile_src
from test_gpu_parity import GLASS

pytestmark = pytest.mark.skipif(not os.path.exists(render.LIB_PATH), reason="libft_hip.so not built")

SCENES = [(c[0], lambda c=c: compile_ft(os.path.join(REGRESSION, c[0] + ".ft"))) for c in REGRESSION_CASES]
SCENES += [("materials", lambda: compile_src(GLASS)),
           ("whitted_1080p", lambda: compile_src(synthetic.whitted_1080p().source)),
           ("depth_512", lambda: compile_src(synthetic.depth_512().source))]


def hexfloats(text):
    return [float.fromhex(m) for m in re.findall(r"\((-?0x[0-9a-f.]+p[+-]?\d+)f\)", text)]


@pytest.mark.parametrize("accel", [AccelMode.Bvh, AccelMode.Naive], ids=["bvh", "naive"])
@pytest.mark.parametrize("name,make", SCENES, ids=[s[0] for s in SCENES])
def test_emitted_depth_kernel_compiles(name, make, accel):
    scene = make()
    render.emit_compile(scene, accel, 0)


def test_emitted_whitted_kernel_compiles():
    scene = compile_src(GLASS)
    for k in (1, 2):
        render.emit_compile(scene, AccelMode.Bvh, k)


def test_literals_are_exact():
    """Every primitive centre of the 1080p scene appears bit-exactly as a hex-float literal."""
    scene = compile_src(synthetic.whitted_1080p().source)
    src = render.emit_scene_source(scene, AccelMode.Bvh)
    lits = set(struct.pack("<f", v) for v in hexfloats(src))
    nodes = scene.nodes
    prims = nodes[nodes["kind"] <= 4]
    assert len(prims) == 64
    for c in prims["center"]:
        for v in c:
            if v != 0.0:
                assert struct.pack("<f", np.float32(v)) in lits, f"centre {v!r} not emitted exactly"


def test_plan_selection():
    """One-leaf scenes (the smooth-union 1080p tree) are emitted whole; multi-leaf BVH scenes
    keep the data-driven traversal and specialise only their Forge functions."""
    big = compile_src(synthetic.whitted_1080p().source)
    src = render.emit_scene_source(big, AccelMode.Bvh)
    assert "#define FT_EMITTED_SCENE 1" in src and "#define FT_EMITTED_FUNCTIONS 1" in src
    multi = compile_src(GLASS)
    src = render.emit_scene_source(multi, AccelMode.Bvh)
    assert "#define FT_EMITTED_SCENE 0" in src and "#define FT_EMITTED_FUNCTIONS 1" in src
    src = render.emit_scene_source(multi, AccelMode.Naive)
    assert "#define FT_EMITTED_SCENE 1" in src


def test_sweep_emission():
    """The shading sweep is emitted for a one-leaf scene whose root is a smooth boolean and
    that has no custom or modifier node; a hard-union root keeps the interpreted sweep."""
    big = compile_src(synthetic.whitted_1080p(n_prims=8).source)
    src = render.emit_scene_source(big, AccelMode.Bvh)
    assert "#define FT_EMITTED_SWEEP 1" in src and "void ftx_sweep(" in src
    assert src.count("tetra_normal(") == 8 and src.count("blend_refs(") == 6  # the split node keeps its refs
    hard = compile_src(synthetic.depth_512().source)
    src = render.emit_scene_source(hard, AccelMode.Naive)
    assert "#define FT_EMITTED_SWEEP 0" in src and "void ftx_sweep(" not in src


def test_emit_rejects_bad_ir():
    from forgedthoughts_amd.scene import RenderError
    bad = compile_src("let scene = Sphere { radius: 1.0 };")
    bad.nodes["kind"][0] = 99
    bad._ir = None
    with pytest.raises(RenderError):
        render.emit_scene_source(bad, AccelMode.Bvh)
    with pytest.raises(RenderError):
        render.emit_compile(compile_src("let scene = Sphere { radius: 1.0 };"), AccelMode.Bvh, 7)
