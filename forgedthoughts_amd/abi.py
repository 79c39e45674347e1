"""ctypes / numpy mirror of include/ft_abi.h (the scene IR and render arguments).

Every struct here must keep the exact field order and size of its C counterpart; the
`test_abi_layout` test compares these sizes with `sizeof` values exported by the built
library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

ABI_VERSION = 6

ACCEL_NAIVE, ACCEL_BVH, ACCEL_BRICKS = 0, 1, 2

NODE_KINDS = [
    "SPHERE", "BOX", "CYLINDER", "TORUS", "EXTRUDE_POLYGON", "CUSTOM", "DOMAIN_MODIFIER",
    "DISTANCE_POST_MODIFIER", "NOISE", "UNION", "INTERSECT", "SUBTRACT", "UNION_ROUND",
    "UNION_CHAMFER", "UNION_COLUMNS", "UNION_STAIRS", "UNION_SOFT", "INTERSECT_ROUND",
    "INTERSECT_CHAMFER", "INTERSECT_COLUMNS", "INTERSECT_STAIRS", "DIFF_ROUND", "DIFF_CHAMFER",
    "DIFF_COLUMNS", "DIFF_STAIRS", "PIPE", "ENGRAVE", "GROOVE", "TONGUE", "SLICE", "SMOOTH",
]
NODE = {name: i for i, name in enumerate(NODE_KINDS)}

MODEL_STANDARD, MODEL_LAMBERT, MODEL_METAL, MODEL_DIELECTRIC = 0, 1, 2, 3
PATTERN_NONE, PATTERN_CHECKER3D = 0, 1
LIGHT_POINT, LIGHT_ENV, LIGHT_SPHERE = 0, 1, 2

AOV = {None: -1, "depth": 0, "normal": 1, "material_id": 2, "ior": 3, "transmission": 4,
       "fresnel": 5, "hit_t": 6}

NODE_DTYPE = np.dtype([
    ("kind", "<u4"), ("lhs", "<i4"), ("rhs", "<i4"), ("object_id", "<u4"), ("material_id", "<u4"),
    ("fn_domain", "<i4"), ("fn_distance", "<i4"), ("fn_post", "<i4"),
    ("center", "<f4", 3), ("rot_deg", "<f4", 3), ("param", "<f4", 6),
    ("bounds_min", "<f4", 3), ("bounds_max", "<f4", 3),
])
assert NODE_DTYPE.itemsize == 104

EXPR_DTYPE = np.dtype([
    ("op", "<u4"), ("type", "<u4"), ("a", "<i4"), ("b", "<i4"), ("c", "<i4"), ("d", "<i4"),
    ("value", "<f4"), ("pad", "<f4"),
])
assert EXPR_DTYPE.itemsize == 32

FUNCTION_DTYPE = np.dtype([("first", "<u4"), ("count", "<u4"), ("n_params", "<u4"), ("result_type", "<u4")])

MATERIAL_DTYPE = np.dtype([
    ("model", "<u4"), ("thin_walled", "<u4"), ("has_medium", "<u4"), ("pattern", "<u4"),
    ("color", "<f4", 3), ("roughness", "<f4"), ("metallic", "<f4"), ("specular", "<f4"),
    ("specular_weight", "<f4"), ("ior", "<f4"), ("specular_color", "<f4", 3), ("clearcoat", "<f4"),
    ("clearcoat_roughness", "<f4"), ("transmission", "<f4"), ("emission_strength", "<f4"),
    ("medium_ior", "<f4"), ("emission_color", "<f4", 3), ("medium_density", "<f4"),
    ("medium_absorption", "<f4", 3), ("pattern_scale", "<f4"), ("pattern_color_a", "<f4", 3),
    ("dynamic_material_id", "<i4"), ("pattern_color_b", "<f4", 3), ("dynamic_override_id", "<i4"),
])
assert MATERIAL_DTYPE.itemsize == 144

HOOKS_DTYPE = np.dtype([
    ("material", "<u4"), ("color", "<i4"), ("roughness", "<i4"), ("ior", "<i4"), ("thin_walled", "<i4"),
    ("emission_color", "<i4"), ("emission_strength", "<i4"), ("metallic", "<i4"), ("specular", "<i4"),
    ("specular_weight", "<i4"), ("transmission", "<i4"), ("specular_color", "<i4"),
    ("clearcoat_roughness", "<i4"), ("clearcoat", "<i4"), ("coat_mask", "<i4"), ("bump", "<i4"),
    ("normal", "<i4"), ("medium_ior", "<i4"), ("medium_absorption", "<i4"), ("medium_density", "<i4"),
    ("eval", "<i4"), ("pad", "<i4"),
])
assert HOOKS_DTYPE.itemsize == 88
HOOK_PARAMS = 13

LIGHT_DTYPE = np.dtype([
    ("kind", "<u4"), ("samples", "<u4"), ("radius", "<f4"), ("pad", "<f4"),
    ("position", "<f4", 3), ("pad1", "<f4"), ("radiance", "<f4", 3), ("pad2", "<f4"),
])
assert LIGHT_DTYPE.itemsize == 48


class Camera(C.Structure):
    _fields_ = [
        ("present", C.c_uint32), ("has_origin", C.c_uint32), ("has_target", C.c_uint32),
        ("has_fov", C.c_uint32), ("origin", C.c_float * 3), ("fov_y_degrees", C.c_float),
        ("target", C.c_float * 3), ("pad", C.c_float),
    ]


class SceneIR(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("root", C.c_uint32), ("n_nodes", C.c_uint32),
        ("n_materials", C.c_uint32), ("n_objects", C.c_uint32), ("n_lights", C.c_uint32),
        ("n_functions", C.c_uint32), ("n_exprs", C.c_uint32), ("env_fn", C.c_int32),
        ("pad", C.c_uint32),
        ("nodes", C.c_void_p), ("materials", C.c_void_p), ("object_transforms", C.c_void_p),
        ("lights", C.c_void_p), ("functions", C.c_void_p), ("exprs", C.c_void_p),
        ("camera", Camera),
        ("n_material_hooks", C.c_uint32), ("pad2", C.c_uint32), ("material_hooks", C.c_void_p),
    ]


class RenderOptionsC(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("max_steps", C.c_uint32),
        ("max_dist", C.c_float), ("epsilon", C.c_float), ("step_scale", C.c_float),
        ("camera_z", C.c_float), ("fov_y_degrees", C.c_float),
    ]


class RaySettingsC(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("tile_size", C.c_uint32),
                ("aa_samples", C.c_uint32), ("debug_aov", C.c_int32)]


class RenderStatsC(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("kernel_launches", C.c_uint64),
                ("last_kernel_ms", C.c_float), ("flags", C.c_uint32), ("sdf_flops", C.c_uint64),
                ("compile_ms", C.c_double), ("atlas_samples", C.c_uint64), ("atlas_bytes", C.c_uint64),
                ("atlas_bricks", C.c_uint64), ("ray_stack", C.c_uint32)]


STATS_EMITTED_SCENE, STATS_EMITTED_FUNCTIONS = 1, 2


class View(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("fov_y_degrees", C.c_float), ("origin", C.c_float * 3),
                ("target", C.c_float * 3)]


BATCH_DEPTH, BATCH_WHITTED, BATCH_SEPARATE, BATCH_GRAPH, BATCH_TILED = 1, 2, 4, 8, 16


class LaunchConfigC(C.Structure):
    _fields_ = [("block_waves", C.c_uint32), ("tile_width", C.c_uint32), ("waves_per_eu", C.c_uint32),
                ("march_unroll", C.c_uint32), ("shade_inline", C.c_uint32)]

    def as_dict(self):
        tile = f"{self.tile_width}x{64 // self.tile_width}" if self.tile_width else "samples of a pixel block"
        return {"block_waves": self.block_waves, "tile": tile,
                "waves_per_eu": self.waves_per_eu, "march_unroll": self.march_unroll,
                "shade_inline": bool(self.shade_inline)}


class TuneReportC(C.Structure):
    _fields_ = [("best", LaunchConfigC), ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("tune_ms", C.c_double), ("candidates", C.c_uint32), ("cached", C.c_uint32),
                ("tune_aa_samples", C.c_uint32), ("pad", C.c_uint32)]


class ProgressC(C.Structure):
    _fields_ = [("tiles_done", C.c_uint32), ("tiles_total", C.c_uint32), ("elapsed_ms", C.c_uint64),
                ("tile_x", C.c_uint32), ("tile_y", C.c_uint32), ("tile_w", C.c_uint32), ("tile_h", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("image", C.POINTER(C.c_uint8))]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.POINTER(ProgressC), C.c_void_p)

KIND_DEPTH, KIND_WHITTED = 0, 1
TUNE_FULL, TUNE_SHAPE_ONLY, TUNE_NO_CACHE = 1, 2, 4

STRUCT_SIZES = {
    "ft_node": NODE_DTYPE.itemsize, "ft_expr": EXPR_DTYPE.itemsize,
    "ft_function": FUNCTION_DTYPE.itemsize, "ft_material": MATERIAL_DTYPE.itemsize,
    "ft_material_hooks": HOOKS_DTYPE.itemsize,
    "ft_light": LIGHT_DTYPE.itemsize, "ft_camera": C.sizeof(Camera),
    "ft_scene_ir": C.sizeof(SceneIR), "ft_render_options": C.sizeof(RenderOptionsC),
    "ft_ray_settings": C.sizeof(RaySettingsC), "ft_render_stats": C.sizeof(RenderStatsC),
    "ft_view": C.sizeof(View), "ft_launch_config": C.sizeof(LaunchConfigC),
    "ft_tune_report": C.sizeof(TuneReportC), "ft_progress": C.sizeof(ProgressC),
}
