"""Host-side mirror of the reference render API over libft_hip.so.

Reference entry points (crates/forgedthoughts/src/renderer.rs):
  render_depth_png_with_accel            1139-1156
  render_preview_progressive_with_accel  1238-1277
  render_ray_progressive_with_accel      1201-1236
  render_ray_png_with_accel              1179-1199
Same names, argument meaning and error behaviour (RenderError kinds); the accelerator
selected by `AccelMode` runs on the GPU.  There is no CPU fallback: if the HIP library
or a GPU is missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Callable, Optional, Tuple

import numpy as np

from . import abi
from .forge.evaluator import EvalState
from .options import AccelMode, RaySettings, RenderOptions
from .scene import CompiledScene, RenderError, compile_scene

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libft_hip.so")
_lib = None

ERROR_KINDS = {1: "MissingSceneRoot", 2: "UnsupportedObjectType", 3: "ExpectedObject", 4: "Image",
               5: "InvalidArgument", 6: "GpuUnavailable", 7: "CompileFailed", 8: "OutOfMemory",
               9: "LaunchFailed", 10: "CallbackAbort"}

EXPORTS = ["ft_last_error", "ft_device_count", "ft_scene_create", "ft_scene_destroy", "ft_render_depth",
           "ft_render_whitted", "ft_render_depth_device", "ft_render_whitted_device",
           "ft_render_whitted_device_noalloc", "ft_scene_synchronize",
           "ft_scene_stream", "ft_scene_stats", "ft_abi_struct_size", "ft_scene_set_flop_counting",
           "ft_emit_scene_source", "ft_emit_compile", "ft_batch_create", "ft_batch_destroy",
           "ft_batch_render_device", "ft_batch_render", "ft_batch_synchronize", "ft_batch_stats",
           "ft_batch_set_flop_counting", "ft_diff_create", "ft_diff_destroy", "ft_diff_param_count",
           "ft_diff_set_params", "ft_diff_get_params", "ft_diff_forward", "ft_diff_forward_device",
           "ft_diff_synchronize", "ft_diff_backward_device", "ft_diff_set_target", "ft_diff_loss_grad",
           "ft_diff_fit", "ft_diff_stats", "ft_diff_set_flop_counting", "ft_scene_build_atlas", "ft_scene_set_launch_config",
           "ft_scene_get_launch_config", "ft_scene_autotune", "ft_render_depth_progressive",
           "ft_render_whitted_progressive"]


class BackendUnavailable(RuntimeError):
    pass


def library():
    """Loads libft_hip.so (built in-tree by forgedthoughts_amd/build.py)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BackendUnavailable(f"{LIB_PATH} is not built; run forgedthoughts_amd/build.py")
        # One HIP runtime per process: when PyTorch is present its runtime library must be
        # the one loaded first (ours then binds to it by soname); loaded the other way round,
        # PyTorch finds no GPU.  Device buffers handed over from torch need the same runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, u32 = C.c_void_p, C.c_uint32
        L.ft_last_error.restype = C.c_char_p
        L.ft_device_count.argtypes = [C.POINTER(C.c_int)]
        L.ft_scene_create.argtypes = [vp, u32, C.POINTER(vp)]
        L.ft_scene_destroy.argtypes = [vp]
        L.ft_render_depth.argtypes = [vp, vp, u32, vp, vp]
        L.ft_render_whitted.argtypes = [vp, vp, vp, vp, vp]
        L.ft_render_depth_device.argtypes = [vp, vp, u32, vp, vp]
        L.ft_render_whitted_device.argtypes = [vp, vp, vp, vp, vp]
        L.ft_render_whitted_device_noalloc.argtypes = [vp, vp, vp, vp, vp]
        L.ft_scene_synchronize.argtypes = [vp]
        L.ft_scene_stream.argtypes = [vp, C.POINTER(vp)]
        L.ft_scene_stats.argtypes = [vp, vp]
        L.ft_abi_struct_size.argtypes = [C.c_char_p]
        L.ft_scene_set_flop_counting.argtypes = [vp, C.c_int]
        L.ft_scene_build_atlas.argtypes = [vp, C.c_float, u32, u32]
        L.ft_scene_set_launch_config.argtypes = [vp, u32, vp]
        L.ft_scene_get_launch_config.argtypes = [vp, u32, vp]
        L.ft_scene_autotune.argtypes = [vp, u32, vp, vp, u32, vp]
        L.ft_render_depth_progressive.argtypes = [vp, vp, u32, u32, abi.PROGRESS_FN, vp, vp]
        L.ft_render_whitted_progressive.argtypes = [vp, vp, vp, abi.PROGRESS_FN, vp, vp]
        L.ft_emit_scene_source.argtypes = [vp, u32, C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64)]
        L.ft_emit_compile.argtypes = [vp, u32, C.c_int, C.POINTER(C.c_double)]
        L.ft_batch_create.argtypes = [C.POINTER(vp), u32, u32, C.POINTER(vp)]
        L.ft_batch_destroy.argtypes = [vp]
        L.ft_batch_render_device.argtypes = [vp, vp, u32, vp, vp, u32, vp, vp]
        L.ft_batch_render.argtypes = [vp, vp, u32, vp, vp, u32, vp, vp]
        L.ft_batch_synchronize.argtypes = [vp]
        L.ft_batch_stats.argtypes = [vp, vp]
        L.ft_batch_set_flop_counting.argtypes = [vp, C.c_int]
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        msg = library().ft_last_error().decode()
        kind = ERROR_KINDS.get(rc, "Unknown")
        if rc == 6:
            raise BackendUnavailable(msg)
        raise RenderError(kind, msg)


def device_count() -> int:
    n = C.c_int(0)
    _check(library().ft_device_count(C.byref(n)))
    return n.value


class GpuScene:
    """One scene uploaded to the GPU for one accel mode (Accelerator::from_scene)."""

    def __init__(self, scene: CompiledScene, accel: AccelMode = AccelMode.Bvh):
        self.scene = scene
        self.accel = AccelMode(accel)
        self._h = C.c_void_p()
        ir = scene.ir()
        _check(library().ft_scene_create(C.addressof(ir), int(self.accel), C.byref(self._h)))

    def close(self):
        if self._h:
            library().ft_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def render_depth(self, options: RenderOptions, aa_samples: int = 1) -> Tuple[np.ndarray, np.ndarray]:
        o = options.to_c()
        rgb = np.zeros((o.height, o.width, 3), np.uint8)
        t = np.zeros((o.height, o.width), np.float32)
        _check(library().ft_render_depth(self._h, C.addressof(o), int(aa_samples), rgb.ctypes.data, t.ctypes.data))
        return rgb, t

    def render_whitted(self, options: RenderOptions, settings: RaySettings) -> Tuple[np.ndarray, np.ndarray]:
        o = options.to_c()
        s = settings.to_c()
        rgb = np.zeros((o.height, o.width, 3), np.uint8)
        hdr = np.zeros((o.height, o.width, 3), np.float32)
        _check(library().ft_render_whitted(self._h, C.addressof(o), C.addressof(s), rgb.ctypes.data, hdr.ctypes.data))
        return rgb, hdr

    # device-resident variants (outputs already in HBM; used for throughput measurement)
    def render_depth_device(self, options: RenderOptions, aa_samples: int, rgb_ptr: int, t_ptr: int = 0):
        o = options.to_c()
        _check(library().ft_render_depth_device(self._h, C.addressof(o), int(aa_samples), rgb_ptr, t_ptr or None))

    def render_whitted_device(self, options: RenderOptions, settings: RaySettings, rgb_ptr: int, hdr_ptr: int = 0):
        o = options.to_c()
        s = settings.to_c()
        _check(library().ft_render_whitted_device(self._h, C.addressof(o), C.addressof(s), rgb_ptr, hdr_ptr or None))

    def synchronize(self):
        _check(library().ft_scene_synchronize(self._h))

    def _progressive(self, call, options: RenderOptions, on_tile) -> np.ndarray:
        o = options.to_c()
        rgb = np.zeros((o.height, o.width, 3), np.uint8)
        raised = []

        def trampoline(pg, user):
            p = pg.contents
            try:
                img = np.ctypeslib.as_array(p.image, shape=(p.height, p.width, 3)).copy()
                on_tile({"tiles_done": p.tiles_done, "tiles_total": p.tiles_total, "elapsed_ms": p.elapsed_ms,
                         "tile": (p.tile_x, p.tile_y, p.tile_w, p.tile_h)}, img)
            except BaseException as exc:  # stops the render; re-raised below
                raised.append(exc)
                return 1
            return 0
        cb = abi.PROGRESS_FN(trampoline) if on_tile is not None else abi.PROGRESS_FN()
        rc = call(C.addressof(o), cb, rgb.ctypes.data)
        if rc == 10 and raised:
            raise raised[0]
        _check(rc)
        return rgb

    def render_depth_progressive(self, options: RenderOptions, aa_samples: int, tile_size: int,
                                 on_tile=None) -> np.ndarray:
        """ft_render_depth_progressive: tiles of tile_size, on_tile(progress, image) per tile."""
        return self._progressive(lambda o, cb, out: library().ft_render_depth_progressive(
            self._h, o, int(aa_samples), int(tile_size), cb, None, out), options, on_tile)

    def render_whitted_progressive(self, options: RenderOptions, settings: RaySettings, on_tile=None) -> np.ndarray:
        s = settings.to_c()
        return self._progressive(lambda o, cb, out: library().ft_render_whitted_progressive(
            self._h, o, C.addressof(s), cb, None, out), options, on_tile)

    def stream(self) -> int:
        p = C.c_void_p()
        _check(library().ft_scene_stream(self._h, C.byref(p)))
        return p.value or 0

    def set_flop_counting(self, enable: bool):
        _check(library().ft_scene_set_flop_counting(self._h, 1 if enable else 0))

    def build_atlas(self, voxel: float = 0.02, bits: int = 16, max_bricks: int = 0):
        """Distance atlas for coarse march steps (bits 16 or 32; 0 removes it)."""
        _check(library().ft_scene_build_atlas(self._h, float(voxel), int(bits), int(max_bricks)))

    def stats(self) -> abi.RenderStatsC:
        st = abi.RenderStatsC()
        _check(library().ft_scene_stats(self._h, C.addressof(st)))
        return st

    def launch_config(self, kind: int) -> abi.LaunchConfigC:
        c = abi.LaunchConfigC()
        _check(library().ft_scene_get_launch_config(self._h, int(kind), C.addressof(c)))
        return c

    def set_launch_config(self, kind: int, block_waves=1, tile_width=None, waves_per_eu=4, march_unroll=1,
                          shade_inline=0):
        """Sets the launch shape of the scene's `kind` kernel (ft_scene_set_launch_config); the
        defaults are the library's: one wave per workgroup, 8x8 pixel tiles for depth and
        sample-major waves (tile width 0) for Whitted."""
        if tile_width is None:
            tile_width = 0 if kind == abi.KIND_WHITTED else 8
        c = abi.LaunchConfigC(block_waves, tile_width, waves_per_eu, march_unroll, shade_inline)
        _check(library().ft_scene_set_launch_config(self._h, int(kind), C.addressof(c)))

    def autotune(self, kind: int, options: RenderOptions, settings: Optional[RaySettings] = None,
                 flags: int = 0) -> abi.TuneReportC:
        """Chooses the fastest launch shape for this scene's `kind` render (ft_scene_autotune)."""
        o = options.to_c()
        s = settings.to_c() if settings is not None else None
        rep = abi.TuneReportC()
        _check(library().ft_scene_autotune(self._h, int(kind), C.addressof(o), C.addressof(s) if s is not None else None,
                                           int(flags), C.addressof(rep)))
        return rep


class GpuBatch:
    """Many scenes rendered from many views per call (ft_batch_*): one launch covers every
    (scene, view) image; FT_BATCH_SEPARATE issues the single-scene launches instead."""

    def __init__(self, scenes, accel: AccelMode = AccelMode.Bvh):
        self.scenes = list(scenes)
        self._irs = [s.ir() for s in self.scenes]
        arr = (C.c_void_p * len(self._irs))(*[C.addressof(ir) for ir in self._irs])
        self._h = C.c_void_p()
        _check(library().ft_batch_create(arr, len(self._irs), int(accel), C.byref(self._h)))

    def close(self):
        if self._h:
            library().ft_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def views(items) -> "C.Array":
        """items: (scene index, origin xyz, target xyz, fov_y degrees)."""
        arr = (abi.View * len(items))()
        for k, (sc, o, t, fov) in enumerate(items):
            arr[k].scene = sc
            arr[k].fov_y_degrees = fov
            arr[k].origin[:] = [float(x) for x in o]
            arr[k].target[:] = [float(x) for x in t]
        return arr

    def render(self, views, options: RenderOptions, settings: Optional[RaySettings], flags: int):
        """Host outputs: (depth[n,h,w,3] or None, whitted[n,h,w,3] or None)."""
        o = options.to_c()
        s = settings.to_c() if settings is not None else None
        n = len(views)
        depth = np.zeros((n, o.height, o.width, 3), np.uint8) if flags & abi.BATCH_DEPTH else None
        whit = np.zeros((n, o.height, o.width, 3), np.uint8) if flags & abi.BATCH_WHITTED else None
        _check(library().ft_batch_render(self._h, C.addressof(views), n, C.addressof(o),
                                         C.addressof(s) if s is not None else None, int(flags),
                                         depth.ctypes.data if depth is not None else None,
                                         whit.ctypes.data if whit is not None else None))
        return depth, whit

    def render_device(self, views, options: RenderOptions, settings: Optional[RaySettings], flags: int,
                      depth_ptr: int = 0, whitted_ptr: int = 0):
        o = options.to_c()
        s = settings.to_c() if settings is not None else None
        _check(library().ft_batch_render_device(self._h, C.addressof(views), len(views), C.addressof(o),
                                                C.addressof(s) if s is not None else None, int(flags),
                                                depth_ptr or None, whitted_ptr or None))

    def synchronize(self):
        _check(library().ft_batch_synchronize(self._h))

    def set_flop_counting(self, enable: bool):
        _check(library().ft_batch_set_flop_counting(self._h, 1 if enable else 0))

    def stats(self) -> abi.RenderStatsC:
        st = abi.RenderStatsC()
        _check(library().ft_batch_stats(self._h, C.addressof(st)))
        return st


def emit_scene_source(scene: CompiledScene, accel: AccelMode = AccelMode.Bvh) -> str:
    """The scene-specialised HIP translation unit the backend would compile for `scene`."""
    ir = scene.ir()
    n = C.c_uint64(0)
    _check(library().ft_emit_scene_source(C.addressof(ir), int(accel), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    _check(library().ft_emit_scene_source(C.addressof(ir), int(accel), buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def emit_compile(scene: CompiledScene, accel: AccelMode = AccelMode.Bvh, kernel: int = 0) -> float:
    """Compiles the scene-specialised kernel with hiprtc (no GPU needed); returns compile ms
    (0 when served from the code-object cache)."""
    ir = scene.ir()
    ms = C.c_double(0.0)
    _check(library().ft_emit_compile(C.addressof(ir), int(accel), int(kernel), C.byref(ms)))
    return ms.value


# ---------------------------------------------------------------- reference API
def render_preview_progressive_with_accel(state: EvalState, options: RenderOptions, accel_mode: AccelMode,
                                          tile_size: int, aa_samples: int,
                                          on_tile: Optional[Callable[[dict, np.ndarray], None]] = None) -> np.ndarray:
    """Depth preview (renderer.rs:1238-1277, 3751-3834); `on_tile(progress, image)` runs on
    the caller's thread once per finished tile of `tile_size`, in the reference's tile
    order, with the image so far.  An exception raised by the callback stops the render
    and propagates, as an Err from the reference's callback is returned."""
    scene = compile_scene(state)
    with GpuScene(scene, accel_mode) as gs:
        return gs.render_depth_progressive(options, max(int(aa_samples), 1), tile_size, on_tile)


def render_ray_progressive_with_accel(state: EvalState, options: RenderOptions, accel_mode: AccelMode,
                                      settings: RaySettings,
                                      on_tile: Optional[Callable[[dict, np.ndarray], None]] = None) -> np.ndarray:
    """Whitted trace (renderer.rs:1201-1236, 3635-3724), tiles of settings.tile_size."""
    scene = compile_scene(state)
    with GpuScene(scene, accel_mode) as gs:
        return gs.render_whitted_progressive(options, settings, on_tile)


def render_depth_png_with_accel(state: EvalState, output_path: str, options: RenderOptions,
                                accel_mode: AccelMode) -> None:
    """render_depth_png_with_accel (renderer.rs:1139-1156): one tile, 1 spp, saved as PNG."""
    rgb = render_preview_progressive_with_accel(state, options, accel_mode, max(options.width, options.height), 1)
    save_png(rgb, output_path)


def render_ray_png_with_accel(state: EvalState, output_path: str, options: RenderOptions,
                              accel_mode: AccelMode, max_depth: int) -> None:
    settings = RaySettings(max_depth=max(max_depth, 1), tile_size=max(options.width, options.height), aa_samples=1)
    save_png(render_ray_progressive_with_accel(state, options, accel_mode, settings), output_path)


def save_png(rgb: np.ndarray, path: str) -> None:
    try:
        from PIL import Image
    except ImportError as exc:  # pragma: no cover
        raise RenderError("Image", f"failed to write png: {exc}")
    try:
        Image.fromarray(rgb, "RGB").save(path)
    except Exception as exc:
        raise RenderError("Image", f"failed to write png: {exc}")
