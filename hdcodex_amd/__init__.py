"""hdcodex_amd — MI355X-native spectral path-tracing core (Python binding).

The native library ``libhdcx.so`` (built in-tree by ``hdcodex_amd.build``)
carries the whole hot path; this module binds its C ABI
(``include/hdcx.h``) with ctypes and mirrors the reference integrator class
``hdcodex::VulkanPathTracer`` (include/hdcodex/gpu/vulkan_path_tracer.h:26-52)
as :class:`PathTracer`. There is no Python or CPU fallback: importing works
without the library, but every call that needs it raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_uint8, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libhdcx.so"

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_INVALID_SCENE = 2
ERR_OUT_OF_MEMORY = 3
ERR_DEVICE = 4
ERR_NO_GEOMETRY = 5
ERR_UNSUPPORTED = 6
ERR_INTERNAL = 7

SHADING_FUSED = 0
SHADING_MODULAR = 1
SHADING_RASTER_PREVIEW = 2

KERNELS = ("raygen", "extend", "shade", "shadow", "film", "cache", "walk", "pair", "miss")

# Every symbol include/hdcx.h declares.
EXPORTED_SYMBOLS = (
    "hdcx_last_error", "hdcx_create", "hdcx_destroy", "hdcx_set_scene", "hdcx_validate_scene", "hdcx_set_shading_mode",
    "hdcx_get_shading_mode", "hdcx_render", "hdcx_read_film", "hdcx_set_sample_offset", "hdcx_has_geometry", "hdcx_trace_closest",
    "hdcx_trace_occluded", "hdcx_scene_create_procedural", "hdcx_scene_view_get", "hdcx_scene_camera_get",
    "hdcx_scene_destroy", "hdcx_scene_flatten", "hdcx_bvh_build_host", "hdcx_set_instrumentation", "hdcx_get_kernel_stats", "hdcx_reset_kernel_stats",
    "hdcx_read_emissive_table", "hdcx_read_dome_importance", "hdcx_read_film_upscaled",
)


class Camera(ctypes.Structure):
    _fields_ = [("origin", c_float * 3), ("lower_left", c_float * 3), ("horizontal", c_float * 3),
                ("vertical", c_float * 3)]

    @classmethod
    def make(cls, origin, lower_left, horizontal, vertical) -> "Camera":
        c = cls()
        for name, value in (("origin", origin), ("lower_left", lower_left), ("horizontal", horizontal),
                            ("vertical", vertical)):
            getattr(c, name)[:] = [float(x) for x in value]
        return c

    def as_array(self) -> np.ndarray:
        return np.array([list(self.origin), list(self.lower_left), list(self.horizontal), list(self.vertical)],
                        dtype=np.float32)


class TextureView(ctypes.Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("format", c_uint32), ("reserved", c_uint32),
                ("texels", c_void_p)]


class SceneView(ctypes.Structure):
    _fields_ = [
        ("vertex_count", c_uint32), ("triangle_count", c_uint32),
        ("positions", POINTER(c_float)), ("indices", POINTER(c_uint32)), ("triangle_materials", POINTER(c_uint32)),
        ("texcoords", POINTER(c_float)), ("normals", POINTER(c_float)), ("emissive", POINTER(c_float)),
        ("material_count", c_uint32), ("light_count", c_uint32),
        ("materials", c_void_p), ("lights", c_void_p),
        ("texture_count", c_uint32), ("opaque", c_uint32),
        ("textures", POINTER(TextureView)),
        ("prototype_count", c_uint32), ("prototypes", c_void_p),
        ("instance_count", c_uint32), ("instances", c_void_p),
        ("material_emission_power", POINTER(c_float)), ("light_triangle_count", c_uint32), ("reserved", c_uint32),
    ]


class KernelStats(ctypes.Structure):
    _fields_ = [
        ("launches", c_uint64 * len(KERNELS)), ("milliseconds", c_double * len(KERNELS)), ("rays", c_uint64 * len(KERNELS)),
        ("nodes_visited", c_uint64), ("triangles_tested", c_uint64), ("extend_rays_counted", c_uint64),
        ("shadow_rays_counted", c_uint64), ("shadow_nodes_visited", c_uint64), ("shadow_triangles_tested", c_uint64),
        ("node_bytes", c_uint64), ("triangle_bytes", c_uint64), ("bvh_nodes", c_uint64), ("bvh_triangles", c_uint64),
        ("extend_hits_counted", c_uint64), ("shadow_queued_paths", c_uint64),
    ]

    def as_dict(self) -> dict:
        out = {}
        for i, k in enumerate(KERNELS):
            out[k] = {"launches": int(self.launches[i]), "ms": float(self.milliseconds[i]), "items": int(self.rays[i])}
        for name, _ in self._fields_[3:]:
            out[name] = int(getattr(self, name))
        return out


class BvhInfo(ctypes.Structure):
    _fields_ = [("node_count", c_uint32), ("slot_count", c_uint32), ("max_stack", c_uint32), ("depth", c_uint32),
                ("builder", ctypes.c_int32), ("top_clusters", c_uint32), ("refits", c_uint32),
                ("reserved", c_uint32), ("build_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class HdcxError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"hdcx error {code}: {message}")
        self.code = code


_lib = None


def library() -> ctypes.CDLL:
    """Loads libhdcx.so; raises RuntimeError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hdcodex_amd.build` "
                           "(there is no fallback implementation)")
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.hdcx_last_error.restype = c_char_p
    lib.hdcx_create.argtypes = [c_int, POINTER(c_void_p)]
    lib.hdcx_destroy.argtypes = [c_void_p]
    lib.hdcx_set_scene.argtypes = [c_void_p, POINTER(SceneView)]
    lib.hdcx_validate_scene.argtypes = [POINTER(SceneView)]
    lib.hdcx_set_shading_mode.argtypes = [c_void_p, c_int]
    lib.hdcx_get_shading_mode.argtypes = [c_void_p, POINTER(c_int)]
    lib.hdcx_render.argtypes = [c_void_p, POINTER(Camera), c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, c_void_p]
    lib.hdcx_read_film.argtypes = [c_void_p, c_void_p]
    lib.hdcx_set_sample_offset.argtypes = [c_void_p, c_uint32]
    lib.hdcx_has_geometry.argtypes = [c_void_p]
    lib.hdcx_trace_closest.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p]
    lib.hdcx_trace_occluded.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p]
    lib.hdcx_scene_create_procedural.argtypes = [c_char_p, c_uint64, c_uint32, c_uint32, POINTER(c_void_p)]
    lib.hdcx_scene_view_get.argtypes = [c_void_p, POINTER(SceneView)]
    lib.hdcx_scene_camera_get.argtypes = [c_void_p, c_uint32, c_uint32, POINTER(Camera)]
    lib.hdcx_scene_destroy.argtypes = [c_void_p]
    lib.hdcx_scene_flatten.argtypes = [c_void_p, POINTER(c_void_p)]
    lib.hdcx_bvh_build_host.argtypes = [POINTER(SceneView), POINTER(BvhInfo), c_void_p, c_void_p]
    lib.hdcx_set_instrumentation.argtypes = [c_void_p, c_uint32]
    lib.hdcx_get_kernel_stats.argtypes = [c_void_p, POINTER(KernelStats)]
    lib.hdcx_read_emissive_table.argtypes = [c_void_p, c_void_p, c_uint32]
    lib.hdcx_read_film_upscaled.argtypes = [c_void_p, c_uint32, c_uint32, c_void_p, c_int]
    lib.hdcx_read_dome_importance.argtypes = [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_uint32), c_void_p,
                                              c_void_p, c_void_p]
    lib.hdcx_reset_kernel_stats.argtypes = [c_void_p]
    _bind_cache(lib)
    _bind_bvh(lib)
    _bind_batch(lib)
    _lib = lib
    return lib


def _check(code: int) -> None:
    if code != OK:
        raise HdcxError(code, library().hdcx_last_error().decode())


class ProceduralScene:
    """A seeded procedural scene packed by the native packer ("cornell",
    "heightfield", "triangle", "materials")."""

    def __init__(self, name: str, seed: int = 0, param: int = 0, dome_width: int = 0):
        lib = library()
        self._h = c_void_p()
        _check(lib.hdcx_scene_create_procedural(name.encode(), seed, param, dome_width, ctypes.byref(self._h)))
        self.name = name
        self.view = SceneView()
        _check(lib.hdcx_scene_view_get(self._h, ctypes.byref(self.view)))

    def camera(self, width: int, height: int) -> Camera:
        cam = Camera()
        _check(library().hdcx_scene_camera_get(self._h, width, height, ctypes.byref(cam)))
        return cam

    def flattened(self) -> "ProceduralScene":
        """The same scene with every instance written out as triangles."""
        lib = library()
        out = ProceduralScene.__new__(ProceduralScene)
        out._h = c_void_p()
        _check(lib.hdcx_scene_flatten(self._h, ctypes.byref(out._h)))
        out.name = self.name + "/flattened"
        out.view = SceneView()
        _check(lib.hdcx_scene_view_get(out._h, ctypes.byref(out.view)))
        return out

    @property
    def instance_count(self) -> int:
        return int(self.view.instance_count)

    @property
    def triangle_count(self) -> int:
        return int(self.view.triangle_count)

    def arrays(self) -> dict:
        """Copies of the packed buffers as numpy arrays."""
        v = self.view
        t, n = v.triangle_count, v.vertex_count
        return {
            "positions": np.ctypeslib.as_array(v.positions, shape=(n, 4)).copy(),
            "indices": np.ctypeslib.as_array(v.indices, shape=(t, 3)).copy(),
            "triangle_materials": np.ctypeslib.as_array(v.triangle_materials, shape=(t,)).copy(),
            "texcoords": np.ctypeslib.as_array(v.texcoords, shape=(t, 3, 2)).copy(),
            "normals": np.ctypeslib.as_array(v.normals, shape=(t, 3, 4)).copy(),
            # None: the emissive-triangle table is built on the device.
            "emissive": np.ctypeslib.as_array(v.emissive, shape=(t, 4)).copy() if v.emissive else None,
            "material_emission_power": (np.ctypeslib.as_array(v.material_emission_power, shape=(v.material_count,)).copy()
                                        if v.material_emission_power else None),
            "light_triangle_count": int(v.light_triangle_count),
            "materials": np.frombuffer(ctypes.string_at(v.materials, 368 * v.material_count), dtype=np.uint8)
            .reshape(v.material_count, 368).copy(),
            "lights": np.frombuffer(ctypes.string_at(v.lights, 224 * max(v.light_count, 1)), dtype=np.uint8)
            .reshape(max(v.light_count, 1), 224).copy(),
            "light_count": int(v.light_count),
            "opaque": bool(v.opaque),
            "textures": [(v.textures[i].width, v.textures[i].height, v.textures[i].format)
                         for i in range(v.texture_count)],
        }

    def close(self) -> None:
        if self._h:
            library().hdcx_scene_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BVH_NODE_DTYPE = np.dtype([("lox", np.float32, 4), ("hix", np.float32, 4), ("loy", np.float32, 4),
                           ("hiy", np.float32, 4), ("loz", np.float32, 4), ("hiz", np.float32, 4),
                           ("child", np.uint32, 4), ("pad", np.uint32, 4)])


def build_bvh_host(scene: "ProceduralScene"):
    """Host-built BVH4 of a scene: (info dict, nodes, slots[(n, 3, 4) float32])."""
    lib = library()
    info = BvhInfo()
    _check(lib.hdcx_bvh_build_host(ctypes.byref(scene.view), ctypes.byref(info), None, None))
    nodes = np.zeros(info.node_count, dtype=BVH_NODE_DTYPE)
    slots = np.zeros((info.slot_count, 3, 4), dtype=np.float32)
    _check(lib.hdcx_bvh_build_host(ctypes.byref(scene.view), ctypes.byref(info), nodes.ctypes.data, slots.ctypes.data))
    return info.as_dict(), nodes, slots


def validate_scene(view: SceneView) -> None:
    """Runs the scene checks of hdcx_set_scene without a device."""
    _check(library().hdcx_validate_scene(ctypes.byref(view)))


def rays_array(origins: np.ndarray, directions: np.ndarray, tmin: float = 0.001, tmax: float = 1e4) -> np.ndarray:
    """Packs rays as the (N, 8) float32 layout of hdcx_ray."""
    n = len(origins)
    r = np.empty((n, 8), dtype=np.float32)
    r[:, 0:3] = origins
    r[:, 3] = tmin
    r[:, 4:7] = directions
    r[:, 7] = tmax
    return r


class PathTracer:
    """Mirror of hdcodex::VulkanPathTracer over the C ABI.

    ``render`` returns linear RGBA32F of shape (height, width, 4), row 0 at
    the bottom, or an empty array when there is no geometry or the size is
    zero (the reference returns an empty vector, :3235).
    """

    def __init__(self, device: int = 0):
        lib = library()
        self._h = c_void_p()
        _check(lib.hdcx_create(device, ctypes.byref(self._h)))

    def set_scene(self, scene: "ProceduralScene | SceneView | None") -> None:
        view = scene.view if isinstance(scene, ProceduralScene) else scene
        _check(library().hdcx_set_scene(self._h, ctypes.byref(view) if view is not None else None))

    def set_shading_mode(self, mode: int) -> None:
        code = library().hdcx_set_shading_mode(self._h, mode)
        if code == ERR_INVALID_ARGUMENT:
            raise ValueError(library().hdcx_last_error().decode())
        _check(code)

    def get_shading_mode(self) -> int:
        m = c_int()
        _check(library().hdcx_get_shading_mode(self._h, ctypes.byref(m)))
        return int(m.value)

    def has_geometry(self) -> bool:
        return bool(library().hdcx_has_geometry(self._h))

    def render(self, camera: Camera, width: int, height: int, sample_index: int, max_bounces: int = 8,
               sample_count: int = 1, readback: bool = True) -> np.ndarray:
        out = np.empty((height, width, 4), dtype=np.float32) if readback else None
        code = library().hdcx_render(self._h, ctypes.byref(camera), width, height, sample_index, max_bounces,
                                     sample_count, out.ctypes.data if out is not None else None)
        if code == ERR_NO_GEOMETRY:
            return np.empty((0,), dtype=np.float32)
        _check(code)
        return out if out is not None else np.empty((0,), dtype=np.float32)

    def set_sample_offset(self, offset: int) -> None:
        """Seeds sample number n as n + offset (independent sample streams)."""
        _check(library().hdcx_set_sample_offset(self._h, offset))

    def read_film(self, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), dtype=np.float32)
        _check(library().hdcx_read_film(self._h, out.ctypes.data))
        return out

    def read_film_upscaled(self, width: int, height: int) -> np.ndarray:
        """The last film upscaled by nearest neighbour on the device."""
        out = np.empty((height, width, 4), dtype=np.float32)
        _check(library().hdcx_read_film_upscaled(self._h, width, height, out.ctypes.data, 0))
        return out

    def emissive_table(self, triangle_count: int) -> np.ndarray:
        """The emissive-triangle table the device samples: (n, 4) float32
        rows (cdf, pdf / area, area, 0)."""
        out = np.empty((triangle_count, 4), dtype=np.float32)
        _check(library().hdcx_read_emissive_table(self._h, out.ctypes.data, triangle_count))
        return out

    def dome_importance(self, light: int):
        """(marginal[h], conditional[h, w], pdf[h, w]) of a dome light, or
        None when it is sampled uniformly."""
        w, h = c_uint32(), c_uint32()
        lib = library()
        _check(lib.hdcx_read_dome_importance(self._h, light, ctypes.byref(w), ctypes.byref(h), None, None, None))
        if w.value == 0:
            return None
        marginal = np.empty(h.value, dtype=np.float32)
        conditional = np.empty((h.value, w.value), dtype=np.float32)
        pdf = np.empty((h.value, w.value), dtype=np.float32)
        _check(lib.hdcx_read_dome_importance(self._h, light, ctypes.byref(w), ctypes.byref(h), marginal.ctypes.data,
                                             conditional.ctypes.data, pdf.ctypes.data))
        return marginal, conditional, pdf

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        """rays: (N, 8) float32 -> structured hits (t, primitive, u, v)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        hits = np.empty(len(rays), dtype=HIT_DTYPE)
        _check(library().hdcx_trace_closest(self._h, rays.ctypes.data, len(rays), hits.ctypes.data))
        return hits

    def trace_occluded(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        out = np.empty(len(rays), dtype=np.uint8)
        _check(library().hdcx_trace_occluded(self._h, rays.ctypes.data, len(rays), out.ctypes.data))
        return out

    def set_instrumentation(self, flags: int) -> None:
        _check(library().hdcx_set_instrumentation(self._h, flags))

    def kernel_stats(self) -> dict:
        s = KernelStats()
        _check(library().hdcx_get_kernel_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def reset_kernel_stats(self) -> None:
        _check(library().hdcx_reset_kernel_stats(self._h))

    def close(self) -> None:
        if self._h:
            library().hdcx_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


HIT_DTYPE = np.dtype([("t", np.float32), ("primitive", np.uint32), ("u", np.float32), ("v", np.float32)])


# ---------------------------------------------------------------- cache ----
class RadianceCacheConfig(ctypes.Structure):
    _fields_ = [("enabled", c_uint32), ("train_samples", c_uint32), ("terminate_bounce", c_uint32),
                ("train_steps", c_uint32), ("learning_rate", c_float), ("beta1", c_float), ("beta2", c_float),
                ("epsilon", c_float), ("ema_decay", c_float), ("spread_threshold", c_float), ("seed", c_uint64),
                ("calibration_decay", c_float), ("flags", c_uint32)]


class RcStats(ctypes.Structure):
    _fields_ = [("frames", c_uint64), ("queries", c_uint64), ("train_samples", c_uint64), ("adam_steps", c_uint64),
                ("last_loss", c_float), ("calibration_scale", c_float), ("infer_ms", c_double), ("train_ms", c_double)]


RC_PARAMS = 24768
RC_PLAIN_TRAINING_PATHS = 1
RC_QUERY_DTYPE = np.dtype([("position", np.float32, 3), ("direction", np.float32, 3), ("lambda0", np.float32),
                           ("roughness", np.float32), ("albedo", np.float32), ("lane", np.uint32),
                           ("path", np.uint32), ("target", np.float32)])


def _bind_cache(lib: ctypes.CDLL) -> None:
    lib.hdcx_set_radiance_cache.argtypes = [c_void_p, POINTER(RadianceCacheConfig)]
    lib.hdcx_rc_get_params.argtypes = [c_void_p, c_void_p]
    lib.hdcx_rc_set_params.argtypes = [c_void_p, c_void_p]
    lib.hdcx_rc_infer.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p]
    lib.hdcx_rc_gradient.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p, POINTER(c_float)]
    lib.hdcx_rc_train.argtypes = [c_void_p, c_void_p, c_uint32, c_uint32, POINTER(c_float)]
    lib.hdcx_rc_gradient_noalloc.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p]
    lib.hdcx_rc_get_stats.argtypes = [c_void_p, POINTER(RcStats)]


def _tracer_set_radiance_cache(self, enabled: bool = True, train_samples: int = 65536, terminate_bounce: int = 1,
                               train_steps: int = 4, learning_rate: float = 1e-2, seed: int = 0,
                               ema_decay: float = 0.99, spread_threshold: float = 0.0,
                               calibration_decay: float = 0.0, plain_training_paths: bool = False) -> None:
    c = RadianceCacheConfig(int(enabled), train_samples, terminate_bounce, train_steps, learning_rate, 0.9, 0.99,
                            1e-8, ema_decay, spread_threshold, seed, calibration_decay,
                            RC_PLAIN_TRAINING_PATHS if plain_training_paths else 0)
    _check(library().hdcx_set_radiance_cache(self._h, ctypes.byref(c)))


def _tracer_rc_params(self) -> np.ndarray:
    out = np.empty(RC_PARAMS, dtype=np.float32)
    _check(library().hdcx_rc_get_params(self._h, out.ctypes.data))
    return out


def _tracer_rc_set_params(self, params: np.ndarray) -> None:
    p = np.ascontiguousarray(params, dtype=np.float32)
    assert p.size == RC_PARAMS
    _check(library().hdcx_rc_set_params(self._h, p.ctypes.data))


def _tracer_rc_infer(self, queries: np.ndarray) -> np.ndarray:
    q = np.ascontiguousarray(queries, dtype=RC_QUERY_DTYPE)
    out = np.empty((len(q), 3), dtype=np.float32)
    _check(library().hdcx_rc_infer(self._h, q.ctypes.data, len(q), out.ctypes.data))
    return out


def _tracer_rc_gradient(self, samples: np.ndarray):
    q = np.ascontiguousarray(samples, dtype=RC_QUERY_DTYPE)
    g = np.empty(RC_PARAMS, dtype=np.float32)
    loss = c_float()
    _check(library().hdcx_rc_gradient(self._h, q.ctypes.data, len(q), g.ctypes.data, ctypes.byref(loss)))
    return float(loss.value), g


def _tracer_rc_gradient_noalloc(self, samples_ptr: int, count: int, grad_ptr: int, loss_ptr: int = 0) -> None:
    """hdcx_rc_gradient_noalloc on device pointers (samples: count x 48
    bytes, grad: RC_PARAMS floats, loss: one float or 0)."""
    _check(library().hdcx_rc_gradient_noalloc(self._h, samples_ptr, count, grad_ptr, loss_ptr or None))


def _tracer_rc_train(self, samples: np.ndarray, steps: int = 1) -> float:
    q = np.ascontiguousarray(samples, dtype=RC_QUERY_DTYPE)
    loss = c_float()
    _check(library().hdcx_rc_train(self._h, q.ctypes.data, len(q), steps, ctypes.byref(loss)))
    return float(loss.value)


def _tracer_rc_stats(self) -> dict:
    s = RcStats()
    _check(library().hdcx_rc_get_stats(self._h, ctypes.byref(s)))
    return {name: getattr(s, name) for name, _ in s._fields_}


PathTracer.set_radiance_cache = _tracer_set_radiance_cache
PathTracer.rc_params = _tracer_rc_params
PathTracer.rc_set_params = _tracer_rc_set_params
PathTracer.rc_infer = _tracer_rc_infer
PathTracer.rc_gradient = _tracer_rc_gradient
PathTracer.rc_train = _tracer_rc_train
PathTracer.rc_gradient_noalloc = _tracer_rc_gradient_noalloc
PathTracer.rc_stats = _tracer_rc_stats
EXPORTED_SYMBOLS = EXPORTED_SYMBOLS + (
    "hdcx_set_radiance_cache", "hdcx_rc_get_params", "hdcx_rc_set_params", "hdcx_rc_infer", "hdcx_rc_gradient",
    "hdcx_rc_train", "hdcx_rc_get_stats", "hdcx_rc_gradient_noalloc")


# ---- BVH builders and the tracer's BVH policy ----------------------------

BVH_HOST_SAH, BVH_DEVICE_LBVH, BVH_SERIAL_LBVH = 0, 1, 2


class BvhOptions(ctypes.Structure):
    _fields_ = [("builder", ctypes.c_int32), ("top_clusters", c_uint32), ("refit", c_uint32), ("reserved", c_uint32)]


def _bind_bvh(lib: ctypes.CDLL) -> None:
    lib.hdcx_bvh_build.argtypes = [POINTER(SceneView), c_int, c_uint32, POINTER(BvhInfo), c_void_p, c_void_p]
    lib.hdcx_set_bvh_options.argtypes = [c_void_p, POINTER(BvhOptions)]
    lib.hdcx_get_bvh_info.argtypes = [c_void_p, POINTER(BvhInfo)]


def build_bvh(scene: "ProceduralScene", builder: int = BVH_HOST_SAH, top_clusters: int = 2048):
    """BVH4 of a scene from the chosen builder, validated on the host:
    (info dict, nodes, slots[(n, 3, 4) float32]). BVH_DEVICE_LBVH needs a GPU."""
    lib = library()
    info = BvhInfo()
    _check(lib.hdcx_bvh_build(ctypes.byref(scene.view), builder, top_clusters, ctypes.byref(info), None, None))
    nodes = np.zeros(info.node_count, dtype=BVH_NODE_DTYPE)
    slots = np.zeros((info.slot_count, 3, 4), dtype=np.float32)
    _check(lib.hdcx_bvh_build(ctypes.byref(scene.view), builder, top_clusters, ctypes.byref(info), nodes.ctypes.data,
                              slots.ctypes.data))
    return info.as_dict(), nodes, slots


def _tracer_set_bvh_options(self, builder: int = BVH_DEVICE_LBVH, top_clusters: int = 2048, refit: bool = True) -> None:
    o = BvhOptions(builder, top_clusters, 1 if refit else 0, 0)
    _check(library().hdcx_set_bvh_options(self._h, ctypes.byref(o)))


def _tracer_bvh_info(self) -> dict:
    info = BvhInfo()
    _check(library().hdcx_get_bvh_info(self._h, ctypes.byref(info)))
    return info.as_dict()


PathTracer.set_bvh_options = _tracer_set_bvh_options
PathTracer.bvh_info = _tracer_bvh_info
EXPORTED_SYMBOLS = EXPORTED_SYMBOLS + ("hdcx_bvh_build", "hdcx_set_bvh_options", "hdcx_get_bvh_info")


# ---- Batched multi-scene renderer -----------------------------------------

class BatchTuning(ctypes.Structure):
    _fields_ = [("wave_paths", c_uint32), ("tile_pixels", c_uint32), ("max_blocks", c_uint32), ("use_graph", c_uint32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [("pixel_samples", c_uint64), ("lane_paths", c_uint64), ("waves", c_uint32),
                ("graph_launches", c_uint32), ("milliseconds", c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchRcStats(ctypes.Structure):
    _fields_ = [("queries", c_uint64), ("train_samples", c_uint64), ("adam_steps", c_uint64), ("train_ms", c_double),
                ("mean_loss", c_float), ("scenes_trained", c_uint32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def _bind_batch(lib: ctypes.CDLL) -> None:
    lib.hdcx_batch_create.argtypes = [c_int, POINTER(c_void_p)]
    lib.hdcx_batch_destroy.argtypes = [c_void_p]
    lib.hdcx_batch_add_scene.argtypes = [c_void_p, POINTER(SceneView), POINTER(BvhOptions), POINTER(c_uint32)]
    lib.hdcx_batch_set_view.argtypes = [c_void_p, c_uint32, POINTER(Camera), c_uint32, c_uint32, c_uint32]
    lib.hdcx_batch_set_tuning.argtypes = [c_void_p, POINTER(BatchTuning)]
    lib.hdcx_batch_get_tuning.argtypes = [c_void_p, POINTER(BatchTuning)]
    lib.hdcx_batch_autotune.argtypes = [c_void_p, c_uint32, c_char_p, POINTER(BatchTuning)]
    lib.hdcx_batch_render.argtypes = [c_void_p, c_uint32, c_uint32]
    lib.hdcx_batch_read_film.argtypes = [c_void_p, c_uint32, c_void_p]
    lib.hdcx_batch_get_stats.argtypes = [c_void_p, POINTER(BatchStats)]
    lib.hdcx_batch_set_radiance_cache.argtypes = [c_void_p, POINTER(RadianceCacheConfig)]
    lib.hdcx_batch_get_rc_stats.argtypes = [c_void_p, POINTER(BatchRcStats)]


class BatchRenderer:
    """Many scenes rendered together (hdcx_batch_*): one wavefront over all
    scenes, per-wave launches replayed from a captured hipGraph."""

    def __init__(self, device: int = 0):
        self._h = c_void_p()
        _check(library().hdcx_batch_create(device, ctypes.byref(self._h)))
        self._sizes = []

    def add_scene(self, scene: "ProceduralScene | SceneView", builder: int = BVH_DEVICE_LBVH,
                  top_clusters: int = 2048) -> int:
        view = scene.view if isinstance(scene, ProceduralScene) else scene
        o = BvhOptions(builder, top_clusters, 1, 0)
        index = c_uint32()
        _check(library().hdcx_batch_add_scene(self._h, ctypes.byref(view), ctypes.byref(o), ctypes.byref(index)))
        self._sizes.append((0, 0))
        return int(index.value)

    def set_view(self, scene: int, camera: Camera, width: int, height: int, samples: int) -> None:
        _check(library().hdcx_batch_set_view(self._h, scene, ctypes.byref(camera), width, height, samples))
        self._sizes[scene] = (width, height)

    def set_tuning(self, wave_paths: int, tile_pixels: int, max_blocks: int, use_graph: bool = True) -> None:
        t = BatchTuning(wave_paths, tile_pixels, max_blocks, 1 if use_graph else 0)
        _check(library().hdcx_batch_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self) -> dict:
        t = BatchTuning()
        _check(library().hdcx_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, max_bounces: int = 8, cache_path: "str | None" = None) -> dict:
        t = BatchTuning()
        path = cache_path.encode() if cache_path else None
        _check(library().hdcx_batch_autotune(self._h, max_bounces, path, ctypes.byref(t)))
        return t.as_dict()

    def render(self, sample_index: int = 0, max_bounces: int = 8) -> None:
        _check(library().hdcx_batch_render(self._h, sample_index, max_bounces))

    def read_film(self, scene: int) -> np.ndarray:
        w, h = self._sizes[scene]
        out = np.empty((h, w, 4), dtype=np.float32)
        _check(library().hdcx_batch_read_film(self._h, scene, out.ctypes.data))
        return out

    def stats(self) -> dict:
        s = BatchStats()
        _check(library().hdcx_batch_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def set_radiance_cache(self, enabled: bool = True, train_samples: int = 65536, terminate_bounce: int = 1,
                           train_steps: int = 4, learning_rate: float = 1e-2, seed: int = 0,
                           ema_decay: float = 0.99, spread_threshold: float = 0.0) -> None:
        """One cache per scene (hdcx_batch_set_radiance_cache); train_samples
        per scene and render."""
        c = RadianceCacheConfig(int(enabled), train_samples, terminate_bounce, train_steps, learning_rate, 0.9, 0.99,
                                1e-8, ema_decay, spread_threshold, seed, 0.0, 0)
        _check(library().hdcx_batch_set_radiance_cache(self._h, ctypes.byref(c)))

    def rc_stats(self) -> dict:
        s = BatchRcStats()
        _check(library().hdcx_batch_get_rc_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def close(self) -> None:
        if self._h:
            _check(library().hdcx_batch_destroy(self._h))
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


EXPORTED_SYMBOLS = EXPORTED_SYMBOLS + (
    "hdcx_batch_create", "hdcx_batch_destroy", "hdcx_batch_add_scene", "hdcx_batch_set_view", "hdcx_batch_set_tuning",
    "hdcx_batch_get_tuning", "hdcx_batch_autotune", "hdcx_batch_render", "hdcx_batch_read_film",
    "hdcx_batch_get_stats", "hdcx_batch_set_radiance_cache", "hdcx_batch_get_rc_stats")
