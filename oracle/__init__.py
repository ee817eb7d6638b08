"""ORACLE — CPU restatement of the reference integrator (test infrastructure).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package, and only as the checker (or, for the baseline, as
the timed CPU port). The product library never links or calls it.

``liboracle.so`` is built from ``oracle/oracle.cpp`` by
``python -m hdcodex_amd.build --oracle``.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
# -O3 builds of the same source per x86-64 level, highest first, with the
# CPU flags each level needs (the CPU baseline loads the best one the host
# runs; results equal the default build's: contraction is off in all).
LEVELS = (("x86-64-v4", ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")),
          ("x86-64-v3", ("avx2", "fma", "bmi2", "movbe")))
_lib = None
_lib_path = LIB_PATH
BUILD_USED = "g++ -O2 -ffp-contract=off"


def host_cpu() -> dict:
    """Model name, logical CPUs and flags of this host (from /proc/cpuinfo)."""
    model, flags, count = "unknown", set(), 0
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            key, _, value = line.partition(":")
            key = key.strip()
            if key == "processor":
                count += 1
            elif key == "model name" and model == "unknown":
                model = value.strip()
            elif key == "flags" and not flags:
                flags = set(value.split())
    except OSError:
        pass
    return {"model": model, "logical_cpus": count, "flags": flags}


def use_fastest_build() -> str:
    """Switches this process to the -O3 build of the highest x86-64 level the
    host supports (call before the first OracleScene); returns its name."""
    global _lib, _lib_path, BUILD_USED
    flags = host_cpu()["flags"]
    for level, needed in LEVELS:
        path = LIB_PATH.parent / f"liboracle_{level}.so"
        if path.exists() and all(f in flags for f in needed):
            if _lib is not None and _lib_path != path:
                raise RuntimeError("the oracle library is already loaded")
            _lib_path = path
            BUILD_USED = f"g++ -O3 -march={level} -ffp-contract=off"
            break
    return BUILD_USED


def library() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not _lib_path.exists():
        raise RuntimeError(f"{_lib_path} missing: run `python -m hdcodex_amd.build --oracle`")
    lib = ctypes.CDLL(str(_lib_path))
    lib.oracle_last_error.restype = ctypes.c_char_p
    lib.oracle_scene_create.argtypes = [c_void_p, POINTER(c_void_p)]
    lib.oracle_scene_destroy.argtypes = [c_void_p]
    lib.oracle_render.argtypes = [c_void_p, c_void_p, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, c_int,
                                  c_void_p, c_uint32, c_uint32, c_int, c_void_p]
    lib.oracle_trace_closest.argtypes = [c_void_p, c_void_p, c_uint32, c_int, c_void_p, c_void_p]
    lib.oracle_trace_occluded.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p]
    lib.oracle_hash.argtypes = [c_uint32]
    lib.oracle_hash.restype = c_uint32
    lib.oracle_random_float.argtypes = [POINTER(c_uint32)]
    lib.oracle_random_float.restype = c_float
    for name in ("oracle_cie_xyz", "oracle_sensor"):
        getattr(lib, name).argtypes = [c_double, c_void_p]
    lib.oracle_reconstruct_rgb.argtypes = [c_void_p, c_double]
    lib.oracle_reconstruct_rgb.restype = c_double
    lib.oracle_illuminant.argtypes = [c_void_p, c_double]
    lib.oracle_illuminant.restype = c_double
    lib.oracle_dispersed_ior.argtypes = [c_double, c_double, c_double, c_double]
    lib.oracle_dispersed_ior.restype = c_double
    lib.oracle_eval_diffuse.argtypes = [c_void_p, c_void_p, c_double, c_double, c_uint32]
    lib.oracle_eval_diffuse.restype = c_double
    lib.oracle_ggx_pdf.argtypes = [c_void_p, c_void_p, c_double, c_double]
    lib.oracle_ggx_pdf.restype = c_double
    lib.oracle_ggx_eval.argtypes = [c_void_p, c_void_p, c_double, c_double, c_double]
    lib.oracle_ggx_eval.restype = c_double
    lib.oracle_ggx_sample.argtypes = [c_void_p, c_double, c_double, c_double, POINTER(c_uint32), c_void_p]
    lib.oracle_ggx_sample.restype = c_double
    lib.oracle_sheen_albedo.argtypes = [c_double, c_double, c_int]
    lib.oracle_sheen_albedo.restype = c_double
    lib.oracle_sheen_response.argtypes = [c_void_p, c_void_p, c_double, c_int]
    lib.oracle_sheen_response.restype = c_double
    lib.oracle_eon_albedo.argtypes = [c_double, c_double]
    lib.oracle_eon_albedo.restype = c_double
    lib.oracle_sample_hg.argtypes = [c_void_p, c_double, POINTER(c_uint32), c_void_p]
    lib.oracle_cosine_hemisphere.argtypes = [c_void_p, POINTER(c_uint32), c_void_p]
    lib.oracle_dome_uv.argtypes = [c_void_p, c_uint32, c_void_p]
    lib.oracle_lat_long_direction.argtypes = [c_double, c_double, c_void_p]
    lib.oracle_sample_dome.argtypes = [c_void_p, c_uint32, POINTER(c_uint32), c_void_p, c_void_p]
    lib.oracle_sample_dome.restype = c_double
    lib.oracle_transmission.argtypes = [c_void_p, c_void_p, c_int, c_double, c_int, POINTER(c_uint32), c_void_p]
    lib.oracle_transmission.restype = c_int
    lib.oracle_dome_pdf.argtypes = [c_void_p, c_uint32, c_void_p]
    lib.oracle_dome_pdf.restype = c_double
    lib.oracle_dome_radiance.argtypes = [c_void_p, c_uint32, c_void_p, c_void_p]
    lib.oracle_texture.argtypes = [c_void_p, c_uint32, c_double, c_double, c_void_p]
    lib.oracle_material_texture.argtypes = [c_void_p, c_uint32, c_double, c_double, c_void_p, c_void_p]
    lib.oracle_normal_map.argtypes = [c_void_p, c_uint32, c_int, c_double, c_double, c_double, c_void_p, c_void_p,
                                      c_void_p, c_void_p]
    lib.oracle_random_walk.argtypes = [c_void_p, c_void_p, c_void_p, c_double, c_double, c_double, POINTER(c_uint32),
                                       c_void_p, c_void_p, c_void_p]
    lib.oracle_random_walk.restype = c_int
    lib.oracle_emissive_table.argtypes = [c_void_p, c_void_p]
    lib.oracle_upscale_nearest.argtypes = [c_void_p, c_uint32, c_uint32, c_uint32, c_uint32, c_void_p]
    lib.oracle_dome_importance.argtypes = [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_uint32), c_void_p,
                                           c_void_p, c_void_p]
    _lib = lib
    return lib


def _vec(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))


class OracleScene:
    """The oracle's copy of a packed scene (takes an hdcx SceneView)."""

    def __init__(self, view):
        self._h = c_void_p()
        self.triangle_count = int(view.triangle_count)
        if library().oracle_scene_create(ctypes.byref(view), ctypes.byref(self._h)) != 0:
            raise RuntimeError(library().oracle_last_error().decode())

    def emissive_table(self) -> np.ndarray:
        """(n, 4) float32 rows (cdf, pdf / area, area, 0): the view's table,
        or the FP64 restatement of the reference's host build."""
        out = np.empty((self.triangle_count, 4), dtype=np.float32)
        library().oracle_emissive_table(self._h, out.ctypes.data)
        return out

    def dome_importance(self, light: int):
        """(marginal[h], conditional[h, w], pdf[h, w]) or None (uniform)."""
        w, h = c_uint32(), c_uint32()
        library().oracle_dome_importance(self._h, light, ctypes.byref(w), ctypes.byref(h), None, None, None)
        if w.value == 0:
            return None
        marginal = np.empty(h.value, dtype=np.float32)
        conditional = np.empty((h.value, w.value), dtype=np.float32)
        pdf = np.empty((h.value, w.value), dtype=np.float32)
        library().oracle_dome_importance(self._h, light, ctypes.byref(w), ctypes.byref(h), marginal.ctypes.data,
                                         conditional.ctypes.data, pdf.ctypes.data)
        return marginal, conditional, pdf

    def render(self, camera, width: int, height: int, sample_index: int, max_bounces: int, sample_count: int,
               precision: int = 64, rows=None, col_step: int = 1, threads: int = 8) -> np.ndarray:
        """Per-pixel SUM over the batch's samples of the sample RGB, for the
        given rows (default: all) and every col_step-th column; shape
        (len(rows), ceil(width/col_step), 3), float64."""
        rows = np.arange(height, dtype=np.uint32) if rows is None else np.ascontiguousarray(rows, dtype=np.uint32)
        cols = (width + col_step - 1) // col_step
        out = np.zeros((len(rows), cols, 3), dtype=np.float64)
        code = library().oracle_render(self._h, ctypes.byref(camera), width, height, sample_index, max_bounces,
                                       sample_count, precision, rows.ctypes.data, len(rows), col_step, threads,
                                       out.ctypes.data)
        if code != 0:
            raise RuntimeError(library().oracle_last_error().decode())
        return out

    def trace_closest(self, rays: np.ndarray, brute_force: bool = False):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        hits = np.empty((len(rays), 4), dtype=np.uint32)
        t64 = np.empty(len(rays), dtype=np.float64)
        library().oracle_trace_closest(self._h, rays.ctypes.data, len(rays), int(brute_force), hits.ctypes.data,
                                       t64.ctypes.data)
        out = np.empty(len(rays), dtype=[("t", np.float32), ("primitive", np.uint32), ("u", np.float32),
                                         ("v", np.float32)])
        out["t"] = hits[:, 0].view(np.float32)
        out["primitive"] = hits[:, 1]
        out["u"] = hits[:, 2].view(np.float32)
        out["v"] = hits[:, 3].view(np.float32)
        return out, t64

    def trace_occluded(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        out = np.empty(len(rays), dtype=np.uint8)
        library().oracle_trace_occluded(self._h, rays.ctypes.data, len(rays), out.ctypes.data)
        return out

    def sample_dome(self, light: int, state: int):
        st = c_uint32(state)
        d = np.zeros(3)
        pdf_eval = np.zeros(1)
        pdf = library().oracle_sample_dome(self._h, light, ctypes.byref(st), d.ctypes.data, pdf_eval.ctypes.data)
        return d, pdf, float(pdf_eval[0]), st.value

    def dome_pdf(self, light: int, d) -> float:
        a = _vec(d)
        return float(library().oracle_dome_pdf(self._h, light, a.ctypes.data))

    def dome_radiance(self, light: int, d) -> np.ndarray:
        a = _vec(d)
        out = np.zeros(3)
        library().oracle_dome_radiance(self._h, light, a.ctypes.data, out.ctypes.data)
        return out

    def texture(self, index: int, u: float, v: float) -> np.ndarray:
        out = np.zeros(4)
        library().oracle_texture(self._h, index, u, v, out.ctypes.data)
        return out

    def material_texture(self, slot: int, u: float, v: float, fallback=(0.0, 0.0, 0.0, 0.0)) -> np.ndarray:
        fb = _vec(fallback)
        out = np.zeros(4)
        library().oracle_material_texture(self._h, slot, u, v, fb.ctypes.data, out.ctypes.data)
        return out

    def normal_map(self, texture: int, flip_y: bool, scale: float, u: float, v: float, gn, positions, uvs) -> np.ndarray:
        g, p, t = _vec(gn), _vec(positions), _vec(uvs)
        out = np.zeros(3)
        library().oracle_normal_map(self._h, texture, int(flip_y), scale, u, v, g.ctypes.data, p.ctypes.data,
                                    t.ctypes.data, out.ctypes.data)
        return out

    def random_walk(self, entry, out_n, mfp: float, albedo: float, aniso: float, state: int):
        """(exited, exit origin, exit direction, attenuation, new state)."""
        a, b = _vec(entry), _vec(out_n)
        st = c_uint32(state)
        eo, ed, att = np.zeros(3), np.zeros(3), np.zeros(1)
        ok = library().oracle_random_walk(self._h, a.ctypes.data, b.ctypes.data, mfp, albedo, aniso, ctypes.byref(st),
                                          eo.ctypes.data, ed.ctypes.data, att.ctypes.data)
        return bool(ok), eo, ed, float(att[0]), st.value

    def close(self):
        if self._h:
            library().oracle_scene_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- pure functions (FP64 unless noted) ----
def hash_state(v: int) -> int:
    return int(library().oracle_hash(v & 0xFFFFFFFF))


def random_float(state: int):
    st = c_uint32(state)
    f = library().oracle_random_float(ctypes.byref(st))
    return float(f), st.value


def cie_xyz(w: float) -> np.ndarray:
    out = np.zeros(3)
    library().oracle_cie_xyz(w, out.ctypes.data)
    return out


def sensor(w: float) -> np.ndarray:
    out = np.zeros(3)
    library().oracle_sensor(w, out.ctypes.data)
    return out


def reconstruct_rgb(rgb, w: float) -> float:
    v = _vec(rgb)
    return float(library().oracle_reconstruct_rgb(v.ctypes.data, w))


def illuminant(rgb, w: float) -> float:
    v = _vec(rgb)
    return float(library().oracle_illuminant(v.ctypes.data, w))


def dispersed_ior(ior: float, scale: float, abbe: float, w: float) -> float:
    return float(library().oracle_dispersed_ior(ior, scale, abbe, w))


def eval_diffuse(v, l, color: float, rough: float, model: int) -> float:
    a, b = _vec(v), _vec(l)
    return float(library().oracle_eval_diffuse(a.ctypes.data, b.ctypes.data, color, rough, model))


def ggx_pdf(v, l, rx: float, ry: float) -> float:
    a, b = _vec(v), _vec(l)
    return float(library().oracle_ggx_pdf(a.ctypes.data, b.ctypes.data, rx, ry))


def ggx_eval(v, l, rx: float, ry: float, f0: float) -> float:
    a, b = _vec(v), _vec(l)
    return float(library().oracle_ggx_eval(a.ctypes.data, b.ctypes.data, rx, ry, f0))


def ggx_sample(inc, rx: float, ry: float, f0: float, state: int):
    a = _vec(inc)
    st = c_uint32(state)
    out = np.zeros(3)
    w = library().oracle_ggx_sample(a.ctypes.data, rx, ry, f0, ctypes.byref(st), out.ctypes.data)
    return out, float(w), st.value


def sheen_albedo(nv: float, rough: float, zeltner: bool) -> float:
    return float(library().oracle_sheen_albedo(nv, rough, int(zeltner)))


def sheen_response(v, l, rough: float, zeltner: bool) -> float:
    a, b = _vec(v), _vec(l)
    return float(library().oracle_sheen_response(a.ctypes.data, b.ctypes.data, rough, int(zeltner)))


def eon_albedo(nd: float, rough: float) -> float:
    return float(library().oracle_eon_albedo(nd, rough))


def sample_hg(fwd, g: float, state: int):
    a = _vec(fwd)
    st = c_uint32(state)
    out = np.zeros(3)
    library().oracle_sample_hg(a.ctypes.data, g, ctypes.byref(st), out.ctypes.data)
    return out, st.value


def cosine_hemisphere(n, state: int):
    a = _vec(n)
    st = c_uint32(state)
    out = np.zeros(3)
    library().oracle_cosine_hemisphere(a.ctypes.data, ctypes.byref(st), out.ctypes.data)
    return out, st.value


def dome_uv(d, fmt: int):
    a = _vec(d)
    uv = np.zeros(2)
    library().oracle_dome_uv(a.ctypes.data, fmt, uv.ctypes.data)
    return uv


def transmission(rd, n, front: bool, ior: float, thin: bool, state: int):
    """(crossed, new direction, new state) of a dielectric boundary event."""
    a, b = _vec(rd), _vec(n)
    st = c_uint32(state)
    out = np.zeros(3)
    crossed = library().oracle_transmission(a.ctypes.data, b.ctypes.data, int(front), ior, int(thin), ctypes.byref(st),
                                            out.ctypes.data)
    return bool(crossed), out, st.value


def upscale_nearest(source: np.ndarray, width: int, height: int) -> np.ndarray:
    """UpscaleNearest of an (h0, w0, 4) float32 film to (height, width, 4)."""
    src = np.ascontiguousarray(source, dtype=np.float32)
    out = np.empty((height, width, 4), dtype=np.float32)
    library().oracle_upscale_nearest(src.ctypes.data, src.shape[1], src.shape[0], width, height, out.ctypes.data)
    return out


def lat_long_direction(u: float, v: float) -> np.ndarray:
    out = np.zeros(3)
    library().oracle_lat_long_direction(u, v, out.ctypes.data)
    return out
