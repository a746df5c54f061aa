"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/build/liboracle.so, the CPU restatement of CudaRT's
per-pixel path tracer (see oracle/oracle.cpp). Only the test suite, the smoke
check and the benchmark's CPU-baseline leg import this module; the product
package never does.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from cudart_amd._capi import Camera, SceneDesc

# PRT_ORACLE_LIB names another build of the same sources (the mutation tests load
# copies of the oracle with one planted mistake each through it).
LIB_PATH = Path(os.environ.get("PRT_ORACLE_LIB") or Path(__file__).resolve().parent / "build" / "liboracle.so")

_lib = None
FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int32)
UP = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
BP = C.POINTER(C.c_uint8)


class Cfg(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("rr_start", C.c_uint32), ("seed", C.c_uint32), ("view_id", C.c_uint32)]


def cfg(max_depth=6, rr_start=3, seed=0x5EED, view_id=0) -> Cfg:
    return Cfg(max_depth, rr_start, seed, view_id)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: build with `python -m cudart_amd.build`")
        l = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        sig = {
            "oracle_scene_create": [C.POINTER(SceneDesc), C.POINTER(P)],
            "oracle_scene_destroy": [P],
            "oracle_scene_info": [P, UP, UP, UP, FP],
            "oracle_scene_export_bvh": [P, FP, IP, UP],
            "oracle_render_f32": [P, C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.POINTER(Cfg), FP, U64P, C.c_int],
            "oracle_render_f64": [P, C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.POINTER(Cfg), DP, U64P, C.c_int],
            "oracle_render_pixels_f32": [P, C.POINTER(Camera), C.c_uint32, C.c_uint32, UP, C.c_uint32, C.c_uint32,
                                         C.c_uint32, C.POINTER(Cfg), FP, U64P, C.c_int],
            "oracle_render_pixels_f64": [P, C.POINTER(Camera), C.c_uint32, C.c_uint32, UP, C.c_uint32, C.c_uint32,
                                         C.c_uint32, C.POINTER(Cfg), DP, U64P, C.c_int],
            "oracle_first_hit": [P, C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Cfg), IP, IP,
                                 FP, C.c_int, C.c_int],
            "oracle_trace": [P, C.c_uint32, FP, FP, IP, FP, FP, FP, C.c_int, C.c_int],
            "oracle_philox": [UP, UP, UP],
            "oracle_sincosf": [C.c_uint32, FP, FP, FP],
            "oracle_triangle_intersect": [FP, FP, FP, FP, FP, FP, FP, FP],
            "oracle_box_hit": [FP, FP, FP, FP],
            "oracle_fresnel": [FP, FP, C.c_float, FP, FP, FP],
            "oracle_diffuse_dir": [FP, C.c_float, C.c_float, FP, FP],
            "oracle_camera_ray": [C.POINTER(Camera), C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float,
                                  C.c_float, FP, FP],
            "oracle_tonemap": [C.c_uint32, FP, BP],
            "oracle_texture_table": [FP],
            "oracle_texture_sample": [C.c_void_p, C.c_uint32, DP, DP, C.c_int],
            "oracle_adjoint": [P, C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Cfg), FP, FP, DP,
                               DP, C.c_double, C.c_int, C.c_int, C.c_int],
            "oracle_scene_set_material_params": [P, C.c_uint32, FP, C.c_float],
        }
        for name, args in sig.items():
            fn = getattr(l, name)
            fn.restype = C.c_int
            fn.argtypes = args
        _lib = l
    return _lib


def _chk(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def _f(a):
    return a.ctypes.data_as(FP)


class OracleScene:
    def __init__(self, desc: SceneDesc):
        h = C.c_void_p()
        _chk(lib().oracle_scene_create(C.byref(desc), C.byref(h)), "oracle_scene_create")
        self._h = h
        self.num_triangles = desc.num_triangles
        self.num_materials = desc.num_materials

    def info(self):
        nn, md, nl = C.c_uint32(), C.c_uint32(), C.c_uint32()
        la = C.c_float()
        _chk(lib().oracle_scene_info(self._h, C.byref(nn), C.byref(md), C.byref(nl), C.byref(la)), "info")
        return dict(num_nodes=nn.value, max_depth=md.value, num_lights=nl.value, lights_area=la.value)

    def export_bvh(self):
        n = self.info()["num_nodes"]
        b = np.zeros((n, 6), np.float32)
        m = np.zeros((n, 3), np.int32)
        o = np.zeros(self.num_triangles, np.uint32)
        _chk(lib().oracle_scene_export_bvh(self._h, _f(b), m.ctypes.data_as(IP), o.ctypes.data_as(UP)), "export")
        return b, m, o

    def render(self, cam: Camera, w: int, h: int, s0: int, n: int, c: Cfg, accum=None, threads: int = 0,
               dtype=np.float32):
        if accum is None:
            accum = np.zeros((h, w, 3), dtype)
        cnt = np.zeros(3, np.uint64)
        if accum.dtype == np.float32:
            rc = lib().oracle_render_f32(self._h, C.byref(cam), w, h, s0, n, C.byref(c), _f(accum),
                                         cnt.ctypes.data_as(U64P), threads)
        else:
            rc = lib().oracle_render_f64(self._h, C.byref(cam), w, h, s0, n, C.byref(c), accum.ctypes.data_as(DP),
                                         cnt.ctypes.data_as(U64P), threads)
        _chk(rc, "oracle_render")
        return accum, cnt

    def render_pixels(self, cam: Camera, w: int, h: int, pixels, s0: int, n: int, c: Cfg, threads: int = 0,
                      dtype=np.float32):
        """Sum of samples s0..s0+n-1 of each listed pixel (index y*w+x), in fp32 or fp64."""
        px = np.ascontiguousarray(pixels, np.uint32)
        out = np.zeros((len(px), 3), dtype)
        cnt = np.zeros(3, np.uint64)
        if out.dtype == np.float32:
            rc = lib().oracle_render_pixels_f32(self._h, C.byref(cam), w, h, px.ctypes.data_as(UP), len(px), s0, n,
                                                C.byref(c), _f(out), cnt.ctypes.data_as(U64P), threads)
        else:
            rc = lib().oracle_render_pixels_f64(self._h, C.byref(cam), w, h, px.ctypes.data_as(UP), len(px), s0, n,
                                                C.byref(c), out.ctypes.data_as(DP), cnt.ctypes.data_as(U64P), threads)
        _chk(rc, "render_pixels")
        return out, cnt

    def first_hit(self, cam: Camera, w: int, h: int, sample: int, c: Cfg, brute: bool = False, threads: int = 0):
        tri = np.zeros((h, w), np.int32)
        mat = np.zeros((h, w), np.int32)
        t = np.zeros((h, w), np.float32)
        _chk(lib().oracle_first_hit(self._h, C.byref(cam), w, h, sample, C.byref(c), tri.ctypes.data_as(IP),
                                    mat.ctypes.data_as(IP), _f(t), int(brute), threads), "first_hit")
        return tri, mat, t

    def trace(self, orig, direc, brute: bool = False, threads: int = 0):
        o = np.ascontiguousarray(orig, np.float32)
        d = np.ascontiguousarray(direc, np.float32)
        n = len(o)
        tri = np.zeros(n, np.int32)
        t, u, v = (np.zeros(n, np.float32) for _ in range(3))
        _chk(lib().oracle_trace(self._h, n, _f(o), _f(d), tri.ctypes.data_as(IP), _f(t), _f(u), _f(v), int(brute),
                                threads), "trace")
        return tri, t, u, v

    def adjoint(self, cam: Camera, w: int, h: int, spp: int, c: Cfg, image, target, fp64: bool = False,
                threads: int = 0, pathwise: bool = False, hess: float | None = None):
        """hess None: the loss's own curvature 2/(w*h*3)."""
        if hess is None:
            hess = 2.0 / (w * h * 3)
        img = np.ascontiguousarray(image, np.float32)
        tgt = np.ascontiguousarray(target, np.float32)
        da = np.zeros((self.num_materials, 3), np.float64)
        de = np.zeros(self.num_materials, np.float64)
        _chk(lib().oracle_adjoint(self._h, C.byref(cam), w, h, spp, C.byref(c), _f(img), _f(tgt),
                                  da.ctypes.data_as(DP), de.ctypes.data_as(DP), float(hess), int(pathwise),
                                  int(fp64), threads),
             "adjoint")
        return da, de

    def set_material_params(self, index: int, kd, emission: float) -> None:
        k = np.asarray(kd, np.float32)
        _chk(lib().oracle_scene_set_material_params(self._h, index, _f(k), float(emission)), "set_material")

    def close(self):
        if self._h:
            lib().oracle_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    _chk(lib().oracle_philox(c.ctypes.data_as(UP), k.ctypes.data_as(UP), o.ctypes.data_as(UP)), "philox")
    return o


def sincosf(x):
    x = np.ascontiguousarray(x, np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    _chk(lib().oracle_sincosf(x.size, _f(x), _f(s), _f(c)), "sincosf")
    return s, c


def triangle_intersect(v1, e1, e2, o, d):
    a = [np.asarray(x, np.float32) for x in (v1, e1, e2, o, d)]
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    hit = lib().oracle_triangle_intersect(*[_f(x) for x in a], C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


def box_hit(lo, hi, o, d) -> bool:
    a = [np.asarray(x, np.float32) for x in (lo, hi, o, d)]
    return bool(lib().oracle_box_hit(*[_f(x) for x in a]))


def fresnel(wo, n, ior):
    a = np.asarray(wo, np.float32)
    b = np.asarray(n, np.float32)
    re, tr = C.c_float(), C.c_float()
    td = np.zeros(3, np.float32)
    _chk(lib().oracle_fresnel(_f(a), _f(b), float(ior), C.byref(re), C.byref(tr), _f(td)), "fresnel")
    return re.value, tr.value, td


def diffuse_dir(n, u1, u2):
    a = np.asarray(n, np.float32)
    d = np.zeros(3, np.float32)
    cw = C.c_float()
    _chk(lib().oracle_diffuse_dir(_f(a), float(u1), float(u2), _f(d), C.byref(cw)), "diffuse_dir")
    return d, cw.value


def camera_ray(cam: Camera, x, y, u0, u1, u2=0.0, u3=0.0):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    _chk(lib().oracle_camera_ray(C.byref(cam), x, y, u0, u1, u2, u3, _f(o), _f(d)), "camera_ray")
    return o, d


def tonemap(rgb):
    a = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
    out = np.zeros((len(a), 4), np.uint8)
    _chk(lib().oracle_tonemap(len(a), _f(a), out.ctypes.data_as(BP)), "tonemap")
    return out


def texture_table():
    t = np.zeros(256, np.float32)
    _chk(lib().oracle_texture_table(_f(t)), "texture_table")
    return t


def texture_sample(rgba, uv, fp64: bool = True):
    """Bilinear, wrapped lookup of an (H, W, 4) uint8 texture at (u, v) pairs -> (n, 3)."""
    from cudart_amd._capi import Texture

    img = np.ascontiguousarray(rgba, np.uint8)
    t = Texture()
    t.width, t.height = img.shape[1], img.shape[0]
    t.rgba = img.ctypes.data_as(BP)
    q = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
    out = np.zeros((len(q), 3), np.float64)
    _chk(lib().oracle_texture_sample(C.byref(t), len(q), q.ctypes.data_as(DP), out.ctypes.data_as(DP), int(fp64)),
         "texture_sample")
    return out
