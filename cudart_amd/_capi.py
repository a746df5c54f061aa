"""ctypes declarations of the C ABI in include/prt.h.

The library is loaded from cudart_amd/lib/libprt.so (built in-tree). There is no
fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

# PRT_LIBRARY selects another build of the same library (e.g. a tuning variant).
LIB_PATH = Path(os.environ.get("PRT_LIBRARY") or Path(__file__).resolve().parent / "lib" / "libprt.so")

PRT_OK = 0
ERRORS = {
    -1: "invalid argument", -2: "out of memory", -3: "HIP error", -4: "no emitters", -5: "unsupported",
    -6: "I/O error", -7: "parse error", -8: "no device", -9: "limit exceeded",
}

DIFFUSE, SPECULAR, DIFFSPEC, MICROFACET, REFRACTIVE, EMISSIVE = range(6)
INTEGRATOR_MEGAKERNEL, INTEGRATOR_WAVEFRONT = 0, 1
NO_TEXTURE = -1


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("dir", C.c_float * 3), ("up", C.c_float * 3), ("right", C.c_float * 3),
                ("fov", C.c_float), ("aspect", C.c_float), ("aperture_width", C.c_float),
                ("focus_distance", C.c_float), ("xpixels", C.c_int32), ("ypixels", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("bsdf", C.c_int32), ("kd", C.c_float * 3), ("ka", C.c_float * 3), ("ks", C.c_float * 3),
                ("ns", C.c_float), ("ni", C.c_float), ("diffuse_coef", C.c_float), ("tex_kd", C.c_int32)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.POINTER(C.c_uint8))]


class SceneDesc(C.Structure):
    _fields_ = [("num_triangles", C.c_uint32), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
                ("material_ids", C.POINTER(C.c_uint32)), ("num_materials", C.c_uint32),
                ("materials", C.POINTER(Material)), ("num_textures", C.c_uint32),
                ("textures", C.POINTER(Texture)), ("camera", Camera)]


class RenderConfig(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("rr_start", C.c_uint32), ("seed", C.c_uint32), ("view_id", C.c_uint32),
                ("device", C.c_int32), ("integrator", C.c_uint32), ("spp_per_launch", C.c_uint32),
                ("count_rays", C.c_uint32), ("wavefront_group", C.c_uint32), ("wavefront_queue", C.c_uint32),
                ("profile_kernels", C.c_uint32), ("wavefront_blocks_per_cu", C.c_uint32),
                ("keep_samples", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("nodes_visited", C.c_uint64), ("tris_tested", C.c_uint64), ("hits", C.c_uint64),
                ("samples", C.c_uint64), ("last_render_ms", C.c_double), ("anyhit_nodes", C.c_uint64),
                ("anyhit_tris", C.c_uint64), ("kernel_ms", C.c_double * 3), ("kernel_launches", C.c_uint32 * 3)]


class BvhInfo(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("num_leaves", C.c_uint32), ("max_depth", C.c_uint32),
                ("max_leaf_size", C.c_uint32), ("node_bytes", C.c_uint64), ("triangle_bytes", C.c_uint64),
                ("shading_bytes", C.c_uint64), ("sah_cost", C.c_float), ("node_read_bytes", C.c_uint32)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("spp", C.c_uint32), ("view_id", C.c_uint32),
                ("out_rgb", C.POINTER(C.c_float))]


class Tune(C.Structure):
    _fields_ = [("queue_slots", C.c_uint32), ("group_side", C.c_uint32), ("block_size", C.c_uint32),
                ("blocks_per_cu", C.c_uint32)]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int32)
UP = C.POINTER(C.c_uint32)
BP = C.POINTER(C.c_uint8)
I = C.c_int
U = C.c_uint32

# name: (restype, argtypes)
SIGNATURES = {
    "prt_last_error": (C.c_char_p, []),
    "prt_abi_version": (I, []),
    "prt_device_count": (I, [C.POINTER(C.c_int)]),
    "prt_mesh_load_obj": (I, [C.c_char_p, PP]),
    "prt_mesh_generate": (I, [C.c_char_p, C.c_char_p, U, PP]),
    "prt_mesh_get_desc": (I, [P, C.POINTER(SceneDesc)]),
    "prt_mesh_set_material": (I, [P, U, C.POINTER(Material)]),
    "prt_mesh_set_camera": (I, [P, C.POINTER(Camera)]),
    "prt_mesh_material_name": (I, [P, U, C.c_char_p, C.c_size_t]),
    "prt_mesh_destroy": (I, [P]),
    "prt_camera_from_json": (I, [C.c_char_p, U, U, C.POINTER(Camera)]),
    "prt_camera_look_at": (I, [FP, FP, FP, C.c_float, U, U, C.POINTER(Camera)]),
    "prt_env_load": (I, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]),
    "prt_scene_create": (I, [C.POINTER(SceneDesc), PP]),
    "prt_scene_bvh_info": (I, [P, C.POINTER(BvhInfo)]),
    "prt_scene_export_bvh": (I, [P, FP, FP, UP]),
    "prt_scene_num_lights": (I, [P, UP, FP]),
    "prt_scene_destroy": (I, [P]),
    "prt_scene_set_material_params": (I, [P, U, FP, C.c_float]),
    "prt_render_config_default": (None, [C.POINTER(RenderConfig)]),
    "prt_renderer_create": (I, [P, C.POINTER(Camera), U, U, U, C.POINTER(RenderConfig), PP]),
    "prt_render_spp": (I, [P, U]),
    "prt_renderer_samples_rendered": (I, [P, UP]),
    "prt_renderer_reset": (I, [P]),
    "prt_get_radiance": (I, [P, FP]),
    "prt_get_radiance_device": (I, [P, P]),
    "prt_copy_image_bytes": (I, [P, BP]),
    "prt_first_hit": (I, [P, U, IP, IP, FP]),
    "prt_get_stats": (I, [P, C.POINTER(Stats)]),
    "prt_reset_stats": (I, [P]),
    "prt_renderer_destroy": (I, [P]),
    "prt_render_batch": (I, [C.POINTER(Job), U, C.POINTER(RenderConfig), C.POINTER(Tune), C.POINTER(Stats)]),
    "prt_autotune": (I, [C.POINTER(Job), U, C.POINTER(RenderConfig), C.POINTER(Tune), C.POINTER(C.c_double)]),
    "prt_adjoint": (I, [P, FP, FP, U, FP, FP, FP]),
    "prt_adjoint_grad": (I, [P, FP, C.c_float, U, FP, FP]),
    "prt_adjoint_grad_static": (I, [P, C.c_void_p, C.c_float, U, FP, FP]),
    "prt_write_png": (I, [C.c_char_p, U, U, BP]),
    "prt_write_pfm": (I, [C.c_char_p, U, U, FP]),
}

_lib = None


class PrtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, 'error')} ({code}): {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Loads libprt.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cudart_amd.build` "
                               "(or __graft_entry__.build()); there is no fallback path")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(rc: int) -> None:
    if rc != PRT_OK:
        raise PrtError(rc, (lib().prt_last_error() or b"").decode())


def fptr(a):
    return a.ctypes.data_as(FP)


def iptr(a):
    return a.ctypes.data_as(IP)


def uptr(a):
    return a.ctypes.data_as(UP)


def bptr(a):
    return a.ctypes.data_as(BP)
