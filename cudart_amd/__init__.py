"""MI355X path tracer: host-side mirror of CudaRT's Scene / Renderer interface.

The classes keep the reference's method names (src/include/renderer.h:78-126,
src/include/scene.h:20-81) over the C ABI of libprt.so (include/prt.h). Every
render runs the HIP kernels; nothing here computes pixels on the CPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _capi
from ._capi import (DIFFSPEC, DIFFUSE, EMISSIVE, INTEGRATOR_MEGAKERNEL, INTEGRATOR_WAVEFRONT, MICROFACET,
                    NO_TEXTURE, REFRACTIVE, SPECULAR, Camera, Material, PrtError, RenderConfig, SceneDesc, Stats,
                    Tune, check, lib)

__all__ = [
    "Mesh", "Scene", "ParallelRenderer", "Camera", "Material", "RenderConfig", "PrtError", "camera_look_at",
    "camera_from_json", "default_config", "device_count", "scene_arrays", "DIFFUSE", "SPECULAR", "DIFFSPEC",
    "MICROFACET", "REFRACTIVE", "EMISSIVE", "INTEGRATOR_MEGAKERNEL", "INTEGRATOR_WAVEFRONT", "NO_TEXTURE",
    "write_png", "write_pfm", "Tune",
]


def device_count() -> int:
    n = C.c_int(0)
    check(lib().prt_device_count(C.byref(n)))
    return n.value


def default_config(**kw) -> RenderConfig:
    c = RenderConfig()
    lib().prt_render_config_default(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    return c


def camera_look_at(eye, target, up, vfov_deg: float, width: int, height: int) -> Camera:
    cam = Camera()
    e = np.asarray(eye, np.float32)
    t = np.asarray(target, np.float32)
    u = np.asarray(up, np.float32)
    check(lib().prt_camera_look_at(_capi.fptr(e), _capi.fptr(t), _capi.fptr(u), float(vfov_deg), width, height,
                                   C.byref(cam)))
    return cam


def camera_from_json(path, width: int, height: int) -> Camera:
    """CameraJsonLoader::getCamera (src/context/json_loader.cpp:26-48)."""
    cam = Camera()
    check(lib().prt_camera_from_json(str(path).encode(), width, height, C.byref(cam)))
    return cam


@dataclass
class SceneArrays:
    """Numpy view of a scene description (copies)."""
    positions: np.ndarray   # (n, 3, 3) float32
    normals: np.ndarray     # (n, 3, 3)
    uvs: np.ndarray         # (n, 3, 2)
    material_ids: np.ndarray  # (n,) uint32
    materials: list
    camera: Camera


def scene_arrays(desc: SceneDesc) -> SceneArrays:
    n = desc.num_triangles
    pos = np.ctypeslib.as_array(desc.positions, shape=(n * 9,)).reshape(n, 3, 3).copy()
    nrm = np.ctypeslib.as_array(desc.normals, shape=(n * 9,)).reshape(n, 3, 3).copy()
    uv = np.ctypeslib.as_array(desc.uvs, shape=(n * 6,)).reshape(n, 3, 2).copy() if desc.uvs else np.zeros((n, 3, 2), np.float32)
    mid = np.ctypeslib.as_array(desc.material_ids, shape=(n,)).copy()
    mats = [Material.from_buffer_copy(desc.materials[i]) for i in range(desc.num_materials)]
    cam = Camera.from_buffer_copy(desc.camera)
    return SceneArrays(pos, nrm, uv, mid, mats, cam)


class Mesh:
    """Host scene description: Scene::loadObj + loadTriangles (src/geometry/scene.cpp:40-138),
    or a procedural scene from a seed."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def load_obj(cls, path) -> "Mesh":
        h = C.c_void_p()
        check(lib().prt_mesh_load_obj(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def generate(cls, kind: str, params: str = "", seed: int = 0x5EED) -> "Mesh":
        h = C.c_void_p()
        check(lib().prt_mesh_generate(kind.encode(), params.encode(), seed, C.byref(h)))
        return cls(h)

    def desc(self) -> SceneDesc:
        d = SceneDesc()
        check(lib().prt_mesh_get_desc(self._h, C.byref(d)))
        return d

    def arrays(self) -> SceneArrays:
        return scene_arrays(self.desc())

    @property
    def num_triangles(self) -> int:
        return self.desc().num_triangles

    @property
    def materials(self) -> list:
        d = self.desc()
        return [Material.from_buffer_copy(d.materials[i]) for i in range(d.num_materials)]

    def material_names(self) -> list:
        names = []
        buf = C.create_string_buffer(1024)
        for i in range(self.desc().num_materials):
            check(lib().prt_mesh_material_name(self._h, i, buf, 1024))
            names.append(buf.value.decode())
        return names

    def set_material(self, index: int, m: Material) -> None:
        check(lib().prt_mesh_set_material(self._h, index, C.byref(m)))

    def set_camera(self, cam: Camera) -> None:
        check(lib().prt_mesh_set_camera(self._h, C.byref(cam)))

    @property
    def camera(self) -> Camera:
        return Camera.from_buffer_copy(self.desc().camera)

    def close(self) -> None:
        if self._h:
            lib().prt_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Built scene: lights, BVH, packed materials (Scene::constructBvh / constructLightList)."""

    def __init__(self, source):
        desc = source.desc() if isinstance(source, Mesh) else source
        self._src = source  # keeps the arrays alive while the scene is built
        h = C.c_void_p()
        check(lib().prt_scene_create(C.byref(desc), C.byref(h)))
        self._h = h
        self.num_triangles = desc.num_triangles
        self.num_materials = desc.num_materials
        self.camera = Camera.from_buffer_copy(desc.camera)
        self._params = {}  # material index -> (kd, emission) of its last set_material_params

    def bvh_info(self) -> _capi.BvhInfo:
        i = _capi.BvhInfo()
        check(lib().prt_scene_bvh_info(self._h, C.byref(i)))
        return i

    def export_bvh(self):
        info = self.bvh_info()
        nodes = np.zeros(info.num_nodes * 16, np.float32)
        tris = np.zeros(self.num_triangles * 12, np.float32)
        order = np.zeros(self.num_triangles, np.uint32)
        check(lib().prt_scene_export_bvh(self._h, _capi.fptr(nodes), _capi.fptr(tris), _capi.uptr(order)))
        return nodes.reshape(-1, 16), tris.reshape(-1, 12), order

    def lights(self):
        n = C.c_uint32()
        a = C.c_float()
        check(lib().prt_scene_num_lights(self._h, C.byref(n), C.byref(a)))
        return n.value, a.value

    def set_material_params(self, index: int, kd, emission: float) -> None:
        k = np.ascontiguousarray(kd, np.float32).reshape(-1)
        if k.size != 3:
            raise ValueError(f"kd must hold 3 values (RGB), got {k.size}")
        if not 0 <= int(index) < self.num_materials:
            raise ValueError(f"material index {index} out of range 0..{self.num_materials - 1}")
        check(lib().prt_scene_set_material_params(self._h, int(index), _capi.fptr(k), float(emission)))
        self._params[int(index)] = (k.copy(), np.float32(emission))

    def has_material_params(self, kd, emission) -> bool:
        """Whether the last set_material_params of every material set exactly these
        kd (M, 3) and emission (M,) values."""
        kd = np.asarray(kd, np.float32).reshape(-1, 3)
        em = np.asarray(emission, np.float32).reshape(-1)
        if len(kd) != self.num_materials or len(em) != self.num_materials:
            return False
        for m in range(self.num_materials):
            cur = self._params.get(m)
            if cur is None or not np.array_equal(cur[0], kd[m]) or cur[1] != em[m]:
                return False
        return True

    def close(self) -> None:
        if self._h:
            lib().prt_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParallelRenderer:
    """ParallelRenderer (src/include/renderer.h:102-126) on one HIP device."""

    def __init__(self, scene: Scene, width: int, height: int, samples: int, camera: Camera | None = None,
                 config: RenderConfig | None = None, **cfg):
        self.scene = scene
        self.config = config if config is not None else default_config(**cfg)
        self.width, self.height, self.samples = width, height, samples
        h = C.c_void_p()
        cam = C.byref(camera) if camera is not None else None
        check(lib().prt_renderer_create(scene._h, cam, width, height, samples, C.byref(self.config), C.byref(h)))
        self._h = h

    # --- reference method names -------------------------------------------
    def renderOneSamplePerPixel(self) -> None:
        check(lib().prt_render_spp(self._h, 1))

    def copyImageBytes(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.uint8)
        check(lib().prt_copy_image_bytes(self._h, _capi.bptr(out)))
        return out

    def getWidth(self) -> int:
        return self.width

    def getHeight(self) -> int:
        return self.height

    def getSamples(self) -> int:
        return self.samples

    def getSamplesRendered(self) -> int:
        n = C.c_uint32()
        check(lib().prt_renderer_samples_rendered(self._h, C.byref(n)))
        return n.value

    # --- additions ---------------------------------------------------------
    def render(self, n: int | None = None) -> None:
        """TerminalLauncher::render's loop (terminal_launcher.cpp:13-15): n samples per pixel."""
        check(lib().prt_render_spp(self._h, self.samples if n is None else n))

    def reset(self) -> None:
        check(lib().prt_renderer_reset(self._h))

    def radiance(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 3), np.float32)
        check(lib().prt_get_radiance(self._h, _capi.fptr(out)))
        return out

    def radiance_to(self, device_ptr: int) -> None:
        check(lib().prt_get_radiance_device(self._h, C.c_void_p(device_ptr)))

    def first_hit(self, sample: int = 0):
        n = self.width * self.height
        tri = np.zeros(n, np.int32)
        mat = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        check(lib().prt_first_hit(self._h, sample, _capi.iptr(tri), _capi.iptr(mat), _capi.fptr(t)))
        shp = (self.height, self.width)
        return tri.reshape(shp), mat.reshape(shp), t.reshape(shp)

    def stats(self) -> Stats:
        s = Stats()
        check(lib().prt_get_stats(self._h, C.byref(s)))
        return s

    def reset_stats(self) -> None:
        check(lib().prt_reset_stats(self._h))

    def _image_arg(self, a, name: str) -> np.ndarray:
        """a as a contiguous float32 (height, width, 3) array; anything else is refused
        (the C side reads width * height * 3 floats)."""
        arr = np.ascontiguousarray(a, np.float32)
        if arr.shape != (self.height, self.width, 3):
            raise ValueError(f"{name} must have shape ({self.height}, {self.width}, 3), got {arr.shape}")
        return arr

    def adjoint(self, target: np.ndarray, spp: int, image: np.ndarray | None = None):
        """Path-replay gradients of mean((I - target)^2) w.r.t. kd (M,3) and emission (M,)."""
        M = self.scene.num_materials
        d_alb = np.zeros((M, 3), np.float32)
        d_em = np.zeros(M, np.float32)
        loss = C.c_float()
        tgt = self._image_arg(target, "target")
        img = None if image is None else self._image_arg(image, "image")
        check(lib().prt_adjoint(self._h, _capi.fptr(tgt), None if img is None else _capi.fptr(img), spp,
                                _capi.fptr(d_alb), _capi.fptr(d_em), C.byref(loss)))
        return d_alb, d_em, loss.value

    def adjoint_grad(self, grad_image: np.ndarray, spp: int, hess: float = 0.0):
        """Path-replay gradients of a loss with dLoss/dI = grad_image and constant diagonal
        curvature d2Loss/dI2 = hess w.r.t. kd (M,3) and emission (M,)."""
        M = self.scene.num_materials
        d_alb = np.zeros((M, 3), np.float32)
        d_em = np.zeros(M, np.float32)
        g = self._image_arg(grad_image, "grad_image")
        check(lib().prt_adjoint_grad(self._h, _capi.fptr(g), float(hess), spp, _capi.fptr(d_alb),
                                     _capi.fptr(d_em)))
        return d_alb, d_em

    def close(self) -> None:
        if self._h:
            lib().prt_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class BatchJob:
    scene: Scene
    camera: Camera
    spp: int
    view_id: int = 0


def render_batch(jobs: list, config: RenderConfig | None = None, tune: "_capi.Tune | None" = None, **cfg):
    """Renders every job (views of possibly different scenes, sizes and spp) in one
    persistent wavefront pass. Returns (list of (h, w, 3) mean-radiance images, Stats)."""
    c = config if config is not None else default_config(**cfg)
    arr = (_capi.Job * len(jobs))()
    outs = []
    for i, j in enumerate(jobs):
        img = np.zeros((j.camera.ypixels, j.camera.xpixels, 3), np.float32)
        outs.append(img)
        arr[i].scene = j.scene._h
        arr[i].camera = j.camera
        arr[i].spp = j.spp
        arr[i].view_id = j.view_id
        arr[i].out_rgb = _capi.fptr(img)
    st = Stats()
    check(lib().prt_render_batch(arr, len(jobs), C.byref(c), C.byref(tune) if tune is not None else None,
                                 C.byref(st)))
    return outs, st


def autotune(jobs: list, config: RenderConfig | None = None, **cfg):
    """Searches traversal occupancy, camera-path grouping and chunk count on an 8-row
    trial of `jobs` (prt_autotune); returns (Tune, trial ms)."""
    c = config if config is not None else default_config(**cfg)
    arr = (_capi.Job * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i].scene = j.scene._h
        arr[i].camera = j.camera
        arr[i].spp = j.spp
        arr[i].view_id = j.view_id
        arr[i].out_rgb = None
    t = _capi.Tune()
    ms = C.c_double()
    check(lib().prt_autotune(arr, len(jobs), C.byref(c), C.byref(t), C.byref(ms)))
    return t, ms.value


def write_png(path, rgba: np.ndarray) -> None:
    a = np.ascontiguousarray(rgba, np.uint8)
    check(lib().prt_write_png(str(path).encode(), a.shape[1], a.shape[0], _capi.bptr(a)))


def write_pfm(path, rgb: np.ndarray) -> None:
    a = np.ascontiguousarray(rgb, np.float32)
    check(lib().prt_write_pfm(str(path).encode(), a.shape[1], a.shape[0], _capi.fptr(a)))
