"""Scene construction helpers for tests (numpy arrays -> prt_scene_desc)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from cudart_amd import _capi
from cudart_amd._capi import Camera, Material, SceneDesc


class OwnedDesc:
    """A SceneDesc whose arrays are kept alive by this object."""

    def __init__(self, positions, normals, material_ids, materials, camera: Camera, uvs=None):
        self.positions = np.ascontiguousarray(positions, np.float32).reshape(-1, 9)
        n = len(self.positions)
        self.normals = np.ascontiguousarray(normals, np.float32).reshape(n, 9)
        self.uvs = np.zeros((n, 6), np.float32) if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(n, 6)
        self.material_ids = np.ascontiguousarray(material_ids, np.uint32).reshape(n)
        self.mats = (Material * len(materials))(*materials)
        self.desc = SceneDesc()
        d = self.desc
        d.num_triangles = n
        d.positions = _capi.fptr(self.positions)
        d.normals = _capi.fptr(self.normals)
        d.uvs = _capi.fptr(self.uvs)
        d.material_ids = _capi.uptr(self.material_ids)
        d.num_materials = len(materials)
        d.materials = C.cast(self.mats, C.POINTER(Material))
        d.num_textures = 0
        d.textures = None
        d.camera = camera


def material(bsdf=0, kd=(0.5, 0.5, 0.5), ka=(0, 0, 0), ks=(0, 0, 0), ni=1.5, dcoef=0.5) -> Material:
    m = Material()
    m.bsdf = bsdf
    m.kd[:] = kd
    m.ka[:] = ka
    m.ks[:] = ks
    m.ns = 0.0
    m.ni = ni
    m.diffuse_coef = dcoef
    m.tex_kd = -1
    return m


def quad(a, b, c, d):
    """Two triangles (a,b,c), (a,c,d) with the face normal on every vertex."""
    a, b, c, d = (np.asarray(x, np.float32) for x in (a, b, c, d))
    n = np.cross(b - a, c - a)
    n = (n / np.linalg.norm(n)).astype(np.float32)
    pos = np.array([[a, b, c], [a, c, d]], np.float32)
    nrm = np.array([[n, n, n], [n, n, n]], np.float32)
    return pos, nrm
