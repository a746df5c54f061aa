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


def read_png_rgba(path) -> np.ndarray:
    """Decodes an 8-bit RGBA, non-interlaced PNG (all five row filters) to (H, W, 4) uint8."""
    import struct
    import zlib

    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n", "not a PNG"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) & 0xFFFFFFFF == crc, "bad CRC"
        if typ == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype, _, _, interlace = hdr
    assert depth == 8 and ctype == 6 and interlace == 0, hdr
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    out = np.zeros((h, 4 * w), np.int32)
    for y in range(h):
        f, row = raw[y, 0], raw[y, 1:].astype(np.int32)
        prev = out[y - 1] if y else np.zeros(4 * w, np.int32)
        cur = out[y]
        for x in range(4 * w):
            a = cur[x - 4] if x >= 4 else 0
            b = prev[x]
            c = prev[x - 4] if x >= 4 else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[x] = (row[x] + p) & 0xFF
    return out.astype(np.uint8).reshape(h, w, 4)


def stacked_quads(k: int, spacing: float = 0.01):
    """k diffuse quads over [-1, 1]^2 at z = 0, -spacing, ..., facing +z, an emissive
    quad above them at z = 3 facing down and one below them facing up. A ray along
    -z meets every child box at every level of the BVH, so its traversal stack
    grows by three entries per level (the three farther children)."""
    zb = -spacing * k - 1.0
    ps, ns, mids = [], [], []
    for i in range(k):
        z = -spacing * i
        p, n = quad([-1, -1, z], [1, -1, z], [1, 1, z], [-1, 1, z])
        ps.append(p)
        ns.append(n)
        mids += [0, 0]
    for corners in (([-0.5, -0.5, 3], [-0.5, 0.5, 3], [0.5, 0.5, 3], [0.5, -0.5, 3]),  # facing -z
                    ([-0.5, -0.5, zb], [0.5, -0.5, zb], [0.5, 0.5, zb], [-0.5, 0.5, zb])):  # facing +z
        p, n = quad(*corners)
        ps.append(p)
        ns.append(n)
        mids += [1, 1]
    return np.concatenate(ps), np.concatenate(ns), mids, [material(0), material(5, ka=(1, 1, 1))]


def stack_at_first_leaf(nodes: np.ndarray, origin, direction) -> int:
    """Traversal stack entries held when a nearest-first walk of the four-wide layout
    (export_bvh's nodes) first reaches a leaf: each visited node pushes its hit
    children but the nearest, and the walk goes on with the nearest. The node
    layout: origin xyz, exponent bytes, 4 child words, then per axis the low and
    high 8-bit planes of the four children."""
    u = nodes.view(np.uint32)
    o = np.asarray(origin, np.float64)
    d = np.asarray(direction, np.float64)
    inv = 1.0 / np.where(np.abs(d) < 2.0 ** -64, np.copysign(2.0 ** -64, d), d)  # the kernels' clamp
    node, depth = 0, 0
    while True:
        org = nodes[node, 0:3].astype(np.float64)
        step = np.array([2.0 ** (int((u[node, 3] >> (8 * a)) & 0xFF) - 127) for a in range(3)])
        hits = []
        for c in range(4):
            w = int(nodes[node, 4:8].view(np.int32)[c])
            if w == 0x7FFFFFFF:
                continue
            lo = np.array([(u[node, 8 + 2 * a] >> (8 * c)) & 0xFF for a in range(3)], np.float64) * step + org
            hi = np.array([(u[node, 9 + 2 * a] >> (8 * c)) & 0xFF for a in range(3)], np.float64) * step + org
            t0, t1 = (lo - o) * inv, (hi - o) * inv
            tn, tf = max(np.minimum(t0, t1).max(), 0.0), np.maximum(t0, t1).min()
            if tn <= tf:
                hits.append((tn, w))
        assert hits, "the walk left the tree without reaching a leaf"
        hits.sort()
        depth += len(hits) - 1
        node = hits[0][1]
        if node < 0:
            return depth
