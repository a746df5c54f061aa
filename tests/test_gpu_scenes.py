"""Awkward scenes through every GPU traversal: a scene that is one leaf (the root
node holding it), many exact duplicates (coincident centroids), tiny clustered
triangles, long thin strips, and zero-area triangles mixed in. On each, first hits
are bit-exact against the oracle, the wavefront image (camera rays traced as
packets, later rays per lane) is bit-identical to the megakernel's with the same
ray counts, and radiance agrees with the oracle's (rtol 1e-5, atol 1e-6)."""
import numpy as np
import pytest

import cudart_amd as ca
import oracle as O
from tests.helpers import OwnedDesc, material, quad

pytestmark = pytest.mark.gpu

W_, H_, SPP, DEPTH = 48, 36, 4, 4


def _face_normals(p):
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where(ln > 0, n / np.where(ln > 0, ln, 1), np.array([0.0, 0.0, 1.0]))
    return np.repeat(n[:, None, :], 3, axis=1).astype(np.float32)


def _scene(kind):
    rng = np.random.default_rng(["leaf", "duplicates", "clustered", "strips", "degenerate"].index(kind))
    light_p, light_n = quad([-1, 2.5, -1], [1, 2.5, -1], [1, 2.5, 1], [-1, 2.5, 1])  # faces -y
    if kind == "leaf":
        p = np.zeros((0, 3, 3), np.float32)
    elif kind == "duplicates":
        base = rng.normal(size=(3, 3, 3))
        p = base[rng.integers(0, 3, 60)]
    elif kind == "clustered":
        p = rng.normal(size=(400, 1, 3)) + rng.normal(size=(400, 3, 3)) * 0.05
    elif kind == "strips":
        p = rng.normal(size=(300, 3, 3)) * np.array([2.0, 1e-3, 0.5])
    else:
        p = rng.normal(size=(300, 3, 3))
        p[::3, 2] = p[::3, 1]                            # a repeated vertex
        p[1::3, 2] = 2 * p[1::3, 1] - p[1::3, 0]         # collinear vertices
    p = p.astype(np.float32)
    n = _face_normals(p) if len(p) else np.zeros((0, 3, 3), np.float32)
    pos = np.concatenate([p, light_p[None].reshape(-1, 3, 3)]).astype(np.float32)
    nrm = np.concatenate([n, light_n.reshape(-1, 3, 3)]).astype(np.float32)
    mids = [int(i) % 3 for i in range(len(p))] + [3, 3]
    mats = [material(0, kd=(0.7, 0.5, 0.3)), material(1, ks=(0.9, 0.9, 0.9)),
            material(2, kd=(0.4, 0.6, 0.5), ks=(0.5, 0.5, 0.5), dcoef=0.6), material(5, ka=(4, 4, 4))]
    # the emitter (y = 2.5) and the triangles around the origin are both in view
    cam = ca.camera_look_at((0.3, 0.8, 6), (0, 1, 0), (0, 1, 0), 45, W_, H_)
    return OwnedDesc(pos, nrm, mids, mats, cam)


def _render(scene, cam, **cfg):
    r = ca.ParallelRenderer(scene, W_, H_, SPP, camera=cam, max_depth=DEPTH, count_rays=1, **cfg)
    r.render(SPP)
    img = r.radiance()
    st = r.stats()
    first = r.first_hit(0)
    r.close()
    return img, (st.primary_rays, st.extension_rays, st.shadow_rays), first


@pytest.mark.parametrize("kind", ["leaf", "duplicates", "clustered", "strips", "degenerate"])
def test_awkward_scene_parity(gpu, kind):
    od = _scene(kind)
    print(f"[scenes] {kind}: {od.desc.num_triangles} triangles", flush=True)
    scene = ca.Scene(od.desc)
    if kind == "leaf":
        assert scene.bvh_info().num_nodes == 1
    cam = od.desc.camera
    mk, cm, (tri, mat, t) = _render(scene, cam)
    wf, cw, _ = _render(scene, cam, integrator=ca.INTEGRATOR_WAVEFRONT)
    assert np.array_equal(mk.view(np.uint32), wf.view(np.uint32))
    assert cm == cw
    orc = O.OracleScene(od.desc)
    c = O.cfg(max_depth=DEPTH)
    otri, omat, ot = orc.first_hit(cam, W_, H_, 0, c)
    assert (tri >= 0).any()
    assert np.array_equal(tri, otri)
    assert np.array_equal(mat, omat)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    acc, _ = orc.render(cam, W_, H_, 0, SPP, c)
    np.testing.assert_allclose(wf, acc / np.float32(SPP), rtol=1e-5, atol=1e-6)
    assert float(wf.max()) > 0.0
