"""The large textured scene on the GPU against the oracle, at its full size:
2,000 displaced icospheres plus a 4M-triangle fBm height field (9,998,796
triangles), 16 RGBA8 1024x1024 textures on UV-mapped diffuse materials, 64
materials, 3840x2160, depth 8 with Russian roulette.

First-hit triangle ids, material ids and distances are compared over the whole
frame (bit-exact); radiance of a 2-sample frame over 20,000 random pixels (the
oracle renders only those), at the same relative tolerance as the other parity
tests and in practice bit-identical.
"""
import time

import numpy as np
import pytest

import cudart_amd as ca
import oracle as O
from cudart_amd import workloads as W

pytestmark = pytest.mark.gpu
W5, H5 = 3840, 2160


@pytest.fixture(scope="module")
def c5(gpu):
    t0 = time.time()
    wl = W.large(W5, H5, 512)
    print(f"\n[c5] generated {wl.mesh.num_triangles} triangles in {time.time() - t0:.1f} s", flush=True)
    t0 = time.time()
    scene = ca.Scene(wl.mesh)
    print(f"[c5] product BVH {scene.bvh_info().num_nodes} nodes in {time.time() - t0:.1f} s", flush=True)
    t0 = time.time()
    orc = O.OracleScene(wl.mesh.desc())
    print(f"[c5] oracle BVH in {time.time() - t0:.1f} s", flush=True)
    yield wl, scene, orc
    orc.close()
    scene.close()


def test_c5_first_hit_bitexact(c5):
    wl, scene, orc = c5
    r = ca.ParallelRenderer(scene, W5, H5, 1, camera=wl.camera, max_depth=8, seed=W.SEED)
    tri, mat, t = r.first_hit(0)
    r.close()
    otri, omat, ot = orc.first_hit(wl.camera, W5, H5, 0, O.cfg(max_depth=8, seed=W.SEED))
    textured = {i for i, m in enumerate(wl.mesh.materials) if m.tex_kd >= 0}
    print(f"[c5] hit fraction {(tri >= 0).mean():.3f}, textured-material hits "
          f"{np.isin(mat, list(textured)).mean():.3f}", flush=True)
    assert (tri >= 0).mean() > 0.5 and np.isin(mat, list(textured)).mean() > 0.05
    assert np.array_equal(tri, otri)
    assert np.array_equal(mat, omat)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))


def test_c5_radiance_matches_oracle_on_pixel_sample(c5):
    wl, scene, orc = c5
    spp = 2
    r = ca.ParallelRenderer(scene, W5, H5, spp, camera=wl.camera, max_depth=8, seed=W.SEED,
                            integrator=ca.INTEGRATOR_WAVEFRONT)
    r.render(spp)
    img = r.radiance().reshape(-1, 3)
    r.close()
    px = np.sort(np.random.default_rng(5).choice(W5 * H5, 20000, replace=False)).astype(np.uint32)
    acc, _ = orc.render_pixels(wl.camera, W5, H5, px, 0, spp, O.cfg(max_depth=8, seed=W.SEED))
    ref = acc / np.float32(spp)
    got = img[px]
    assert np.isfinite(got).all()
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)
    print(f"[c5] bit-identical pixels: {np.mean(got.view(np.uint32) == ref.view(np.uint32)):.6f}", flush=True)
