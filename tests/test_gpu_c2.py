"""The 1M-triangle scene on the GPU against the oracle, at its full size: 200
displaced L4 icospheres + ground + 2 emissive quads (1,024,006 triangles; 19
diffuse, 8 specular and 5 refractive materials), 1920x1080, depth 8 with Russian
roulette from bounce 3.

First-hit triangle ids, material ids and distances over the whole frame
(bit-exact, renderer.cu:283-330 and camera.cu:32-57 against their restatement);
radiance of a 2-sample wavefront frame over 20,000 random pixels (the oracle
renders only those) at the parity tests' relative tolerance, and that frame equal
to the megakernel's bit for bit.
"""
import time

import numpy as np
import pytest

import cudart_amd as ca
import oracle as O
from cudart_amd import workloads as W

pytestmark = pytest.mark.gpu
W2, H2 = 1920, 1080


@pytest.fixture(scope="module")
def c2(gpu):
    wl = W.spheres1m(W2, H2, 128)
    scene = ca.Scene(wl.mesh)
    t0 = time.time()
    orc = O.OracleScene(wl.mesh.desc())
    print(f"\n[c2] {wl.mesh.num_triangles} triangles; oracle BVH in {time.time() - t0:.1f} s", flush=True)
    yield wl, scene, orc
    orc.close()
    scene.close()


def test_c2_first_hit_bitexact_full_frame(c2):
    wl, scene, orc = c2
    assert wl.mesh.num_triangles == 1024006
    for sample in (0, 77):
        r = ca.ParallelRenderer(scene, W2, H2, 1, camera=wl.camera, max_depth=8, seed=W.SEED)
        tri, mat, t = r.first_hit(sample)
        r.close()
        otri, omat, ot = orc.first_hit(wl.camera, W2, H2, sample, O.cfg(max_depth=8, seed=W.SEED))
        print(f"[c2] sample {sample}: hit fraction {(tri >= 0).mean():.3f}, "
              f"materials hit {len(np.unique(mat[mat >= 0]))}", flush=True)
        assert (tri >= 0).mean() > 0.5 and len(np.unique(mat[mat >= 0])) >= 30
        assert np.array_equal(tri, otri)
        assert np.array_equal(mat, omat)
        assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))


def test_c2_radiance_matches_oracle_on_pixel_sample(c2):
    wl, scene, orc = c2
    spp = 2
    kw = dict(camera=wl.camera, max_depth=8, rr_start=3, seed=W.SEED)
    r = ca.ParallelRenderer(scene, W2, H2, spp, integrator=ca.INTEGRATOR_WAVEFRONT, **kw)
    r.render(spp)
    img = r.radiance().reshape(-1, 3)
    r.close()
    m = ca.ParallelRenderer(scene, W2, H2, spp, integrator=ca.INTEGRATOR_MEGAKERNEL, **kw)
    m.render(spp)
    assert np.array_equal(m.radiance().reshape(-1, 3).view(np.uint32), img.view(np.uint32))
    m.close()
    px = np.sort(np.random.default_rng(2).choice(W2 * H2, 20000, replace=False)).astype(np.uint32)
    acc, _ = orc.render_pixels(wl.camera, W2, H2, px, 0, spp, O.cfg(max_depth=8, rr_start=3, seed=W.SEED))
    ref = acc / np.float32(spp)
    got = img[px]
    assert np.isfinite(got).all()
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)
    print(f"[c2] bit-identical pixels: {np.mean(got.view(np.uint32) == ref.view(np.uint32)):.6f}", flush=True)
