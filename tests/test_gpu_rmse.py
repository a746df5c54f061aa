"""Image error of the 1M-triangle 1080p frame against the fp64 oracle.

The GPU's 128-sample image (wavefront; the megakernel's is bit-identical, so
"matched RMSE" holds by construction) is compared on 1,024 random pixels with
the oracle's fp64 estimate from 4,096 other samples of the same pixels
(independent sample indices). The RMSE must fall like 1/sqrt(spp) between 32 and
128 samples (no bias beyond the noise), and the mean error must be within the
noise of the mean.
"""
import numpy as np
import pytest

import cudart_amd as ca
import oracle as O
from cudart_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_rmse_against_fp64_oracle_spheres1m(gpu):
    wl = W.spheres1m()
    scene = ca.Scene(wl.mesh)
    r = ca.ParallelRenderer(scene, wl.width, wl.height, wl.spp, camera=wl.camera, max_depth=wl.max_depth,
                            seed=W.SEED, integrator=ca.INTEGRATOR_WAVEFRONT)
    r.render(32)
    img32 = r.radiance().reshape(-1, 3).astype(np.float64)
    r.render(wl.spp - 32)
    img128 = r.radiance().reshape(-1, 3).astype(np.float64)
    r.close()
    px = np.sort(np.random.default_rng(11).choice(wl.width * wl.height, 1024, replace=False)).astype(np.uint32)
    orc = O.OracleScene(wl.mesh.desc())
    ref_spp = 4096
    acc, _ = orc.render_pixels(wl.camera, wl.width, wl.height, px, 1 << 20, ref_spp,
                               O.cfg(max_depth=wl.max_depth, seed=W.SEED), dtype=np.float64)
    ref = acc / ref_spp
    e128 = img128[px] - ref
    e32 = img32[px] - ref
    rmse128 = float(np.sqrt(np.mean(e128 ** 2)))
    rmse32 = float(np.sqrt(np.mean(e32 ** 2)))
    mean_rad = float(ref.mean())
    print(f"RMSE vs fp64 oracle ({ref_spp} spp, {len(px)} px): 128 spp {rmse128:.5f}, 32 spp {rmse32:.5f} "
          f"(ratio {rmse128 / rmse32:.3f}); mean radiance {mean_rad:.4f}; mean error {e128.mean():+.2e}")
    assert 0.35 < rmse128 / rmse32 < 0.65
    assert abs(e128.mean()) < 4 * e128.std() / np.sqrt(e128.size)
