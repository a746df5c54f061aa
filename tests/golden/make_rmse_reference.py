"""Writes spheres1m_1080p_fp64_reference.npz: the oracle's fp64 estimate of the
1M-triangle 1920x1080 frame (200 displaced icospheres, depth 8, roulette from
bounce 3, seed 0x5EED) at 4,096 fixed random pixels from 8,192 samples each,
sample indices 2^20 .. 2^20 + 8191 (disjoint from the 128 a rendered frame uses).
bench.py reports the RMSE of its 128-sample image against it; the reference's own
noise (about 1/8 of the 128-sample image's) is stated beside it.

Run from the repository root: python tests/golden/make_rmse_reference.py
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle as O  # noqa: E402
from cudart_amd import workloads as W  # noqa: E402

OUT = Path(__file__).resolve().parent / "spheres1m_1080p_fp64_reference.npz"
PIXELS, SPP, S0 = 4096, 8192, 1 << 20


def main():
    wl = W.spheres1m()
    px = np.sort(np.random.default_rng(20261203).choice(wl.width * wl.height, PIXELS, replace=False)).astype(np.uint32)
    orc = O.OracleScene(wl.mesh.desc())
    c = O.cfg(max_depth=wl.max_depth, rr_start=wl.rr_start, seed=W.SEED)
    t0 = time.time()
    acc = np.zeros((PIXELS, 3))
    sq = np.zeros((PIXELS, 3))
    # In 16 slices, so the per-pixel spread of the estimate can be stated too.
    k = SPP // 16
    for j in range(16):
        a, _ = orc.render_pixels(wl.camera, wl.width, wl.height, px, S0 + j * k, k, c, dtype=np.float64)
        acc += a
        sq += (a / k) ** 2
        print(f"slice {j + 1}/16, {time.time() - t0:.0f} s", flush=True)
    mean = acc / SPP
    # standard error of the 8192-sample mean from the 16 slice means
    var_slices = sq / 16 - mean ** 2
    stderr = np.sqrt(np.maximum(var_slices, 0) * 16 / 15 / 16)
    np.savez_compressed(OUT, pixels=px, rgb=mean, stderr=stderr, spp=SPP, s0=S0, width=wl.width, height=wl.height,
                        seed=W.SEED, max_depth=wl.max_depth, rr_start=wl.rr_start)
    print(f"wrote {OUT} ({time.time() - t0:.0f} s); mean radiance {mean.mean():.5f}, "
          f"rms stderr {np.sqrt((stderr ** 2).mean()):.5f}")


if __name__ == "__main__":
    main()
