"""Adjoint vs central finite differences on the differentiable-render workload
(100k-triangle Cornell-like scene, 16 materials, 512x512, 64 samples forward and
replay, L2 loss against a target rendered from parameters perturbed by +-0.1).

For every seed the image, the adjoint and the finite differences (h = 1e-3) use the
same random numbers; the target is rendered once, at `target_spp` samples from a
seed no image uses, like a reference photograph independent of the image's noise. Central
differences also see the Russian-roulette decisions that flip inside +-h; their
contribution is unbiased but noisy (variance ~ 1 / (paths * h)), so the comparison
is of the means over seeds, and the error of the mean falls as 1/sqrt(seeds) until
it reaches whatever bias remains.

Run standalone for the long validation (one line per seed):
    python tests/adjoint_fd.py --seeds 48
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

if __name__ == "__main__":
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))

import cudart_amd as ca  # noqa: E402
from cudart_amd import workloads as W  # noqa: E402

H_FD = 1e-3


def _materials(wl, seed=1):
    mats = wl.mesh.materials
    kd0 = np.array([list(m.kd) for m in mats], np.float32)
    em0 = np.array([m.ka[0] for m in mats], np.float32)
    rng = np.random.default_rng(seed)
    kdT = np.clip(kd0 + rng.uniform(-0.1, 0.1, kd0.shape), 0.05, 0.95).astype(np.float32)
    emT = (em0 + np.where(em0 > 0, rng.uniform(-0.1, 0.1, len(mats)), 0)).astype(np.float32)
    return kd0, em0, kdT, emT


def compare(seeds, w=512, h=512, spp=64, target_spp=512, rr_start=3, integrator=ca.INTEGRATOR_WAVEFRONT,
            log=print):
    """Returns (relative L2 error of the mean adjoint vs the mean FD after each seed,
    mean adjoint, mean FD, mean first-order adjoint, relative L2 error of each seed).
    rr_start >= the depth (6) renders without Russian roulette."""
    wl = W.cornell100k(w, h)
    scene = ca.Scene(wl.mesh)
    kd0, em0, kdT, emT = _materials(wl)
    M = len(em0)

    def set_params(kd, em):
        for i in range(M):
            scene.set_material_params(i, kd[i], float(em[i]))

    # One target for every seed: a photograph of the perturbed scene, rendered from
    # random numbers no image below uses.
    kw0 = dict(camera=wl.camera, max_depth=wl.max_depth, rr_start=rr_start, integrator=integrator)
    rt = ca.ParallelRenderer(scene, w, h, target_spp, seed=0x7A5C3E1, **kw0)
    set_params(kdT, emT)
    rt.render(target_spp)
    target = rt.radiance().astype(np.float64)
    rt.close()
    set_params(kd0, em0)

    g_sum = fd_sum = g1_sum = None
    errs, each = [], []
    for k, sd in enumerate(seeds):
        t0 = time.time()
        r = ca.ParallelRenderer(scene, w, h, spp, seed=sd, **kw0)

        def loss_at(kd, em):
            set_params(kd, em)
            r.reset()
            r.render(spp)
            return float(np.mean((r.radiance().astype(np.float64) - target) ** 2))

        set_params(kd0, em0)
        r.render(spp)
        image = r.radiance()
        d_alb, d_em, _ = r.adjoint(target.astype(np.float32), spp, image=image)
        grad = ((2.0 / image.size) * (image.astype(np.float64) - target)).astype(np.float32)
        a1, e1 = r.adjoint_grad(grad, spp, hess=0.0)
        fd = np.zeros(4 * M)
        for m in range(M):
            for c in range(3):
                kp, km = kd0.copy(), kd0.copy()
                kp[m, c] += H_FD
                km[m, c] -= H_FD
                fd[3 * m + c] = (loss_at(kp, em0) - loss_at(km, em0)) / (2 * H_FD)
            ep, en = em0.copy(), em0.copy()
            ep[m] += H_FD
            en[m] -= H_FD
            fd[3 * M + m] = (loss_at(kd0, ep) - loss_at(kd0, en)) / (2 * H_FD)
        set_params(kd0, em0)
        r.close()
        g = np.concatenate([d_alb.ravel(), d_em]).astype(np.float64)
        g1 = np.concatenate([a1.ravel(), e1]).astype(np.float64)
        g_sum = g if g_sum is None else g_sum + g
        g1_sum = g1 if g1_sum is None else g1_sum + g1
        fd_sum = fd if fd_sum is None else fd_sum + fd
        each.append(float(np.linalg.norm(g - fd) / np.linalg.norm(fd)))
        errs.append(float(np.linalg.norm(g_sum - fd_sum) / np.linalg.norm(fd_sum)))
        log(f"seed {k + 1}/{len(seeds)} ({sd}): this seed {each[-1]:.4f}, "
            f"mean so far {errs[-1]:.4f}, first-order mean "
            f"{np.linalg.norm(g1_sum - fd_sum) / np.linalg.norm(fd_sum):.4f} ({time.time() - t0:.1f} s)")
        sys.stdout.flush()
    n = len(seeds)
    scene.close()
    return errs, g_sum / n, fd_sum / n, g1_sum / n, each


def seed_list(n):
    return [W.SEED + 7919 * k for k in range(n)]


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=48)
    ap.add_argument("--target-spp", type=int, default=512)
    ap.add_argument("--rr-start", type=int, default=3, help="6 or more: no Russian roulette")
    args = ap.parse_args()
    errs, g, fd, g1, _ = compare(seed_list(args.seeds), target_spp=args.target_spp, rr_start=args.rr_start)
    print(f"relative L2 error of the mean adjoint vs the mean central FD over {args.seeds} seeds: {errs[-1]:.5f}")
    print(f"first-order adjoint (no roulette curvature term): {np.linalg.norm(g1 - fd) / np.linalg.norm(fd):.5f}")
    big = np.abs(fd) >= 0.1 * np.abs(fd).max()
    print("parameters above 10% of the largest: max relative difference "
          f"{np.max(np.abs(g[big] - fd[big]) / np.abs(fd[big])):.5f}")
