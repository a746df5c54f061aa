"""The oracle's path-replay adjoint against finite differences of its own fp64 renders.

With common random numbers and a step small enough that no Russian-roulette
decision flips, a central difference measures the pathwise derivative: the adjoint
in pathwise mode (1/p differentiated at every roulette step, including the
tie-aware derivative of max(mask) on gray albedos) must reproduce it. Without
roulette the estimator is smooth and the full gradient (albedo and emission) must
match. The expected-loss estimator differs from the pathwise one only at roulette
steps with p < 1, which the last test checks by switching roulette off.
"""
import numpy as np
import pytest

import oracle as O
from cudart_amd import workloads as W

W_, H_, SPP = 24, 24, 4


@pytest.fixture(scope="module")
def setup():
    wl = W.cornell100k(W_, H_)
    orc = O.OracleScene(wl.mesh.desc())
    mats = wl.mesh.materials
    kd0 = np.array([list(m.kd) for m in mats], np.float32)
    em0 = np.array([m.ka[0] for m in mats], np.float32)
    rng = np.random.default_rng(3)
    kdT = np.clip(kd0 + rng.uniform(-0.1, 0.1, kd0.shape), 0.05, 0.95).astype(np.float32)
    return wl, orc, kd0, em0, kdT


def _render(wl, orc, kd, em, cfg, seed_shift=0):
    for i in range(len(em)):
        orc.set_material_params(i, kd[i], float(em[i]))
    c = O.cfg(max_depth=cfg.max_depth, rr_start=cfg.rr_start, seed=cfg.seed + seed_shift)
    acc, _ = orc.render(wl.camera, W_, H_, 0, SPP, c, dtype=np.float64)
    return acc / SPP


def _fd(wl, orc, kd0, em0, target, cfg, h_kd, h_em, emission=True):
    t64 = target.astype(np.float64)

    def loss(kd, em):
        return float(np.mean((_render(wl, orc, kd, em, cfg) - t64) ** 2))

    M = len(em0)
    out = []
    for m in range(M):
        for c in range(3):
            kp, km = kd0.copy(), kd0.copy()
            kp[m, c] = np.float32(kd0[m, c] + h_kd)
            km[m, c] = np.float32(kd0[m, c] - h_kd)
            out.append((loss(kp, em0) - loss(km, em0)) / (float(kp[m, c]) - float(km[m, c])))
    if emission:
        for m in range(M):
            ep, en = em0.copy(), em0.copy()
            ep[m] = np.float32(em0[m] + h_em)
            en[m] = np.float32(em0[m] - h_em)
            out.append((loss(kd0, ep) - loss(kd0, en)) / (float(ep[m]) - float(en[m])))
    return np.array(out)


def _target(wl, orc, kdT, em0, cfg):
    return _render(wl, orc, kdT, em0, cfg, seed_shift=101).astype(np.float32)


def test_pathwise_adjoint_matches_small_step_differences(setup):
    wl, orc, kd0, em0, kdT = setup
    cfg = O.cfg(max_depth=wl.max_depth, seed=W.SEED)
    target = _target(wl, orc, kdT, em0, cfg)
    image = _render(wl, orc, kd0, em0, cfg).astype(np.float32)
    ga, _ = orc.adjoint(wl.camera, W_, H_, SPP, cfg, image, target, fp64=True, pathwise=True)
    fd = _fd(wl, orc, kd0, em0, target, cfg, 1e-5, 0, emission=False)
    g = ga.ravel()
    rel = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    assert rel < 1e-4, rel
    # the roulette terms matter here: without them the gradient is off
    ge, _ = orc.adjoint(wl.camera, W_, H_, SPP, O.cfg(max_depth=wl.max_depth, rr_start=99, seed=W.SEED), image,
                        target, fp64=True)
    assert np.linalg.norm(ge.ravel() - fd) / np.linalg.norm(fd) > 100 * rel


def test_adjoint_without_roulette_matches_differences(setup):
    wl, orc, kd0, em0, kdT = setup
    cfg = O.cfg(max_depth=4, rr_start=8, seed=W.SEED)
    target = _target(wl, orc, kdT, em0, cfg)
    image = _render(wl, orc, kd0, em0, cfg).astype(np.float32)
    ga, ge = orc.adjoint(wl.camera, W_, H_, SPP, cfg, image, target, fp64=True)
    # emission enters through ka = 3 e A_L rounded to fp32: a step of 1e-3 keeps that
    # rounding below 1e-4 relative
    fd = _fd(wl, orc, kd0, em0, target, cfg, 1e-4, 1e-3)
    g = np.concatenate([ga.ravel(), ge])
    rel = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    assert rel < 1e-4, rel


def test_expected_and_pathwise_agree_without_roulette(setup):
    wl, orc, kd0, em0, kdT = setup
    cfg = O.cfg(max_depth=5, rr_start=5, seed=W.SEED)
    target = _target(wl, orc, kdT, em0, cfg)
    image = _render(wl, orc, kd0, em0, cfg).astype(np.float32)
    a0, e0 = orc.adjoint(wl.camera, W_, H_, SPP, cfg, image, target, fp64=True)
    a1, e1 = orc.adjoint(wl.camera, W_, H_, SPP, cfg, image, target, fp64=True, pathwise=True)
    # equal up to the order of the per-thread sums
    np.testing.assert_allclose(a0, a1, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(e0, e1, rtol=1e-12, atol=1e-15)
    # with roulette they differ (the p < 1 steps)
    cfg = O.cfg(max_depth=wl.max_depth, seed=W.SEED)
    b0, _ = orc.adjoint(wl.camera, W_, H_, SPP, cfg, image, target, fp64=True)
    b1, _ = orc.adjoint(wl.camera, W_, H_, SPP, cfg, image, target, fp64=True, pathwise=True)
    assert np.linalg.norm(b0 - b1) > 1e-3 * np.linalg.norm(b1)
