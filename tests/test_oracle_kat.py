"""Known-answer and closed-form checks that pin the CPU oracle.

Each test compares an oracle function with a value fixed by something other than
the oracle's own formula: published test vectors (Philox), the mathematics of the
operation (sin/cos, ray/triangle and ray/box geometry, cosine-sampling moments),
hand derivations of the reference's quirky formulas (Fresnel Q5), or statistical
expectations of the estimator (Q1/Q2/Q9/Q4). A plausible slip (a dropped factor,
a wrong sign, a swapped operand) changes at least one of them.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from cudart_amd._capi import Camera
from tests.helpers import OwnedDesc, material, quad

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_philox_known_answers():
    kat = json.loads((GOLDEN / "philox4x32_10_kat.json").read_text())
    for v in kat["vectors"]:
        out = O.philox([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]])
        assert [f"{x:08x}" for x in out] == v["out"]


def test_sincos_within_two_ulp():
    x = np.linspace(-4.0, 8.0, 400001, dtype=np.float32)
    s, c = O.sincosf(x)
    xd = x.astype(np.float64)
    for got, ref in ((s, np.sin(xd)), (c, np.cos(xd))):
        err = np.abs(got.astype(np.float64) - ref)
        ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
        assert (err <= 2.0 * ulp + 1e-45).all(), float((err / ulp).max())
    # exact special values
    s0, c0 = O.sincosf(np.array([0.0], np.float32))
    assert s0[0] == 0.0 and c0[0] == 1.0


def test_moller_trumbore_closed_forms():
    v1, e1, e2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    # ray through the centroid along -z: u = v = 1/3, t = distance
    hit, t, u, v = O.triangle_intersect(v1, e1, e2, [1 / 3, 1 / 3, 2.0], [0, 0, -1])
    assert hit and t == pytest.approx(2.0, abs=1e-7) and u == pytest.approx(1 / 3, abs=1e-7)
    assert v == pytest.approx(1 / 3, abs=1e-7)
    # axis-aligned exact t
    hit, t, u, v = O.triangle_intersect(v1, e1, e2, [0.25, 0.25, 5.0], [0, 0, -1])
    assert hit and t == 5.0 and u == 0.25 and v == 0.25
    # outside (u + v > 1), behind, parallel, tiny determinant and t <= 1e-5 are misses
    assert not O.triangle_intersect(v1, e1, e2, [0.6, 0.6, 1.0], [0, 0, -1])[0]
    assert not O.triangle_intersect(v1, e1, e2, [0.2, 0.2, 1.0], [0, 0, 1])[0]
    assert not O.triangle_intersect(v1, e1, e2, [0.2, 0.2, 1.0], [1, 0, 0])[0]
    tiny = [0.001, 0, 0], [0, 0.001, 0]  # |det| = 1e-6 < 1e-5: rejected (absolute threshold, Q13)
    assert not O.triangle_intersect(v1, tiny[0], tiny[1], [0.0002, 0.0002, 1.0], [0, 0, -1])[0]
    assert not O.triangle_intersect(v1, e1, e2, [0.2, 0.2, 5e-6], [0, 0, -1])[0]
    # hits on the edge u = 0 are accepted
    assert O.triangle_intersect(v1, e1, e2, [0.0, 0.5, 1.0], [0, 0, -1])[0]


def test_box_test_is_the_infinite_line():
    lo, hi = [0, 0, 0], [1, 1, 1]
    assert O.box_hit(lo, hi, [0.5, 0.5, -2], [0, 0, 1])
    # the box behind the origin still counts (the reference ignores tMin/tMax, Q14)
    assert O.box_hit(lo, hi, [0.5, 0.5, 3], [0, 0, 1])
    assert not O.box_hit(lo, hi, [2.5, 0.5, -2], [0, 0, 1])
    # a zero direction component with the origin on the slab plane gives 0*inf = NaN,
    # which the comparisons ignore
    assert O.box_hit(lo, hi, [0.0, 0.5, -2], [0, 0, 1])


def test_fresnel_quirk_normal_incidence():
    # IOR 1.5, entering along -n: R0 = ((1-1.5)^2 / 2.5) * 2.5 = 0.25 (operator precedence),
    # t = -n, c = 1 - dot(t, n) = 2, Re = R0 + (1-R0) * 2^5 = 24.25 (renderer.cu:464-493).
    re, tr, td = O.fresnel([0, 0, -1], [0, 0, 1], 1.5)
    assert re == pytest.approx(24.25, rel=1e-6)
    assert tr == pytest.approx(-23.25, rel=1e-6)
    np.testing.assert_allclose(td, [0, 0, -1], atol=1e-7)
    P = 0.25 + 0.5 * re
    assert P == pytest.approx(12.375, rel=1e-6)
    assert re / P == pytest.approx(1.9596, rel=1e-4)
    assert tr / (1 - P) == pytest.approx(2.0440, rel=1e-4)


def test_fresnel_total_internal_reflection_and_exit():
    # leaving glass (dot(wo, n) > 0) at a grazing angle: k < 0 -> Re = 1
    wo = np.array([0.9, 0.0, math.sqrt(1 - 0.81)], np.float32)
    re, tr, _ = O.fresnel(wo, [0, 0, 1], 1.5)
    assert re == 1.0 and tr == 0.0
    # leaving at normal incidence: etai = 1.5, etat = 1, n flipped; t = +n; c = 1 - 1 = 0 -> Re = R0 = 0.25
    re, tr, td = O.fresnel([0, 0, 1], [0, 0, 1], 1.5)
    assert re == pytest.approx(0.25, rel=1e-6)
    np.testing.assert_allclose(td, [0, 0, 1], atol=1e-6)


def test_diffuse_sampler_moments():
    rng = np.random.default_rng(3)
    n = np.array([0.3, 0.5, -0.81], np.float64)
    n = (n / np.linalg.norm(n)).astype(np.float32)
    dots = []
    for u1, u2 in rng.random((20000, 2)):
        d, cw = O.diffuse_dir(n, u1, u2)
        assert abs(np.linalg.norm(d) - 1) < 1e-5
        assert cw >= -1e-6
        dots.append(cw)
    dots = np.array(dots)
    # cosine-weighted hemisphere: E[cos] = 2/3, E[cos^2] = 1/2
    assert dots.mean() == pytest.approx(2 / 3, abs=0.01)
    assert (dots ** 2).mean() == pytest.approx(0.5, abs=0.01)


def _cam(eye, target, w, h, vfov=40.0):
    import cudart_amd as ca
    return ca.camera_look_at(eye, target, (0, 1, 0), vfov, w, h)


def test_camera_center_ray_closed_form():
    import cudart_amd as ca
    cam = ca.camera_look_at((1, 2, 3), (0, 0, 0), (0, 1, 0), 50.0, 64, 48)
    for (i, j) in [(0, 0), (10, 20), (63, 47)]:
        o, d = O.camera_ray(cam, i, j, 0.5, 0.5)  # tent(1.0) = 0: no jitter
        ref = (np.array(cam.dir) + np.array(cam.right) * cam.fov * cam.aspect * (i / 64 - 0.5)
               + np.array(cam.up) * cam.fov * (0.5 - j / 48))
        np.testing.assert_allclose(d, ref / np.linalg.norm(ref), atol=1e-6)
        np.testing.assert_allclose(o, cam.eye, atol=0)
    # the image spans the full vertical field of view
    o, top = O.camera_ray(cam, 32, 0, 0.5, 0.5)
    ang = math.degrees(math.acos(float(np.dot(top, cam.dir))))
    assert ang == pytest.approx(25.0, abs=0.2)


def test_tonemap_exhaustive_table():
    # float3ToUchar4 then gammaCorrectPixel: q = trunc(min(v,1)*255), out = trunc((q/255)^(1/2.2)*255)
    v = (np.arange(256, dtype=np.float64) + 0.5) / 255.0
    rgb = np.repeat(v[:, None], 3, 1).astype(np.float32)
    out = O.tonemap(rgb)[:, 0].astype(int)
    q = np.floor(np.minimum(rgb[:, 0].astype(np.float64), 1.0) * 255.0)
    g = np.power(q / 255.0, 1 / 2.2) * 255.0
    ref = np.floor(g).astype(int)
    near = np.abs(g - np.round(g)) < 1e-4  # ties decided by float rounding
    assert (out[~near] == ref[~near]).all()
    # clamp above 1, negative and NaN map to 0
    t = O.tonemap(np.array([[2.0, -1.0, np.nan]], np.float32))[0]
    assert list(t) == [255, 0, 0, 255]


def test_texture_table_is_gamma_22():
    t = O.texture_table()
    ref = np.power(np.arange(256) / 255.0, 2.2)
    np.testing.assert_allclose(t, ref, rtol=2e-6, atol=1e-9)


def _emitter_scene(camera):
    # Floor quad y=0 (diffuse kd) and a 0.5 x 0.5 emitter quad at y = 2 facing down.
    fp, fn = quad([-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1])
    lp, ln = quad([-0.25, 2, -0.25], [0.25, 2, -0.25], [0.25, 2, 0.25], [-0.25, 2, 0.25])
    pos = np.concatenate([fp, lp])
    nrm = np.concatenate([fn, ln])
    mats = [material(0, kd=(0.6, 0.4, 0.2)), material(5, kd=(1, 1, 1), ka=(5, 5, 5))]
    return OwnedDesc(pos, nrm, [0, 0, 1, 1], mats, camera)


def test_sky_pixels_are_exact():
    cam = _cam((0, 1, 5), (0, 1, 10), 8, 8)  # looking away from everything
    od = _emitter_scene(cam)
    s = O.OracleScene(od.desc)
    acc, cnt = s.render(cam, 8, 8, 0, 3, O.cfg())
    img = acc / np.float32(3)
    assert (img == np.array([0.0, 0.025, 0.05], np.float32)).all()
    assert cnt[0] == 8 * 8 * 3 and cnt[1] == 0 and cnt[2] == 0


def test_direct_emitter_pixel_is_3ka_plus_sky():
    # Bounce 0 on an EMISSIVE triangle adds ka / A_L = 3 Ka; its stale zero direction
    # then misses and adds the sky (Q3, Q4).
    cam = _cam((0, 1.0, 0), (0, 2, 0.0001), 4, 4, vfov=2.0)
    od = _emitter_scene(cam)
    s = O.OracleScene(od.desc)
    acc, cnt = s.render(cam, 4, 4, 0, 2, O.cfg())
    img = acc / np.float32(2)
    np.testing.assert_allclose(img, np.broadcast_to(np.array([15.0, 15.025, 15.05]), img.shape), rtol=1e-6)
    assert cnt[1] == cnt[0]  # one extension ray (the stale zero direction) per path


def test_next_event_estimate_expectation():
    # Depth 1, camera straight down at the floor centre: the pixel value is
    # NEE only. The light colour is ka = 3 Ka A_L (scene.cpp:161-164) and the
    # weight divides by A_L again (renderer.cu:441), so one pick of light
    # triangle l contributes Le A_l cos_l / d^2. Its expectation is
    # (2 kd E[cos]) * (1/numLights) * int Le cos_l / d^2 dA
    # with E[cos] = 2/3 (Q1: pdf fixed at 0.5) and no factor numLights (Q2).
    cam = _cam((0, 1.0, 0), (0, 0, 0.0001), 1, 1, vfov=0.01)
    od = _emitter_scene(cam)
    s = O.OracleScene(od.desc)
    spp = 200000
    acc, _ = s.render(cam, 1, 1, 0, spp, O.cfg(max_depth=1), threads=1)
    est = acc[0, 0] / spp
    # integral over the 0.5 x 0.5 light at height 2 above the shading point
    g = np.linspace(-0.25, 0.25, 801)
    X, Z = np.meshgrid(g, g)
    d2 = X ** 2 + Z ** 2 + 4.0
    integrand = 15.0 * (2.0 / np.sqrt(d2)) / d2
    integral = np.trapezoid(np.trapezoid(integrand, g, axis=1), g)
    kd = np.array([0.6, 0.4, 0.2])
    expect = 2 * kd * (2 / 3) * (1 / 2) * integral
    np.testing.assert_allclose(est, expect, rtol=0.01)


def test_oracle_deterministic_across_thread_counts():
    import cudart_amd as ca
    from cudart_amd import workloads as W
    wl = W.cornell(48, 32, spp=4)
    s = O.OracleScene(wl.mesh.desc())
    cam = ca.camera_look_at((2.2, 0.548, 0.0), (0.0, 0.548, 0.0), (0, 1, 0), 39.3, 48, 32)
    a, ca_ = s.render(cam, 48, 32, 0, 4, O.cfg(), threads=1)
    b, cb = s.render(cam, 48, 32, 0, 4, O.cfg(), threads=8)
    assert np.array_equal(a, b) and np.array_equal(ca_, cb)
    # sample ranges compose: [0,2) + [2,4) == [0,4) in the same order
    c, _ = s.render(cam, 48, 32, 0, 2, O.cfg())
    c, _ = s.render(cam, 48, 32, 2, 2, O.cfg(), accum=c)
    assert np.array_equal(a, c)


def test_f32_and_f64_oracles_agree():
    import cudart_amd as ca
    from cudart_amd import workloads as W
    wl = W.cornell(32, 32, spp=8)
    s = O.OracleScene(wl.mesh.desc())
    cam = wl.camera
    a, ca_ = s.render(cam, 32, 32, 0, 8, O.cfg(max_depth=5))
    b, cb = s.render(cam, 32, 32, 0, 8, O.cfg(max_depth=5), dtype=np.float64)
    # the same random numbers drive both; paths can part only where fp32 rounding
    # flips a discrete decision, so compare image-level error, not every pixel
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    assert np.median(rel) < 1e-5
    assert np.mean(rel < 1e-3) > 0.97


def test_texture_sampling_identities():
    # Bilinear with wrap on both axes, texel (i, j) centred at ((i + .5) / W, (j + .5) / H),
    # rows indexed by v. Checked by identities a wrong axis, offset or wrap would break.
    rng = np.random.default_rng(4)
    H, W = 5, 7
    img = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    lin = O.texture_table()[img[..., :3]].astype(np.float64)  # the gamma-2.2 table, pinned above
    # texel centres return the texel
    ij = [(i, j) for j in range(H) for i in range(W)]
    uv = np.array([((i + 0.5) / W, (j + 0.5) / H) for i, j in ij])
    got = O.texture_sample(img, uv)
    np.testing.assert_allclose(got, np.array([lin[j, i] for i, j in ij]), rtol=1e-12, atol=1e-12)
    # half-way between horizontal neighbours: their mean, including the wrap from W-1 to 0
    uv = np.array([((i + 1.0) / W, (j + 0.5) / H) for i, j in ij])
    want = np.array([(lin[j, i] + lin[j, (i + 1) % W]) / 2 for i, j in ij])
    np.testing.assert_allclose(O.texture_sample(img, uv), want, rtol=1e-12, atol=1e-12)
    # and vertical neighbours
    uv = np.array([((i + 0.5) / W, (j + 1.0) / H) for i, j in ij])
    want = np.array([(lin[j, i] + lin[(j + 1) % H, i]) / 2 for i, j in ij])
    np.testing.assert_allclose(O.texture_sample(img, uv), want, rtol=1e-12, atol=1e-12)
    # periodic in u and v; a constant texture is reproduced; fp32 within rounding of fp64
    q = rng.uniform(-3, 3, (200, 2))
    base = O.texture_sample(img, q)
    np.testing.assert_allclose(O.texture_sample(img, q + [1.0, -2.0]), base, rtol=1e-9, atol=1e-12)
    const = np.full((H, W, 4), 200, np.uint8)
    c = float(O.texture_table()[200])
    np.testing.assert_allclose(O.texture_sample(const, q), np.full((200, 3), c), rtol=1e-12)
    np.testing.assert_allclose(O.texture_sample(img, q, fp64=False), base, rtol=1e-5, atol=1e-6)
