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


# ---------------------------------------------------------------------------
# Path-level closed forms: whole renders through samplePixel, on scenes whose
# every pixel has a value fixed by the reference's estimator
# (src/renderer/renderer.cu:158-161 emission, :171-181 DIFFSPEC, :183-205 and
# :464-493 REFRACTIVE, :239-243 SPECULAR, :253-259 Russian roulette). Where a
# branch depends on a random number, the test reads that number itself from
# Philox4x32-10 (pinned above by its published vectors) at the slot the draw
# schedule assigns it: counter (pixel, sample, block, 0), key (seed, view);
# block 1 + 2b = {lobe, bsdf r1, bsdf r2, light pick} and block 2 + 2b =
# {light u, light v, roulette, spare} at bounce b; uniform = (x >> 8) 2^-24.
# ---------------------------------------------------------------------------
SKY = np.array([0.0, 0.025, 0.05])
SEED = 0x5EED


def _draw(pixel, sample, block, slot, seed=SEED, view=0):
    out = O.philox([pixel, sample, block, 0], [seed, view])
    return np.float32((out[slot] >> 8) * 2.0 ** -24)


def _mirror_emitter_desc(bsdf=1, ks=(0.8, 0.5, 0.3), kd=(0.6, 0.4, 0.2), dcoef=0.5, emit=5.0, w=16, h=16):
    """A mirror floor (y = 0, normal +y) seen from (0, 1, -1) at 45 degrees and a
    2 x 2 emitter at y = 2 (normal -y) that every reflected camera ray meets at
    z ~ 2. The camera's field of view is 0.5 degrees, so every pixel sees the
    same mirror-then-emitter path."""
    import cudart_amd as ca
    cam = ca.camera_look_at((0, 1, -1), (0, 0, 0), (0, 1, 0), 0.5, w, h)
    fp, fn = quad([-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1])
    lp, ln = quad([-1, 2, 1], [1, 2, 1], [1, 2, 3], [-1, 2, 3])
    mats = [material(bsdf, kd=kd, ks=ks, dcoef=dcoef), material(5, kd=(1, 1, 1), ka=(emit, emit, emit))]
    return OwnedDesc(np.concatenate([fp, lp]), np.concatenate([fn, ln]), [0, 0, 1, 1], mats, cam), cam


def _render1(od, cam, w, h, c):
    s = O.OracleScene(od.desc)
    acc, cnt = s.render(cam, w, h, 0, 1, c)
    s.close()
    return acc.reshape(h * w, 3).astype(np.float64), cnt


def test_mirror_then_emitter_is_ks_times_3ka():
    # Bounce 0 on the mirror adds its (zero) emission and sets mask = ks / pdf(1)
    # (:239-243); bounce 1 follows a SPECULAR vertex, so it adds mask * ka / A_L =
    # ks * 3 Ka (ka = 3 Ka A_L, :158-161). At depth 3 the emitter's stale direction
    # (EMISSIVE has no branch, Q4) leaves upward and adds mask * sky.
    ks, ka = np.array([0.8, 0.5, 0.3]), 5.0
    od, cam = _mirror_emitter_desc(ks=tuple(ks), emit=ka)
    img, cnt = _render1(od, cam, 16, 16, O.cfg(max_depth=2))
    np.testing.assert_allclose(img, np.broadcast_to(ks * 3 * ka, img.shape), rtol=2e-6)
    assert cnt[1] == 256 and cnt[2] == 0  # one extension ray per path, no shadow rays
    img, cnt = _render1(od, cam, 16, 16, O.cfg(max_depth=3))
    np.testing.assert_allclose(img, np.broadcast_to(ks * (3 * ka + SKY), img.shape), rtol=2e-6)
    # a diffuse floor in the same place adds no emission at bounce 1 (previous BSDF diffuse)
    od, cam = _mirror_emitter_desc(bsdf=0, kd=(0, 0, 0))
    img, _ = _render1(od, cam, 16, 16, O.cfg(max_depth=2))
    assert (img == 0).all()


def test_refractive_normal_incidence_takes_one_of_two_closed_forms():
    # A refractive quad (IOR 1.5) at z = 0 facing the camera at z = 5, an emitter
    # behind it at z = -2. At normal incidence Re = 24.25, Tr = -23.25, P = 12.375,
    # RP = Re / P = 1.95960, TP = Tr / (1 - P) = 2.04396 (:464-493, :190-205).
    # The lobe draw u = slot 0 of block 1: u > 0.25 transmits straight through,
    # mask = ks * TP, and the emitter adds ks * TP * 3 Ka (the previous BSDF is
    # REFRACTIVE); otherwise mask = RP, the vertex turns SPECULAR (mask *= ks) and
    # the mirrored ray (+z) leaves to the sky: RP * ks * sky.
    import cudart_amd as ca
    W = H = 64
    cam = ca.camera_look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), 0.01, W, H)
    gp, gn = quad([-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0])
    lp, ln = quad([-1, -1, -2], [1, -1, -2], [1, 1, -2], [-1, 1, -2])
    assert gn[0, 0, 2] == 1 and ln[0, 0, 2] == 1
    ks, ka = np.array([0.9, 0.7, 0.5]), 2.0
    mats = [material(4, ks=tuple(ks), ni=1.5), material(5, ka=(ka, ka, ka))]
    od = OwnedDesc(np.concatenate([gp, lp]), np.concatenate([gn, ln]), [0, 0, 1, 1], mats, cam)
    img, cnt = _render1(od, cam, W, H, O.cfg(max_depth=2))
    Re, Tr = 24.25, -23.25
    P = 0.25 + 0.5 * Re
    RP, TP = Re / P, Tr / (1 - P)
    transmit = ks * TP * 3 * ka
    reflect = RP * ks * SKY
    u = np.array([_draw(p, 0, 1, 0) for p in range(W * H)])
    through = u > np.float32(0.25)
    np.testing.assert_allclose(img[through], np.broadcast_to(transmit, img[through].shape), rtol=1e-5)
    np.testing.assert_allclose(img[~through], np.broadcast_to(reflect, img[~through].shape), rtol=1e-5)
    # the transmitted share is 0.75 (not P, not 0.25), within 5 binomial deviations
    frac = through.mean()
    assert abs(frac - 0.75) < 5 * math.sqrt(0.75 * 0.25 / (W * H)), frac
    assert cnt[1] == W * H


def test_diffspec_is_the_weighted_mixture_of_its_lobes():
    # DIFFSPEC (:171-181): u = slot 0 of block 1; u < p takes the diffuse lobe with
    # mask *= 1/p, else the specular lobe with mask *= 1/(1-p). Every later draw
    # uses the same slots in the pure DIFFUSE and pure SPECULAR renders of the same
    # material, so each DIFFSPEC pixel equals the DIFFUSE pixel / p or the SPECULAR
    # pixel / (1 - p), as its own lobe draw says.
    W = H = 32
    p = 0.3
    kw = dict(ks=(0.8, 0.5, 0.3), kd=(0.6, 0.4, 0.2), dcoef=p, w=W, h=H)
    c = O.cfg(max_depth=3)
    ds, _ = _render1(*_mirror_emitter_desc(bsdf=2, **kw), W, H, c)
    dd, _ = _render1(*_mirror_emitter_desc(bsdf=0, **kw), W, H, c)
    dspec, _ = _render1(*_mirror_emitter_desc(bsdf=1, **kw), W, H, c)
    u = np.array([_draw(q, 0, 1, 0) for q in range(W * H)])
    diffuse = u < np.float32(p)
    np.testing.assert_allclose(ds[diffuse], dd[diffuse] / p, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(ds[~diffuse], dspec[~diffuse] / (1 - p), rtol=1e-5)
    assert (dd[diffuse] > 0).any() and not np.allclose(dd[diffuse], dspec[diffuse])
    frac = diffuse.mean()
    assert abs(frac - p) < 5 * math.sqrt(p * (1 - p) / (W * H)), frac


def test_roulette_kills_by_its_own_draw_and_rescales_survivors():
    # Roulette from bounce 0 (rr_start = 0) on the mirror-then-emitter scene: after
    # the mirror mask = ks, p = max(ks) = 0.8; the path ends if the roulette draw
    # (slot 2 of block 2) is above p (:253-259), else mask = ks / p and the emitter
    # adds (ks / p) * 3 Ka. The mean over pixels is ks * 3 Ka (unbiased).
    W = H = 32
    ks, ka = np.array([0.8, 0.5, 0.3]), 5.0
    od, cam = _mirror_emitter_desc(ks=tuple(ks), emit=ka, w=W, h=H)
    img, cnt = _render1(od, cam, W, H, O.cfg(max_depth=2, rr_start=0))
    r = np.array([_draw(q, 0, 2, 2) for q in range(W * H)])
    p = np.float32(0.8)
    alive = ~(r > p)
    np.testing.assert_allclose(img[alive], np.broadcast_to(ks / 0.8 * 3 * ka, img[alive].shape), rtol=2e-6)
    assert (img[~alive] == 0).all()
    assert cnt[1] == alive.sum()  # killed paths trace no extension ray
    frac = alive.mean()
    assert abs(frac - 0.8) < 5 * math.sqrt(0.8 * 0.2 / (W * H)), frac
    mean = img.mean(0)
    np.testing.assert_allclose(mean, ks * 3 * ka, rtol=5 * math.sqrt(0.2 / 0.8 / (W * H)))


def test_roulette_starts_at_bounce_three_between_two_mirrors():
    # A camera between two parallel mirrors (y = 0 facing up, y = 2 facing down)
    # looking straight down: every path bounces until the depth limit or roulette.
    # With ks = (0.8, 0.5, 0.3) the mask after bounce b is ks^(b+1) rescaled by the
    # earlier roulette steps, so p = 0.8^4 at bounce 3 and 0.8 after. The extension
    # ray count must equal the count predicted from each path's own roulette draws
    # with roulette at bounces >= 3 (the reference's constant).
    import cudart_amd as ca
    W = H = 16
    D = 8
    cam = ca.camera_look_at((0, 1, 0), (0, 0, 0), (0, 0, 1), 0.01, W, H)
    fp, fn = quad([-5, 0, -5], [-5, 0, 5], [5, 0, 5], [5, 0, -5])
    cp, cn = quad([-5, 2, -5], [5, 2, -5], [5, 2, 5], [-5, 2, 5])
    lp, ln = quad([49, 0, -1], [49, 0, 1], [51, 0, 1], [51, 0, -1])  # an emitter no path meets
    assert fn[0, 0, 1] == 1 and cn[0, 0, 1] == -1
    ks = np.array([0.8, 0.5, 0.3], np.float32)
    mats = [material(1, ks=tuple(ks)), material(5, ka=(1, 1, 1))]
    od = OwnedDesc(np.concatenate([fp, cp, lp]), np.concatenate([fn, cn, ln]), [0, 0, 0, 0, 1, 1], mats, cam)
    img, cnt = _render1(od, cam, W, H, O.cfg(max_depth=D))
    assert (img == 0).all()
    want = 0
    for q in range(W * H):
        mask = np.ones(3, np.float32)
        for b in range(D):
            if b > 0:
                want += 1  # bounce b traces an extension ray
            mask = mask * ks
            if b >= 3:
                pb = mask.max()
                if _draw(q, 0, 2 + 2 * b, 2) > pb:
                    break
                mask = mask / pb
    assert cnt[1] == want, (int(cnt[1]), want)
    assert want < W * H * (D - 1)  # some paths ended by roulette
