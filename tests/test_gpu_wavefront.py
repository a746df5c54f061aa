"""The persistent wavefront integrator and batching: every image must be
bit-identical to the megakernel's for the same seed, view and depth (same RNG
slots, same vertex arithmetic), whatever the queue size and chunking."""
import numpy as np
import pytest

import cudart_amd as ca
import oracle as O
from cudart_amd import workloads as W

pytestmark = pytest.mark.gpu


def _render(scene, cam, w, h, spp, **cfg):
    r = ca.ParallelRenderer(scene, w, h, spp, camera=cam, count_rays=1, **cfg)
    r.render(spp)
    img = r.radiance()
    st = r.stats()
    r.close()
    return img, (st.primary_rays, st.extension_rays, st.shadow_rays)


def _bits(a):
    return a.view(np.uint32)


@pytest.mark.parametrize("queue", [0, 5000, 20000])
def test_wavefront_equals_megakernel_cornell(gpu, queue):
    wl = W.cornell(96, 64, spp=6, max_depth=6)
    scene = ca.Scene(wl.mesh)
    mk, cm = _render(scene, wl.camera, 96, 64, 6, max_depth=6)
    wf, cw = _render(scene, wl.camera, 96, 64, 6, max_depth=6, integrator=ca.INTEGRATOR_WAVEFRONT,
                     wavefront_queue=queue)
    assert np.array_equal(_bits(mk), _bits(wf))
    assert cm == cw


def test_wavefront_equals_megakernel_spheres(gpu):
    wl = W.spheres1m(160, 90, spp=4, max_depth=8, instances=40)
    scene = ca.Scene(wl.mesh)
    mk, cm = _render(scene, wl.camera, 160, 90, 4, max_depth=8)
    wf, cw = _render(scene, wl.camera, 160, 90, 4, max_depth=8, integrator=ca.INTEGRATOR_WAVEFRONT)
    assert np.array_equal(_bits(mk), _bits(wf))
    assert cm == cw


def test_wavefront_matches_oracle_and_progressive(gpu):
    wl = W.cornell(64, 48, spp=5, max_depth=5)
    scene = ca.Scene(wl.mesh)
    r = ca.ParallelRenderer(scene, 64, 48, 5, camera=wl.camera, max_depth=5, integrator=ca.INTEGRATOR_WAVEFRONT)
    r.render(2)
    r.render(3)  # continues at sample index 2
    img = r.radiance()
    orc = O.OracleScene(wl.mesh.desc())
    acc, _ = orc.render(wl.camera, 64, 48, 0, 5, O.cfg(max_depth=5))
    np.testing.assert_allclose(img, acc / np.float32(5), rtol=1e-5, atol=1e-6)


def test_batch_of_uneven_views_equals_megakernel(gpu):
    a = W.cornell(40, 30)
    b = W.spheres1m(50, 34, instances=12)
    sa, sb = ca.Scene(a.mesh), ca.Scene(b.mesh)
    cam_a2 = ca.camera_look_at((2.0, 0.6, 0.3), (0, 0.5, 0), (0, 1, 0), 45, 23, 17)
    jobs = [ca.BatchJob(sa, a.camera, 7, view_id=1), ca.BatchJob(sb, b.camera, 3, view_id=2),
            ca.BatchJob(sa, cam_a2, 2, view_id=3)]
    for queue in (0, 1000):
        tune = ca._capi.Tune(queue, 0, 64, 0)
        imgs, st = ca.render_batch(jobs, tune=tune, max_depth=7, count_rays=1)
        tot = 0
        for j, img in zip(jobs, imgs):
            ref, cnt = _render(j.scene, j.camera, j.camera.xpixels, j.camera.ypixels, j.spp, max_depth=7,
                               view_id=j.view_id)
            assert np.array_equal(_bits(img), _bits(ref))
            tot += sum(cnt)
        assert st.primary_rays + st.extension_rays + st.shadow_rays == tot


def test_autotune_returns_a_working_tuning(gpu):
    a = W.cornell(64, 64)
    s = ca.Scene(a.mesh)
    jobs = [ca.BatchJob(s, a.camera, 4)]
    tune, ms = ca.autotune(jobs, max_depth=5)
    assert tune.queue_slots > 0 and tune.block_size == 64 and ms > 0
    assert tune.group_side in (0, 1, 2, 4, 8) and tune.blocks_per_cu in (0, 20, 24, 28)
    imgs, _ = ca.render_batch(jobs, tune=tune, max_depth=5)
    ref, _ = _render(s, a.camera, 64, 64, 4, max_depth=5)
    assert np.array_equal(_bits(imgs[0]), _bits(ref))
    # every forced camera-path grouping and a small chunk size give the same image
    for g in (1, 2, 4, 8):
        t = ca.Tune()
        t.queue_slots, t.group_side = 4096, g
        imgs, _ = ca.render_batch(jobs, tune=t, max_depth=5)
        assert np.array_equal(_bits(imgs[0]), _bits(ref)), g
    bad = ca.Tune()
    bad.group_side = 3
    with pytest.raises(ca.PrtError):
        ca.render_batch(jobs, tune=bad, max_depth=5)
    bad = ca.Tune()
    bad.block_size = 128
    with pytest.raises(ca.PrtError):
        ca.render_batch(jobs, tune=bad, max_depth=5)


def test_deep_stack_spill_equals_megakernel_and_oracle(gpu):
    """40,000 stacked quads with a light above and one below: primary rays and the
    shadow rays to the lower light hold more than 16 traversal stack entries (the
    walk is checked on the CPU in test_bvh.py), so wf_extend and wf_shadow push into
    and pop from the global spill area. The image must still be bit-identical to the
    megakernel's (32 LDS entries, no spill here) and match the oracle."""
    from tests.helpers import OwnedDesc, stacked_quads
    p, n, m, mats = stacked_quads(40000)
    cam = ca.camera_look_at((0, 0, 2), (0, 0, 0), (0, 1, 0), 40, 64, 48)
    od = OwnedDesc(p, n, m, mats, cam)
    scene = ca.Scene(od.desc)
    mk, cm = _render(scene, cam, 64, 48, 4, max_depth=4)
    wf, cw = _render(scene, cam, 64, 48, 4, max_depth=4, integrator=ca.INTEGRATOR_WAVEFRONT)
    assert np.array_equal(_bits(mk), _bits(wf))
    assert cm == cw and cw[2] > 0
    orc = O.OracleScene(od.desc)
    acc, cnt = orc.render(cam, 64, 48, 0, 4, O.cfg(max_depth=4))
    np.testing.assert_allclose(wf, acc / np.float32(4), rtol=1e-5, atol=1e-6)
    assert cw == tuple(int(x) for x in cnt)
    assert float(wf.mean()) > 0.0
