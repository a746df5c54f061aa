"""The batched persistent wavefront on the full batched workload: 64 square views
(128..1024 pixels per side, 8/32/128/256 spp) of 8 seeded sphere scenes of 10k-500k
triangles, rendered in one autotuned call.

Every view must equal, bit for bit, the megakernel's render of that view alone (same
seed, view id and depth); that is checked on the eight views with the fewest paths
and the largest one. The autotuned batch's time is reported beside the default
tuning's (it must be no slower) and beside rendering the 64 views one call at a
time with the same integrator (it must be at least 5% faster).
"""
import time

import numpy as np
import pytest

import cudart_amd as ca
from cudart_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_full_batch_equals_per_view_megakernel(gpu):
    meshes, views, depth = W.batch()
    scenes = [ca.Scene(m) for m in meshes]
    jobs = [ca.BatchJob(scenes[s], cam, spp, view_id=v) for s, cam, spp, v in views]
    paths = sum(j.camera.xpixels * j.camera.ypixels * j.spp for j in jobs)
    t0 = time.perf_counter()
    tune, trial_ms = ca.autotune(jobs, max_depth=depth)
    tune_s = time.perf_counter() - t0
    imgs, st = ca.render_batch(jobs, tune=tune, max_depth=depth)
    batch_ms = st.last_render_ms
    imgs0, st0 = ca.render_batch(jobs, max_depth=depth)  # default queue, grouping and occupancy
    for a, b in zip(imgs, imgs0):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # the tuned and default runs again, interleaved in alternating order, best of
    # three each (the first run after the trial renders is often a few percent slower)
    tuned_ms, default_ms = [batch_ms], [st0.last_render_ms]
    for rep in range(2):
        for which in ((0, 1) if rep else (1, 0)):
            if which == 0:
                tuned_ms.append(ca.render_batch(jobs, tune=tune, max_depth=depth)[1].last_render_ms)
            else:
                default_ms.append(ca.render_batch(jobs, max_depth=depth)[1].last_render_ms)
    batch_ms, dflt = min(tuned_ms), min(default_ms)
    print(f"\n[batch] {len(jobs)} views, {sum(m.num_triangles for m in meshes)} triangles in {len(scenes)} scenes, "
          f"{paths / 1e9:.2f} G paths; autotune {tune_s:.1f} s (trial {trial_ms:.1f} ms) -> queue "
          f"{tune.queue_slots}, group {tune.group_side or 'auto'}, blocks/CU {tune.blocks_per_cu or 'max'}; "
          f"batch {batch_ms:.0f} ms ({paths / batch_ms / 1e6:.2f} G paths/s); default tuning {dflt:.0f} ms "
          f"(tuned / default {batch_ms / dflt:.3f})", flush=True)
    # The search starts from the default and moves only for a 2% trial gain: the
    # tuned batch is no slower than the default beyond run-to-run spread, which
    # reached 4.3% between two runs of one and the same tuning on this workload.
    assert batch_ms <= 1.08 * dflt
    # one view at a time, same integrator
    t_views = 0.0
    for j in jobs:
        r = ca.ParallelRenderer(j.scene, j.camera.xpixels, j.camera.ypixels, j.spp, camera=j.camera,
                                max_depth=depth, view_id=j.view_id, integrator=ca.INTEGRATOR_WAVEFRONT)
        r.render(j.spp)
        t_views += r.stats().last_render_ms
        r.close()
    print(f"[batch] the same views one call each: {t_views:.0f} ms (batch speed-up {t_views / batch_ms:.2f}x)",
          flush=True)
    assert t_views >= 1.05 * batch_ms
    order = sorted(range(len(jobs)), key=lambda i: jobs[i].camera.xpixels ** 2 * jobs[i].spp)
    for i in order[:8] + order[-1:]:
        j = jobs[i]
        r = ca.ParallelRenderer(j.scene, j.camera.xpixels, j.camera.ypixels, j.spp, camera=j.camera,
                                max_depth=depth, view_id=j.view_id, spp_per_launch=8)
        r.render(j.spp)
        ref = r.radiance()
        r.close()
        assert np.array_equal(imgs[i].view(np.uint32), ref.view(np.uint32)), f"view {i}"
