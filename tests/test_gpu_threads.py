"""Renderers driven from several host threads at once. Each wavefront render captures
its chunk's launches into a HIP graph on the renderer's own stream; a legacy-stream
copy or fill issued by another thread during that capture makes HIP fail the
capture, so the library keeps every copy and fill on a renderer's stream. Two
threads rendering on two renderers at the same time must give exactly the images
that the same calls give one after the other."""
import threading

import numpy as np
import pytest

import cudart_amd as ca
from cudart_amd import workloads as W

pytestmark = pytest.mark.gpu

CALLS, SPP = 24, 1


def _renderers(scene, wl):
    kw = dict(camera=wl.camera, max_depth=6, integrator=ca.INTEGRATOR_WAVEFRONT)
    return [ca.ParallelRenderer(scene, wl.width, wl.height, CALLS * SPP, seed=W.SEED + k, **kw) for k in range(2)]


def _work(r, errors):
    """Renders, reading the image back after each call (a readback's temporary
    device buffer is allocated and freed while the other thread may be capturing)."""
    try:
        for _ in range(CALLS):
            r.render(SPP)
            r.radiance()
    except Exception as e:  # reported by the main thread
        errors.append(e)


def test_two_threads_render_as_one_after_the_other(gpu):
    wl = W.cornell(96, 64, spp=CALLS * SPP, max_depth=6)
    scene = ca.Scene(wl.mesh)
    seq = _renderers(scene, wl)
    for r in seq:
        _work(r, [])
    want = [r.radiance() for r in seq]
    for r in seq:
        r.close()
    con = _renderers(scene, wl)
    errors = []
    threads = [threading.Thread(target=_work, args=(r, errors)) for r in con]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    got = [r.radiance() for r in con]
    for r in con:
        r.close()
    for a, b in zip(want, got):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert not np.array_equal(want[0], want[1])  # the two views differ (seeds)


def test_two_threads_batch_as_one_after_the_other(gpu):
    """Two prt_render_batch calls at once (each on a stream of its own, each
    capturing its chunk's launches) give the images of the same calls made alone."""
    a = W.cornell(40, 30)
    b = W.spheres1m(50, 34, instances=12)
    sa, sb = ca.Scene(a.mesh), ca.Scene(b.mesh)
    batches = [[ca.BatchJob(sa, a.camera, 3, view_id=1), ca.BatchJob(sb, b.camera, 2, view_id=2)],
               [ca.BatchJob(sb, b.camera, 3, view_id=3), ca.BatchJob(sa, a.camera, 2, view_id=4)]]
    want = [ca.render_batch(jobs, max_depth=6)[0] for jobs in batches]
    got, errors = [None, None], []

    def run(k):
        try:
            for _ in range(8):
                got[k] = ca.render_batch(batches[k], max_depth=6)[0]
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for w, g in zip(want, got):
        for x, y in zip(w, g):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
