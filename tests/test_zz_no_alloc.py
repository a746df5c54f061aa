"""prt_adjoint_grad_static, the adjoint entry point that allocates nothing, against
prt_adjoint_grad on the same inputs, through the C ABI.

Two calls in a row, the second on a smaller image (so it reuses a smaller part of
the device's static scratch the first call filled): the gradients must equal
prt_adjoint_grad's bit for bit. The first call replays radiance the forward pass
kept; the second traces each sample's primal path inside the replay (the entry
point writes no per-sample buffer), where prt_adjoint_grad renders a per-sample
buffer first; both give the same paths. Also: a host pointer and an image beyond
the scratch's limit are refused with clear errors.
"""
import ctypes as C

import numpy as np
import pytest

import cudart_amd as ca
from cudart_amd import _capi, workloads as W

pytestmark = pytest.mark.gpu


def _err() -> str:
    return (ca.lib().prt_last_error() or b"").decode()


class _DeviceBuffer:
    """Device memory from the HIP runtime libprt.so itself uses (torch bundles another
    one, and two HIP runtimes cannot share a process's GPU), through its C API."""

    def __init__(self, host: np.ndarray):
        ca.lib()
        path = next(l.split()[-1] for l in open("/proc/self/maps")
                    if "libamdhip64.so" in l and "torch" not in l)
        self.hip = C.CDLL(path)
        self.ptr = C.c_void_p()
        n = host.nbytes
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(n)) == 0
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(n), 1) == 0  # H2D

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _grads(r, scene, grad, spp, hess):
    M = scene.num_materials
    ref_a, ref_e = r.adjoint_grad(grad, spp, hess=hess)
    dev = _DeviceBuffer(np.ascontiguousarray(grad, np.float32))
    d_alb = np.zeros((M, 3), np.float32)
    d_em = np.zeros(M, np.float32)
    rc = ca.lib().prt_adjoint_grad_static(r._h, dev.ptr, float(hess), spp, _capi.fptr(d_alb), _capi.fptr(d_em))
    dev.close()
    assert rc == 0, _err()
    return (ref_a, ref_e), (d_alb, d_em)


def _case(w, h, spp, keep):
    wl = W.cornell100k(w, h)
    scene = ca.Scene(wl.mesh)
    r = ca.ParallelRenderer(scene, w, h, spp, camera=wl.camera, max_depth=wl.max_depth, seed=W.SEED,
                            integrator=ca.INTEGRATOR_WAVEFRONT, keep_samples=int(keep))
    r.render(spp)
    img = r.radiance()
    target = np.clip(img + np.random.default_rng(w).normal(0, 0.05, img.shape), 0, None).astype(np.float32)
    grad = (2.0 * (img.astype(np.float64) - target) / img.size).astype(np.float32)
    return scene, r, grad, 2.0 / img.size


def test_static_scratch_adjoint_equals_the_allocating_one(gpu):
    for (w, h, spp, keep) in ((256, 192, 8, True), (96, 64, 4, False)):
        print(f"case {w}x{h} spp={spp} keep_samples={keep}", flush=True)
        scene, r, grad, hess = _case(w, h, spp, keep)
        (ra, re), (sa, se) = _grads(r, scene, grad, spp, hess)
        assert np.abs(ra).max() > 0 and np.abs(re).max() > 0
        assert np.array_equal(ra, sa) and np.array_equal(re, se), (np.abs(ra - sa).max(), np.abs(re - se).max())
        r.close()


def test_static_scratch_adjoint_refuses_bad_input(gpu):
    scene, r, grad, hess = _case(64, 64, 2, False)
    M = scene.num_materials
    d_alb = np.zeros((M, 3), np.float32)
    d_em = np.zeros(M, np.float32)
    host = np.ascontiguousarray(grad)
    rc = ca.lib().prt_adjoint_grad_static(r._h, C.c_void_p(host.ctypes.data), float(hess), 2,
                                          _capi.fptr(d_alb), _capi.fptr(d_em))
    assert rc == -1 and "device memory" in _err()
    r.close()
    # 4096 x 4096: 262,144 tiles x 4 x 16 parameters is past the 2^21-double scratch.
    wl = W.cornell100k(4096, 4096)
    scene = ca.Scene(wl.mesh)
    big = ca.ParallelRenderer(scene, 4096, 4096, 1, camera=wl.camera, max_depth=wl.max_depth, seed=W.SEED)
    dev = _DeviceBuffer(np.zeros(4096 * 4096 * 3, np.float32))
    rc = ca.lib().prt_adjoint_grad_static(big._h, dev.ptr, 0.0, 1, _capi.fptr(d_alb), _capi.fptr(d_em))
    dev.close()
    assert rc == -9 and "static scratch" in _err()
    big.close()
