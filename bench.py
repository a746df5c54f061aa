"""Benchmark: Mrays/s and ms/frame of the path tracer on the 1M-triangle scene,
the megakernel beside it, and the adjoint's grad-samples/s.

Workload (one "step" = one frame): 200 displaced icospheres + ground + 2 emissive
quads (1,024,006 triangles), 1920x1080, 128 spp, depth 8 with Russian roulette
from bounce 3, seed 0x5EED. The scene is generated, its BVH built and uploaded
before timing; the timed region renders K whole frames with the persistent
wavefront integrator (one HIP graph per chunk of paths; accumulation cleared per
frame), bracketed by device synchronisation. `value` = rays traced per second.

Rays = primary + extension + shadow rays, counted by the kernels on an
instrumented frame (the counts of a frame are deterministic for a seed).

Megakernel leg: the same frame with the one-kernel-per-launch integrator; its
image must equal the wavefront's bit for bit (so the RMSE against any reference
is the same), and the ratio of the two rates is reported.

Roofline of the dominant kernel: the extend launches (closest-hit traversal of
every primary and extension ray; wf_extend_packet traces the first iteration's
camera rays 64 at a time as packets, wf_extend the other iterations' rays one per
lane). The traversal is irregular pointer chasing with no dense
contraction, so the roofline is HBM's (peak 8 TB/s). `achieved` = measured HBM
bytes per extend launch (rocprofv3 FETCH_SIZE and WRITE_SIZE passes, each scaled
by the same counter on a plain 1 GiB copy; read from the committed summary under
profiles/) divided by the average extend launch time measured live with HIP
events around every extend launch on the renderer's stream (a frame rendered with
the chunk's kernels launched directly instead of through the graph); `frac` =
achieved / peak. Beside it: the kernel's algorithmic bytes per launch
    64 B x four-wide node visits + 48 B x triangle records tested
    + 48 B x hits (the winning record re-read for u, v, material)
    + 52 B x rays (32 B queue entry read, 20 B hit record written)
(device counters of the instrumented frame, shadow rays removed; a packet's node
and triangle tests count once per lane that makes them, although the packet reads
each node and triangle once), most of which the L2 and the MALL serve; and the instruction-issue fraction: the kernel's VALU
work per launch (committed SQ counter summary) per second of the
live launch time, over the rate a calibration kernel of independent fp32 FMAs
reaches on the same GPU (bench.py --issue-probe, committed summary), counted in
VALU lane operations (SQ_THREAD_CYCLES_VALU: instructions times active lanes).

Adjoint leg: path-replay gradients of the L2 image loss w.r.t. 16 materials'
albedo and emission on the 100k-triangle Cornell-like scene at 512x512, 64
samples replayed (the forward render keeps each sample's radiance): grad-samples/s
= pixels x samples / adjoint device time.

CPU baseline: the oracle (C++ restatement of the reference --cpu renderer, fp32,
OpenMP over every core this process may run on, as the reference's
`omp parallel for` uses all threads) on a bounded sample of pixels of the same
frame, in Mrays/s, with the thread count, the host's CPU count and model.

Image error: RMSE of the timed frame against the oracle's fp64 8,192-sample
estimate of 4,096 fixed pixels (tests/golden/spheres1m_1080p_fp64_reference.npz,
made by tests/golden/make_rmse_reference.py), with that estimate's own noise.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s + ms/frame @1080p 128spp 1M tris; adjoint grad-samples/s; CPU-relative"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "wf_extend_spheres1m_1080p_hbm_traffic.txt"
ISSUE_FILE = ROOT / "profiles" / "wf_extend_wf_shadow_spheres1m_1080p_sq_counters.txt"
RMSE_FILE = ROOT / "tests" / "golden" / "spheres1m_1080p_fp64_reference.npz"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", choices=["spheres1m", "large", "cornell", "batch"], default="spheres1m",
                   help="spheres1m: the 1M-triangle 1080p frame (default); large: ~10M triangles with "
                        "textures at 3840x2160, 512 spp (add --no-compare --no-adjoint to time it alone); "
                        "cornell: the Cornell box fixture, 256x256, 16 spp, depth 5; "
                        "batch: 64 views of 8 scenes (10k-500k triangles) in one autotuned batched call")
    p.add_argument("--width", type=int, default=None)
    p.add_argument("--height", type=int, default=None)
    p.add_argument("--spp", type=int, default=None)
    p.add_argument("--depth", type=int, default=None)
    p.add_argument("--instances", type=int, default=200)
    p.add_argument("--spp-per-launch", type=int, default=4, help="megakernel samples per launch")
    p.add_argument("--integrator", choices=["wavefront", "megakernel"], default="wavefront")
    p.add_argument("--no-compare", action="store_true", help="skip the megakernel leg")
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--threads", type=int, default=0)
    p.add_argument("--frames-only", action="store_true",
                   help="only the timed frames of the headline integrator (for profilers)")
    p.add_argument("--direct", action="store_true",
                   help="with --frames-only: launch the wavefront chunks' kernels directly, not through graphs")
    p.add_argument("--copy-probe", action="store_true",
                   help="run a 1 GiB device-to-device copy 5 times and exit (counter calibration)")
    p.add_argument("--l1-probe", action="store_true",
                   help="run the vector-load gather calibration kernel and exit (L1 / L2 line-rate ceiling)")
    p.add_argument("--issue-probe", action="store_true",
                   help="run the fp32-FMA issue calibration kernel 5 times and exit (counter calibration)")
    a = p.parse_args()
    size = {"spheres1m": (1920, 1080, 128, 8), "large": (3840, 2160, 512, 8), "cornell": (256, 256, 16, 5),
            "batch": (0, 0, 0, 8)}
    w, h, spp, depth = size[a.workload]
    a.width = a.width or w
    a.height = a.height or h
    a.spp = a.spp or spp
    a.depth = a.depth or depth
    return a


def copy_probe():
    import torch

    n = 1 << 28  # 2^28 float32 = 1 GiB read + 1 GiB written per copy
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "launches": 5}))


def issue_probe():
    """The VALU issue rate of independent fp32 FMAs (8 chains per thread, full
    occupancy, no memory traffic): 16,384 workgroups x 4 waves x 4,096 steps x 8 FMAs."""
    import ctypes as C

    # HIP runtime of the in-tree libraries (not torch's), so no torch in this process.
    lib = C.CDLL(str(ROOT / "cudart_amd" / "lib" / "libprt_probe.so"))
    lib.prt_probe_valu.restype = C.c_int
    lib.prt_probe_valu.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]
    blocks, iters, launches = 16384, 4096, 5
    ms = C.c_float()
    rc = lib.prt_probe_valu(blocks, iters, 1, C.byref(ms))  # warm-up
    rc = rc or lib.prt_probe_valu(blocks, iters, launches, C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"issue probe failed: hipError {-rc}")
    fma = blocks * 4 * iters * 8
    print(json.dumps({"issue_probe_fma_instructions_per_launch": fma, "launches": launches,
                      "ms_per_launch": round(ms.value / launches, 4),
                      "fma_instructions_per_s": fma / (ms.value / launches * 1e-3)}))


def l1_probe():
    """16-byte vector loads from pseudo-random 64-byte lines (4 independent loads per
    lane and step, 32 waves per compute unit): buffers of 16 KiB (L1), 1 MiB (L2) and
    64 MiB; the 64 lanes of a load on 64 lines (each lane the four parts of its own
    line), all on one line, 16 to a line, 4 to a line (the four lanes of a quad on the
    four parts of one line), or 64 lines with every other lane masked off. Rates are
    load instructions and active lane loads per compute unit per ns (each lane's 16
    bytes count once)."""
    import ctypes as C

    lib = C.CDLL(str(ROOT / "cudart_amd" / "lib" / "libprt_probe.so"))
    lib.prt_probe_l1.restype = C.c_int
    lib.prt_probe_l1.argtypes = [C.c_uint32] * 5 + [C.POINTER(C.c_float)]
    lib.prt_probe_cu_count.restype = C.c_int
    cus = lib.prt_probe_cu_count()
    if cus <= 0:
        raise RuntimeError(f"l1 probe: no device (hipError {-cus})")
    blocks, iters, launches = cus * 8, 2000, 3
    rows = []
    layouts = ["64 lines", "1 line", "16 lanes per line", "4 lanes per line (quad, one part each)",
               "64 lines, every other lane masked"]
    for kib in (16, 1024, 65536):
        for mode in range(5):
            ms = C.c_float()
            rc = lib.prt_probe_l1(kib * 16, mode, blocks, iters, 1, C.byref(ms))  # warm-up
            rc = rc or lib.prt_probe_l1(kib * 16, mode, blocks, iters, launches, C.byref(ms))
            if rc != 0:
                raise RuntimeError(f"l1 probe failed: {rc}")
            instr = blocks * 4 * iters * 4 * launches  # wave64 load instructions
            lanes = 32 if mode == 4 else 64
            ns = ms.value * 1e6
            rows.append({"buffer_kib": kib, "mode": mode, "layout": layouts[mode], "ms": round(ms.value / launches, 4),
                         "load_instructions_per_cu_per_ns": round(instr / cus / ns, 5),
                         "lane_loads_per_cu_per_ns": round(instr * lanes / cus / ns, 4)})
    print(json.dumps({"l1_probe": rows, "compute_units": cus}))


def host_cores() -> int:
    """Cores this process may run on (the box's CPU share), as OpenMP's default uses."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, target_s: float, threads: int):
    import oracle as O
    from cudart_amd import workloads as W

    orc = O.OracleScene(wl.mesh.desc())
    c = O.cfg(max_depth=wl.max_depth, rr_start=wl.rr_start, seed=W.SEED)
    npx = wl.width * wl.height
    rng = np.random.Generator(np.random.Philox(key=7))
    # Calibrate on a small sample, then size the measured sample (pixels x samples
    # of the same frame) to about target_s of work.
    probe = np.sort(rng.choice(npx, 16384, replace=False)).astype(np.uint32)
    t0 = time.perf_counter()
    orc.render_pixels(wl.camera, wl.width, wl.height, probe, 0, 1, c, threads)
    dt = time.perf_counter() - t0
    want = len(probe) * target_s / max(dt, 1e-4)  # pixel-samples
    spp = int(max(1, min(wl.spp, want // npx)))
    n = int(min(npx, max(16384, want / spp)))
    pix = np.sort(rng.choice(npx, n, replace=False)).astype(np.uint32)
    t0 = time.perf_counter()
    _, cnt = orc.render_pixels(wl.camera, wl.width, wl.height, pix, 0, spp, c, threads)
    dt = time.perf_counter() - t0
    rays = int(cnt.sum())
    frame_ms = dt * (npx / n) * (wl.spp / spp) * 1e3
    return {"value": round(rays / dt / 1e6, 2), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{n} random pixels x samples 0..{spp - 1} of the same {wl.width}x{wl.height} frame: "
                      f"{rays} rays in {dt:.1f} s on {threads} OpenMP threads",
            "ms_per_frame_extrapolated": round(frame_ms, 1),
            "host_cpus": os.cpu_count(), "affinity_cpus": host_cores(), "cpu_model": cpu_model()}


def summary_value(path: Path, key: str):
    """`key = value` from a committed rocprofv3 summary under profiles/, or None."""
    if not path.exists():
        return None
    m = re.search(rf"^{re.escape(key)}\s*=\s*([0-9.eE+-]+)", path.read_text(), re.M)
    return float(m.group(1)) if m else None


def measured_traffic():
    """HBM bytes per extend launch (mean of the frame's 8) from the committed rocprofv3 summary, or None."""
    return summary_value(TRAFFIC_FILE, "traffic_bytes_per_launch")


def rmse_against_reference(image, wl):
    """RMSE of a width x height x 3 image against the committed fp64 reference."""
    if not RMSE_FILE.exists():
        return None
    ref = np.load(RMSE_FILE, allow_pickle=False)
    if (int(ref["width"]), int(ref["height"]), int(ref["max_depth"])) != (wl.width, wl.height, wl.max_depth):
        return None
    px = ref["pixels"]
    got = image.reshape(-1, 3)[px].astype(np.float64)
    err = got - ref["rgb"]
    return {"value": round(float(np.sqrt(np.mean(err ** 2))), 6), "pixels": int(len(px)),
            "reference_spp": int(ref["spp"]), "reference_noise_rms": round(float(np.sqrt(np.mean(ref["stderr"] ** 2))), 6),
            "mean_radiance": round(float(ref["rgb"].mean()), 6), "mean_error": float(err.mean()),
            "reference": "oracle fp64, tests/golden/spheres1m_1080p_fp64_reference.npz"}


def timed_frames(r, spp, warmup, steps):
    for _ in range(warmup):
        r.reset()
        r.render(spp)
    dev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        r.reset()
        r.render(spp)
        dev.append(r.stats().last_render_ms)
    wall = time.perf_counter() - t0
    return wall, float(np.mean(dev))


def adjoint_leg(cpu_threads: int = 0):
    """cpu_threads > 0: also time the oracle's adjoint (primal + replay of every
    pixel-sample, fp32) on 2 samples per pixel of the same image and target."""
    import cudart_amd as ca
    from cudart_amd import workloads as W

    wl = W.cornell100k(512, 512, 64)
    scene = ca.Scene(wl.mesh)
    M = scene.num_materials
    kd0 = np.array([list(m.kd) for m in wl.mesh.materials], np.float32)
    em0 = np.array([m.ka[0] for m in wl.mesh.materials], np.float32)
    rng = np.random.default_rng(1)
    kdT = np.clip(kd0 + rng.uniform(-0.1, 0.1, kd0.shape), 0.05, 0.95).astype(np.float32)
    kw = dict(camera=wl.camera, max_depth=wl.max_depth, integrator=ca.INTEGRATOR_WAVEFRONT)
    rt = ca.ParallelRenderer(scene, 512, 512, 64, seed=W.SEED ^ 0x7A5C3E1, **kw)
    for i in range(M):
        scene.set_material_params(i, kdT[i], float(em0[i]))
    rt.render(64)
    target = rt.radiance()
    rt.close()
    for i in range(M):
        scene.set_material_params(i, kd0[i], float(em0[i]))
    # The forward pass keeps each sample's radiance, so the adjoint only replays paths.
    r = ca.ParallelRenderer(scene, 512, 512, 64, seed=W.SEED, keep_samples=1, **kw)
    r.render(64)
    image = r.radiance()
    r.adjoint(target, 64, image=image)  # warm-up
    times = []
    for _ in range(3):
        r.adjoint(target, 64, image=image)
        times.append(r.stats().last_render_ms)
    ms = float(np.median(times))
    r.close()
    out = {"value": round(512 * 512 * 64 / (ms * 1e-3) / 1e6, 2), "unit": "Mgrad-samples/s",
           "ms": round(ms, 3), "config": {"workload": "cornell100k: 100k-triangle Cornell-like scene, 16 materials",
                                          "width": 512, "height": 512, "spp_replayed": 64, "max_depth": 6,
                                          "parameters": 4 * M}}
    if cpu_threads > 0:
        import oracle as O

        orc = O.OracleScene(wl.mesh.desc())
        c = O.cfg(max_depth=wl.max_depth, rr_start=wl.rr_start, seed=W.SEED)
        spp_cpu = 2
        t0 = time.perf_counter()
        orc.adjoint(wl.camera, 512, 512, spp_cpu, c, image, target, threads=cpu_threads)
        dt = time.perf_counter() - t0
        orc.close()
        out["cpu_baseline"] = {"value": round(512 * 512 * spp_cpu / dt / 1e6, 3), "unit": "Mgrad-samples/s",
                               "cores": cpu_threads, "kind": "port",
                               "sample": f"the oracle's adjoint (primal and replay of every pixel-sample, fp32) on "
                                         f"{spp_cpu} samples per pixel of the same 512x512 image: {dt:.1f} s"}
    return out


def batch_cpu_baseline(meshes, views, depth, target_s: float, threads: int):
    """The oracle on random pixel-samples of every view of the batched workload (each
    view's own scene, camera, view id and depth), sized to about target_s."""
    import oracle as O
    from cudart_amd import workloads as W

    orcs = [O.OracleScene(m.desc()) for m in meshes]
    rng = np.random.Generator(np.random.Philox(key=11))

    def run(budget):
        # per view: up to `budget` pixel-samples, random pixels x samples 0..n-1
        rays, t = 0, 0.0
        for s_idx, cam, spp, vid in views:
            npx = cam.xpixels * cam.ypixels
            pix = np.sort(rng.choice(npx, min(npx, budget), replace=False)).astype(np.uint32)
            ns = int(min(spp, max(1, budget // npx)))
            c = O.cfg(max_depth=depth, rr_start=3, seed=W.SEED, view_id=vid)
            t0 = time.perf_counter()
            _, cnt = orcs[s_idx].render_pixels(cam, cam.xpixels, cam.ypixels, pix, 0, ns, c, threads)
            t += time.perf_counter() - t0
            rays += int(cnt.sum())
        return rays, t

    budget = 256
    rays, dt = run(budget)  # calibration: per-view overheads dominate small budgets, so rescale until
    for _ in range(3):      # the measured sample holds most of target_s
        if dt >= 0.6 * target_s:
            break
        budget = int(max(budget + 1, budget * target_s / max(dt, 1e-3)))
        rays, dt = run(budget)
    return {"value": round(rays / dt / 1e6, 2), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"up to {budget} random pixel-samples of each of the {len(views)} views (own scene, camera "
                      f"and view id): {rays} rays in {dt:.1f} s on {threads} OpenMP threads",
            "host_cpus": os.cpu_count(), "affinity_cpus": host_cores(), "cpu_model": cpu_model()}


def batch_main(a):
    """The batched workload: 64 square views (128..1024 px, 8..256 spp) of 8 seeded
    scenes of 10k-500k triangles, rendered in one autotuned prt_render_batch call."""
    import cudart_amd as ca
    from cudart_amd import workloads as W

    t0 = time.perf_counter()
    meshes, views, depth = W.batch(max_depth=a.depth)
    scenes = [ca.Scene(m) for m in meshes]
    jobs = [ca.BatchJob(scenes[s], cam, spp, view_id=v) for s, cam, spp, v in views]
    build_s = time.perf_counter() - t0
    paths = sum(j.camera.xpixels * j.camera.ypixels * j.spp for j in jobs)
    _, st = ca.render_batch(jobs, max_depth=depth, count_rays=1)  # ray counts of the same work
    rays = st.primary_rays + st.extension_rays + st.shadow_rays
    t0 = time.perf_counter()
    tune, trial_ms = ca.autotune(jobs, max_depth=depth)
    tune_s = time.perf_counter() - t0
    for _ in range(a.warmup):
        ca.render_batch(jobs, tune=tune, max_depth=depth)
    dev = 0.0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        dev += ca.render_batch(jobs, tune=tune, max_depth=depth)[1].last_render_ms
    wall = time.perf_counter() - t0
    value = rays * a.steps / wall / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batch: {len(jobs)} square views (128-1024 px, 8-256 spp) of {len(scenes)} seeded "
                               f"scenes of 10k-500k triangles in one prt_render_batch call, seed 0x5EED",
                   "triangles": [s.num_triangles for s in scenes], "paths": int(paths), "max_depth": depth,
                   "rr_start": 3, "integrator": "wavefront (batched)"},
        "ms_per_batch": round(dev / a.steps, 3),
        "paths_per_s": round(paths * a.steps / wall, 1),
        "rays_per_batch": int(rays),
        "rays_by_kind": {"primary": int(st.primary_rays), "extension": int(st.extension_rays),
                         "shadow": int(st.shadow_rays)},
        "tune": {"queue_slots": tune.queue_slots, "group_side": tune.group_side,
                 "blocks_per_cu": tune.blocks_per_cu, "autotune_s": round(tune_s, 2), "trial_ms": round(trial_ms, 2)},
        "roofline": {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None, "achieved": None,
                     "frac": None, "basis": "no HBM counter summary for this workload (profiles/ holds the "
                                            "default frame's)"},
        "build_s": round(build_s, 2),
    }
    if not a.no_compare:
        # The same views, one prt_render call each, same integrator.
        t_views = 0.0
        for j in jobs:
            r = ca.ParallelRenderer(j.scene, j.camera.xpixels, j.camera.ypixels, j.spp, camera=j.camera,
                                    max_depth=depth, view_id=j.view_id, integrator=ca.INTEGRATOR_WAVEFRONT)
            r.render(j.spp)
            t_views += r.stats().last_render_ms
            r.close()
        out["one_call_per_view"] = {"ms": round(t_views, 3), "batch_speedup": round(t_views / (dev / a.steps), 3)}
    if not a.no_cpu:
        threads = a.threads or host_cores()
        out["cpu_baseline"] = batch_cpu_baseline(meshes, views, depth, a.cpu_seconds, threads)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    for s_ in scenes:
        s_.close()
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: the renderer runs on one "
              f"device (the path would shard by sample range and tile across GPUs)", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe()
        return
    if a.issue_probe:
        issue_probe()
        return
    if a.l1_probe:
        l1_probe()
        return
    if a.workload == "batch":
        batch_main(a)
        return
    import cudart_amd as ca
    from cudart_amd import workloads as W

    t0 = time.perf_counter()
    if a.workload == "cornell":
        wl = W.cornell(a.width, a.height, a.spp, a.depth)
        described = "cornell: meshes/cornell.obj (36 triangles), ceiling quad emissive, seed 0x5EED"
    elif a.workload == "large":
        wl = W.large(a.width, a.height, a.spp, a.depth)
        described = ("large: fBm terrain (4M triangles) + 2,000 textured UV spheres (~10M triangles), "
                     "16 1024x1024 RGBA8 textures, 64 materials, seed 0x5EED")
    else:
        wl = W.spheres1m(a.width, a.height, a.spp, a.depth, instances=a.instances)
        described = "spheres1m: 200 displaced L4 icospheres + ground + 2 emissive quads, seed 0x5EED"
    scene = ca.Scene(wl.mesh)
    build_s = time.perf_counter() - t0
    info = scene.bvh_info()
    nb = info.node_read_bytes  # bytes one node visit reads
    wave = a.integrator == "wavefront"
    integ = ca.INTEGRATOR_WAVEFRONT if wave else ca.INTEGRATOR_MEGAKERNEL
    per = max(1, a.spp_per_launch)
    base = dict(camera=wl.camera, max_depth=a.depth, rr_start=wl.rr_start, seed=W.SEED, spp_per_launch=per)

    r = ca.ParallelRenderer(scene, a.width, a.height, a.spp, integrator=integ,
                            profile_kernels=int(a.frames_only and a.direct), **base)
    if a.frames_only:
        wall, dev_ms = timed_frames(r, a.spp, a.warmup, a.steps)
        print(json.dumps({"frames": a.steps, "ms_per_frame": round(dev_ms, 3)}))
        return

    # Instrumented frame: ray / node / triangle counts of one frame.
    rc = ca.ParallelRenderer(scene, a.width, a.height, a.spp, integrator=integ, count_rays=1, **base)
    rc.render(a.spp)
    st = rc.stats()
    rc.close()
    rays = st.primary_rays + st.extension_rays + st.shadow_rays

    wall, dev_ms = timed_frames(r, a.spp, a.warmup, a.steps)
    ms_per_step = wall * 1e3 / a.steps
    value = rays * a.steps / wall / 1e6
    image = r.radiance()
    r.close()

    # Roofline of the dominant kernel.
    if wave:
        rp = ca.ParallelRenderer(scene, a.width, a.height, a.spp, integrator=integ, profile_kernels=1, **base)
        rp.render(a.spp)
        sp = rp.stats()
        rp.close()
        launches = sp.kernel_launches[0]
        launch_ms = sp.kernel_ms[0] / launches
        ext_rays = st.primary_rays + st.extension_rays
        ext_bytes = (nb * (st.nodes_visited - st.anyhit_nodes) + 48 * (st.tris_tested - st.anyhit_tris)
                     + 48 * st.hits + 52 * ext_rays)
        bytes_per_launch = ext_bytes / launches
        kernel = "extend launches (wf_extend_packet + wf_extend)"
        kernel_times = {"extend": round(sp.kernel_ms[0], 2), "shade": round(sp.kernel_ms[1], 2),
                        "shadow": round(sp.kernel_ms[2], 2), "launches_each": launches,
                        "profiled_frame_ms": round(sp.last_render_ms, 2)}
        per_ray = {"nodes": round((st.nodes_visited - st.anyhit_nodes) / ext_rays, 2),
                   "tris": round((st.tris_tested - st.anyhit_tris) / ext_rays, 2)}
    else:
        launches = (a.spp + per - 1) // per
        launch_ms = dev_ms / launches
        bytes_per_launch = (nb * st.nodes_visited + 48 * st.tris_tested + 64 * st.hits) / launches
        kernel = "megakernel"
        kernel_times = None
        per_ray = {"nodes": round(st.nodes_visited / rays, 2), "tris": round(st.tris_tested / rays, 2)}
    algorithmic = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    # The counter summaries are of the default frame (1M scene, 1080p, 128 spp).
    default_frame = (wave and a.workload == "spheres1m" and (a.width, a.height, a.spp, a.depth, a.instances)
                     == (1920, 1080, 128, 8, 200))
    traffic = measured_traffic() if default_frame else None
    roofline = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": traffic, "kernel": kernel,
                "launch_ms": round(launch_ms, 4),
                "algorithmic_bytes_per_launch": int(bytes_per_launch),
                "algorithmic_GBs": round(algorithmic, 1), "per_ray": per_ray, "kernel_ms_per_frame": kernel_times}
    if traffic is not None:
        achieved = traffic / (launch_ms * 1e-3) / 1e9
        roofline.update(achieved=round(achieved, 1), frac=round(achieved / HBM_PEAK_GBS, 4),
                        basis="measured HBM bytes per launch (profiles/) / live launch time",
                        traffic_over_algorithmic=round(traffic / bytes_per_launch, 4))
        lane_ops = summary_value(ISSUE_FILE, "wf_extend_valu_lane_ops_per_launch")
        peak = summary_value(ISSUE_FILE, "issue_probe_valu_lane_ops_per_s")
        lanes = summary_value(ISSUE_FILE, "wf_extend_active_lanes_per_valu")
        if lane_ops and peak:
            rate = lane_ops / (launch_ms * 1e-3)
            roofline["issue"] = {"achieved": round(rate / 1e12, 2), "peak": round(peak / 1e12, 2),
                                 "unit": "T VALU lane operations/s", "frac": round(rate / peak, 4),
                                 "active_lanes_per_valu": lanes,
                                 "peak_source": "calibration kernel of independent fp32 FMAs, all lanes active, on the "
                                                "same GPU: one wave64 VALU instruction per CU per cycle "
                                                "(bench.py --issue-probe; profiles/)"}
    else:
        # No counter summary for this configuration: HBM bytes are not measured, and the
        # algorithmic rate (much of it served by the caches) is no fraction of the HBM peak.
        roofline.update(achieved=None, frac=None,
                        basis="no HBM counter summary for this configuration (profiles/ holds the default "
                              "frame's); see algorithmic_GBs")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": described,
                   "triangles": scene.num_triangles, "width": a.width, "height": a.height, "spp": a.spp,
                   "max_depth": a.depth, "rr_start": 3, "integrator": a.integrator},
        "ms_per_frame": round(dev_ms, 3),
        "rays_per_frame": int(rays),
        "rays_by_kind": {"primary": int(st.primary_rays), "extension": int(st.extension_rays),
                         "shadow": int(st.shadow_rays)},
        "roofline": roofline,
        "bvh": {"nodes": info.num_nodes, "node_read_bytes": nb, "max_depth": info.max_depth,
                "sah_cost": round(info.sah_cost, 2), "build_s": round(build_s, 2)},
    }
    if wave and not a.no_compare:
        rm = ca.ParallelRenderer(scene, a.width, a.height, a.spp, integrator=ca.INTEGRATOR_MEGAKERNEL, **base)
        m_wall, m_dev = timed_frames(rm, a.spp, 1, max(1, min(2, a.steps)))
        m_steps = max(1, min(2, a.steps))
        m_image = rm.radiance()
        rm.close()
        m_value = rays * m_steps / m_wall / 1e6
        out["megakernel"] = {"value": round(m_value, 2), "unit": "Mrays/s", "ms_per_frame": round(m_dev, 3),
                             "spp_per_launch": per, "steps": m_steps,
                             "image_equal_to_wavefront": bool(np.array_equal(image, m_image))}
        out["wavefront_over_megakernel"] = round(value / m_value, 3)
    if a.workload == "spheres1m":
        rm_ = rmse_against_reference(image, wl)
        if rm_ is not None:
            out["rmse_vs_fp64_oracle"] = rm_
    threads = a.threads or host_cores()
    if not a.no_adjoint:
        out["adjoint"] = adjoint_leg(0 if a.no_cpu else threads)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_seconds, threads)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
