"""Benchmark: Mrays/s and ms/frame of the path tracer on the 1M-triangle scene.

Workload (one "step" = one frame): 200 displaced icospheres + ground + 2 emissive
quads (1,024,006 triangles), 1920x1080, 128 spp, depth 8 with Russian roulette
from bounce 3, seed 0x5EED. The scene is generated and its BVH built and uploaded
before timing; the timed region renders K whole frames (accumulation buffer
cleared per frame), bracketed by device synchronisation.

Rays = primary + extension + shadow rays traced, counted by the kernel on an
instrumented warm-up frame (the counts of a frame are deterministic for a seed).

Roofline: the dominant kernel is the megakernel launch (spp_per_launch samples of
every pixel). Its algorithmic bytes per launch are
    64 B x child-pair nodes fetched + 48 B x triangle records fetched + 64 B x
    shading records fetched,
counted by the kernel; achieved = bytes / (average launch time measured with HIP
events on the renderer's stream), peak = 8 TB/s HBM3E.

CPU baseline: the oracle (C++ restatement of the reference --cpu renderer, fp32,
OpenMP) on rank 0 over a fixed sample of pixels of the same frame at 1 spp,
reported in Mrays/s with the thread count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s + ms/frame @1080p 128spp 1M tris; adjoint grad-samples/s; CPU-relative"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--spp", type=int, default=128)
    p.add_argument("--depth", type=int, default=8)
    p.add_argument("--instances", type=int, default=200)
    p.add_argument("--spp-per-launch", type=int, default=4)
    p.add_argument("--integrator", choices=["megakernel", "wavefront"], default="megakernel")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--threads", type=int, default=0)
    return p.parse_args()


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
    used = threads if threads > 0 else (os.cpu_count() or 1)
    frame_ms = dt * (npx / n) * (wl.spp / spp) * 1e3
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": used, "kind": "port",
            "sample": f"{n} random pixels x samples 0..{spp - 1} of the same {wl.width}x{wl.height} frame: "
                      f"{rays} rays in {dt:.1f} s on {used} OpenMP threads",
            "ms_per_frame_extrapolated": round(frame_ms, 1)}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the path shards by sample range, "
              f"see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import cudart_amd as ca
    from cudart_amd import workloads as W

    t0 = time.perf_counter()
    wl = W.spheres1m(a.width, a.height, a.spp, a.depth, instances=a.instances)
    scene = ca.Scene(wl.mesh)
    build_s = time.perf_counter() - t0
    info = scene.bvh_info()
    integ = ca.INTEGRATOR_WAVEFRONT if a.integrator == "wavefront" else ca.INTEGRATOR_MEGAKERNEL
    per = max(1, a.spp_per_launch)
    launches = (a.spp + per - 1) // per

    # Instrumented warm-up frame: ray / node / triangle counts of one frame.
    rc = ca.ParallelRenderer(scene, a.width, a.height, a.spp, camera=wl.camera, max_depth=a.depth, seed=W.SEED,
                             integrator=integ, spp_per_launch=per, count_rays=1)
    rc.render(a.spp)
    st = rc.stats()
    counted_ms = st.last_render_ms
    rays = st.primary_rays + st.extension_rays + st.shadow_rays
    alg_bytes = 64 * st.nodes_visited + 48 * st.tris_tested + 64 * st.hits
    rc.close()

    r = ca.ParallelRenderer(scene, a.width, a.height, a.spp, camera=wl.camera, max_depth=a.depth, seed=W.SEED,
                            integrator=integ, spp_per_launch=per)
    for _ in range(a.warmup):
        r.reset()
        r.render(a.spp)
    frame_ms = []
    t_start = time.perf_counter()
    for _ in range(a.steps):
        r.reset()
        r.render(a.spp)
        frame_ms.append(r.stats().last_render_ms)
    wall = time.perf_counter() - t_start
    dev_ms = float(np.mean(frame_ms))
    ms_per_step = wall * 1e3 / a.steps
    value = rays * a.steps / wall / 1e6
    launch_ms = dev_ms / launches
    achieved = alg_bytes / launches / (launch_ms * 1e-3) / 1e9

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
        "config": {"workload": "spheres1m: 200 displaced L4 icospheres + ground + 2 emissive quads, seed 0x5EED",
                   "triangles": scene.num_triangles, "width": a.width, "height": a.height, "spp": a.spp,
                   "max_depth": a.depth, "rr_start": 3, "integrator": a.integrator, "spp_per_launch": per},
        "ms_per_frame": round(dev_ms, 3),
        "rays_per_frame": int(rays),
        "rays_by_kind": {"primary": int(st.primary_rays), "extension": int(st.extension_rays),
                         "shadow": int(st.shadow_rays)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "megakernel" if integ == ca.INTEGRATOR_MEGAKERNEL else "wavefront",
                     "launch_ms": round(launch_ms, 4), "bytes_per_launch": int(alg_bytes // launches),
                     "per_ray": {"nodes": round(st.nodes_visited / rays, 2), "tris": round(st.tris_tested / rays, 2)}},
        "bvh": {"nodes": info.num_nodes, "max_depth": info.max_depth, "sah_cost": round(info.sah_cost, 2),
                "build_s": round(build_s, 2)},
        "counted_frame_ms": round(counted_ms, 3),
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_seconds, a.threads or min(16, os.cpu_count() or 1))
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
