"""The named workloads the path is measured and checked on (all seeded, synthetic).

  cornell   the reference's meshes/cornell.obj (36 triangles) with the ceiling
            quad `lambert4SG` made EMISSIVE (radiance 15) and albedos brought
            into [0.1, 0.8]; pinhole camera into the open +x side; 256x256,
            16 spp, depth 5. The reference's own cornell has no emitter and no
            camera file, so it cannot render as shipped.
  spheres1m 200 displaced icospheres (5,120 triangles each) over a ground quad,
            2 emissive quads, 32 materials (19 diffuse / 8 specular / 5
            refractive IOR 1.5); 1,024,006 triangles; 1920x1080, 128 spp,
            depth 8 with Russian roulette from bounce 3.
  cornell100k  closed Cornell-like box with 19 displaced icospheres (97,292
            triangles), 16 materials; 512x512, 64 spp, depth 6.
  batch     8 scenes of 10k-500k triangles, 64 views of 128..1024 pixels per
            side and 8/32/128/256 spp.
  large     fBm terrain + 2,000 UV spheres with 16 RGBA8 textures, 64 materials;
            3840x2160, 512 spp, depth 8.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import EMISSIVE, Camera, Material, Mesh, camera_look_at

ASSETS = Path(__file__).resolve().parent / "assets"
SEED = 0x5EED


@dataclass
class Workload:
    name: str
    mesh: Mesh
    width: int
    height: int
    spp: int
    max_depth: int
    camera: Camera
    rr_start: int = 3
    extra: dict = field(default_factory=dict)


def cornell(width: int = 256, height: int = 256, spp: int = 16, max_depth: int = 5) -> Workload:
    m = Mesh.load_obj(ASSETS / "cornell.obj")
    names = m.material_names()
    mats = m.materials
    albedo = {"initialShadingGroup": (0.5, 0.5, 0.5), "lambert2SG": (0.8, 0.1, 0.1),
              "lambert3SG": (0.1, 0.8, 0.1), "lambert4SG": (0.8, 0.8, 0.8)}
    for i, (nm, mt) in enumerate(zip(names, mats)):
        if nm in albedo:
            mt.kd[:] = albedo[nm]
        if nm == "lambert4SG":
            mt.bsdf = EMISSIVE
            mt.ka[:] = (5.0, 5.0, 5.0)  # radiance LIGHTS_GAIN * Ka = 15
        m.set_material(i, mt)
    cam = camera_look_at((2.2, 0.548, 0.0), (0.0, 0.548, 0.0), (0.0, 1.0, 0.0), 39.3, width, height)
    m.set_camera(cam)
    return Workload("cornell", m, width, height, spp, max_depth, cam)


def spheres1m(width: int = 1920, height: int = 1080, spp: int = 128, max_depth: int = 8, seed: int = SEED,
              instances: int = 200) -> Workload:
    m = Mesh.generate("spheres", f"instances={instances};level=4;width={width};height={height}", seed)
    return Workload("spheres1m", m, width, height, spp, max_depth, m.camera)


def cornell100k(width: int = 512, height: int = 512, spp: int = 64, max_depth: int = 6, seed: int = SEED) -> Workload:
    m = Mesh.generate("cornell-spheres", f"width={width};height={height}", seed)
    return Workload("cornell100k", m, width, height, spp, max_depth, m.camera)


def large(width: int = 3840, height: int = 2160, spp: int = 512, max_depth: int = 8, seed: int = SEED,
          params: str = "") -> Workload:
    m = Mesh.generate("terrain-spheres", f"width={width};height={height};{params}", seed)
    return Workload("large", m, width, height, spp, max_depth, m.camera)


def batch_plan(seed: int = SEED, views: int = 64, scenes: int = 8):
    """Scene sizes and per-view (scene, side, spp) drawn from a seeded generator."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    tris = np.exp(rng.uniform(np.log(10_000), np.log(500_000), scenes)).astype(int)
    plan = []
    for v in range(views):
        s = int(rng.integers(0, scenes))
        side = int(rng.integers(128, 1025))
        spp = int(rng.choice([8, 32, 128, 256]))
        plan.append((s, side, spp))
    return [int(t) for t in tris], plan


def batch(seed: int = SEED, views: int = 64, scenes: int = 8, max_depth: int = 8):
    """The batched workload: `scenes` sphere scenes of batch_plan's sizes and `views`
    square views of them from cameras on a circle around the scene centre, each with
    its own resolution, spp and view id. Returns (meshes, [(scene index, Camera, spp,
    view_id)], max_depth)."""
    tris, plan = batch_plan(seed, views, scenes)
    meshes = [Mesh.generate("mixed", f"tris={t}", seed + 101 * i) for i, t in enumerate(tris)]
    jobs = []
    for v, (s, side, spp) in enumerate(plan):
        th = 2.0 * np.pi * v / views
        eye = (35.0 * np.sin(th), 12.0 + 4.0 * np.cos(3.0 * th), 35.0 * np.cos(th))
        cam = camera_look_at(eye, (0.0, 8.0, 0.0), (0.0, 1.0, 0.0), 45.0, side, side)
        jobs.append((s, cam, spp, v + 1))
    return meshes, jobs, max_depth
