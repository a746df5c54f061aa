"""Host code under AddressSanitizer and UBSan (with float-to-int overflow checks), on
the CPU (g++ only, no GPU code):

- the OBJ / MTL / camera-JSON / env loaders fed malformed and random inputs, and
  the procedural scene generators fed malformed and out-of-range parameters
  (tests/host_fuzz_loaders.cpp): every input is loaded or refused with an error
  code, with no invalid access, undefined behaviour or leak;
- the command-line front end: random argument lists through its parser, then its
  whole load / render / save sequence (PNG, PFM, progressive preview, refusals)
  over a backend that returns a fixed image (tests/host_fuzz_cli.cpp);
- the scene API (tests/host_fuzz_scene.cpp): random scenes of every BSDF with
  textures, degenerate triangles and coordinates near the float range must build;
  descriptions with a material, texture or BSDF out of range, an empty texture,
  129 materials, no triangles or non-finite vertex data must be refused;
- the oracle's render paths (tests/host_sanitize_oracle.cpp): generated scenes
  with every BSDF, textures and area lights, fp32 and fp64 renders, first hits
  (BVH and brute force) and both adjoint estimators;
- the BVH builder on random scenes, leaf sizes 1 and 2 and merged leaves of up to
  8 triangles (the SAH collapse's leaves).

The five drivers are compiled in parallel once per module.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HOST = ROOT / "cudart_amd" / "csrc" / "host"
TESTS = ROOT / "tests"
SAN = ["-std=c++17", "-O1", "-g", "-fopenmp", "-fsanitize=address,undefined,float-cast-overflow",
       "-fno-omit-frame-pointer", "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}",
       f"-I{HOST}", f"-I{ROOT / 'oracle'}"]

BVH_DRIVER = r"""
#include "bvh.hpp"
#include <cstdio>
#include <random>
#include <vector>
int main() {
    std::mt19937 g(3);
    std::normal_distribution<float> N(0.f, 1.f);
    for (int trial = 0; trial < 24; ++trial) {
        const uint32_t n = 1 + g() % 6000;
        std::vector<float> p(size_t(n) * 9);
        for (auto& x : p) x = N(g) * (trial % 3 == 0 ? 1e-3f : 1.f);
        for (int leaf = 1; leaf <= 2; ++leaf) {
            prt::BvhParams bp;
            bp.leaf_size = leaf;
            bp.merge_max = 2 + trial % 7;
            prt::BvhLayout out;
            prt::build_bvh(p.data(), n, bp, out);
            if (out.order.size() != n || out.num_nodes == 0) return 1;
        }
    }
    std::puts("ok");
    return 0;
}
"""

# driver name -> sources (the bvh driver's own source is written at build time)
DRIVERS = {
    "loaders": [HOST / "mesh_io.cpp", HOST / "procgen.cpp", HOST / "errors.cpp", TESTS / "host_fuzz_loaders.cpp"],
    "cli": [HOST / "cli.cpp", HOST / "mesh_io.cpp", HOST / "procgen.cpp", HOST / "image_io.cpp", HOST / "errors.cpp",
            TESTS / "host_fuzz_cli.cpp"],
    "scene": [HOST / "scene.cpp", HOST / "bvh.cpp", HOST / "errors.cpp", TESTS / "host_fuzz_scene.cpp"],
    "oracle": [ROOT / "oracle" / "oracle.cpp", HOST / "mesh_io.cpp", HOST / "procgen.cpp", HOST / "errors.cpp",
               TESTS / "host_sanitize_oracle.cpp"],
    "bvh": [HOST / "bvh.cpp"],
}


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    out = tmp_path_factory.mktemp("sanitized")
    (out / "bvh_driver.cpp").write_text(BVH_DRIVER)

    def build(name):
        srcs = DRIVERS[name] + ([out / "bvh_driver.cpp"] if name == "bvh" else [])
        exe = out / name
        p = subprocess.run([gxx, *SAN, *map(str, srcs), "-o", str(exe)], capture_output=True, text=True, timeout=900)
        assert p.returncode == 0, p.stderr[-3000:]
        return name, exe

    with cf.ThreadPoolExecutor(max_workers=min(5, os.cpu_count() or 1)) as ex:
        return dict(ex.map(build, DRIVERS))


def _run(exe, *args, cwd):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    p = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=300, cwd=cwd, env=env)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-3000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out, out[-3000:]
    return out


def test_loaders_under_sanitizers(drivers, tmp_path):
    out = _run(drivers["loaders"], tmp_path, cwd=tmp_path)
    assert "loaded" in out and "refused" in out


def test_cli_under_sanitizers(drivers, tmp_path):
    out = _run(drivers["cli"], tmp_path, ROOT / "cudart_amd" / "assets" / "cornell.obj", cwd=tmp_path)
    assert "run 0" in out
    assert (tmp_path / "out.png").stat().st_size > 0 and (tmp_path / "out.pfm").stat().st_size > 0


def test_scene_api_under_sanitizers(drivers, tmp_path):
    assert "wrong 0" in _run(drivers["scene"], cwd=tmp_path)


def test_oracle_under_sanitizers(drivers, tmp_path):
    assert "oracle ok: 4 scenes" in _run(drivers["oracle"], cwd=tmp_path)


def test_bvh_builder_under_sanitizers(drivers, tmp_path):
    assert _run(drivers["bvh"], cwd=tmp_path).strip().endswith("ok")
