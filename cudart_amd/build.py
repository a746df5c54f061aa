"""Builds libprt.so (HIP kernels for gfx950 + host C++), the `prt` CLI, and the
CPU oracle (liboracle.so + the `prt_cpu` CLI) in-tree, without a GPU.

Flags that matter for results:
  -ffp-contract=off   no fused multiply-add unless written explicitly, on host
                      and device, so the estimator rounds like the CPU oracle;
  no -ffast-math      HIP keeps correctly rounded fp32 division and sqrt;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keeps the compiler from forming packed-fp32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shlex
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
ORACLE = ROOT / "oracle"
OBJ_DIR = PKG / "_build"
LIB_DIR = PKG / "lib"
BIN_DIR = PKG / "bin"
ORACLE_BUILD = ORACLE / "build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("PRT_OFFLOAD_ARCH", "gfx950")

CXX_FLAGS = ["-O2", "-fPIC", "-fopenmp", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra",
             "-Wno-unused-parameter", f"-I{INCLUDE}"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-parameter", "-Wno-unused-command-line-argument",
             f"-I{INCLUDE}", "-D__HIP_PLATFORM_AMD__"]

HOST_SOURCES = ["host/errors.cpp", "host/mesh_io.cpp", "host/procgen.cpp", "host/bvh.cpp", "host/scene.cpp",
                "host/image_io.cpp"]
HIP_SOURCES = ["kernels/render.hip", "kernels/device.hip", "kernels/wavefront.hip", "kernels/adjoint.hip"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}):\n{shlex.join(cmd)}\n{proc.stdout}")
    out = proc.stdout.strip()
    # hipcc's host pass does not know the packed-fp32 feature: drop that warning only.
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
    if lines:
        print("\n".join(lines), file=sys.stderr)


def _compile(src: Path, obj: Path, hip: bool) -> Path:
    obj.parent.mkdir(parents=True, exist_ok=True)
    if hip:
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [CXX, *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_library(jobs: int | None = None) -> Path:
    """Compiles every source (always) and links cudart_amd/lib/libprt.so."""
    jobs = jobs or min(8, os.cpu_count() or 4)
    tasks = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for s in HOST_SOURCES:
            tasks.append(ex.submit(_compile, CSRC / s, OBJ_DIR / (Path(s).stem + ".o"), False))
        for s in HIP_SOURCES:
            tasks.append(ex.submit(_compile, CSRC / s, OBJ_DIR / (Path(s).stem + ".hip.o"), True))
        objs = [t.result() for t in tasks]
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    lib = LIB_DIR / "libprt.so"
    tmp = LIB_DIR / "libprt.so.tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs), "-lgomp"])
    os.replace(tmp, lib)
    return lib


def build_probe() -> Path:
    """libprt_probe.so: the instruction-issue calibration kernel of bench.py --issue-probe."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    lib = LIB_DIR / "libprt_probe.so"
    tmp = LIB_DIR / "libprt_probe.so.tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", str(CSRC / "probe" / "issue_probe.hip"), "-o", str(tmp)])
    os.replace(tmp, lib)
    return lib


def build_cli() -> Path:
    """The product CLI (`-f -s -w -h`), linked against libprt.so."""
    BIN_DIR.mkdir(parents=True, exist_ok=True)
    exe = BIN_DIR / "prt"
    _run([CXX, *CXX_FLAGS, str(CSRC / "host" / "cli.cpp"), str(CSRC / "host" / "cli_main_gpu.cpp"),
          "-o", str(exe), f"-L{LIB_DIR}", "-lprt", "-Wl,-rpath,$ORIGIN/../lib"])
    return exe


def build_oracle() -> Path:
    """The CPU oracle: liboracle.so (test infrastructure)."""
    ORACLE_BUILD.mkdir(parents=True, exist_ok=True)
    lib = ORACLE_BUILD / "liboracle.so"
    tmp = ORACLE_BUILD / "liboracle.so.tmp"
    _run([CXX, *CXX_FLAGS, "-shared", str(ORACLE / "oracle.cpp"), "-o", str(tmp)])
    os.replace(tmp, lib)
    return lib


def build_oracle_cli() -> Path:
    """`prt_cpu`: the product CLI's front end over the oracle's renderer (test infrastructure)."""
    exe = ORACLE_BUILD / "prt_cpu"
    _run([CXX, *CXX_FLAGS, str(CSRC / "host" / "cli.cpp"), str(ORACLE / "prt_cpu.cpp"), "-o", str(exe),
          f"-L{LIB_DIR}", f"-L{ORACLE_BUILD}", "-lprt", "-loracle",
          f"-Wl,-rpath,$ORIGIN/../../cudart_amd/lib", "-Wl,-rpath,$ORIGIN"])
    return exe


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        f_lib = ex.submit(build_library)
        f_orc = ex.submit(build_oracle)
        f_probe = ex.submit(build_probe)
        f_lib.result()
        f_orc.result()
        f_probe.result()
    build_cli()
    build_oracle_cli()


if __name__ == "__main__":
    build_all()
    print("built", LIB_DIR / "libprt.so", BIN_DIR / "prt", ORACLE_BUILD / "liboracle.so", ORACLE_BUILD / "prt_cpu")
