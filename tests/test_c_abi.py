"""The C ABI from C: `include/prt.h` compiles as C99 with warnings as errors, and a C
program linked against `libprt.so` drives the host-only entry points (version,
mesh generation, scene build and BVH statistics, error reporting, image output) as
an FFI binding (cgo, ctypes, JNI) would. No GPU is touched."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "cudart_amd" / "lib"

C_SOURCE = r"""
#include "prt.h"
#include <stdio.h>
#include <string.h>

int main(void) {
    prt_mesh *mesh = NULL;
    prt_scene_desc desc;
    prt_scene *scene = NULL;
    prt_bvh_info info;
    prt_render_config cfg;
    uint32_t lights = 0;
    float area = 0.f;
    unsigned char rgba[2 * 2 * 4];
    int rc;

    printf("abi %d\n", prt_abi_version());
    prt_render_config_default(&cfg);
    printf("max_depth %u\n", cfg.max_depth);
    rc = prt_mesh_generate("cornell-spheres", "instances=3;level=2;width=32;height=24", 7u, &mesh);
    if (rc != PRT_OK) { printf("generate %d %s\n", rc, prt_last_error()); return 1; }
    if (prt_mesh_get_desc(mesh, &desc) != PRT_OK) return 2;
    printf("triangles %u\n", desc.num_triangles);
    if (prt_scene_create(&desc, &scene) != PRT_OK) { printf("scene %s\n", prt_last_error()); return 3; }
    if (prt_scene_bvh_info(scene, &info) != PRT_OK) return 4;
    printf("nodes %u leaves %u node_read_bytes %u\n", info.num_nodes, info.num_leaves, info.node_read_bytes);
    if (prt_scene_num_lights(scene, &lights, &area) != PRT_OK) return 5;
    printf("lights %u\n", lights);
    rc = prt_mesh_generate("no-such-kind", "", 1u, NULL);
    printf("bad kind %d %s\n", rc == PRT_OK ? 0 : 1, strlen(prt_last_error()) > 0 ? "message" : "none");
    memset(rgba, 128, sizeof rgba);
    if (prt_write_png(OUT_PNG, 2, 2, rgba) != PRT_OK) return 6;
    prt_scene_destroy(scene);
    prt_mesh_destroy(mesh);
    printf("ok\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_c_program_drives_the_abi(tmp_path):
    lib = LIB / "libprt.so"
    if not lib.exists():
        pytest.skip("libprt.so not built")
    src = tmp_path / "abi.c"
    png = tmp_path / "out.png"
    src.write_text(C_SOURCE)
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", f"-I{ROOT / 'include'}",
                    f'-DOUT_PNG="{png}"', str(src), "-o", str(exe), f"-L{LIB}", "-lprt", f"-Wl,-rpath,{LIB}"],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120).stdout
    assert out.strip().endswith("ok"), out
    lines = dict(l.split(" ", 1) for l in out.strip().splitlines() if " " in l)
    assert int(lines["abi"]) >= 1
    assert int(lines["max_depth"]) > 0
    assert int(lines["triangles"]) > 12
    assert lines["nodes"].split()[0] != "0" and "node_read_bytes 64" in out
    assert int(lines["lights"]) == 2
    assert lines["bad"] == "kind 1 message"
    assert png.read_bytes()[:8] == b"\x89PNG\r\n\x1a\n"
