"""The command-line renderers: option handling and the fixes of main.cpp's defects
(CPU), and the rendered files (prt_cpu here; prt on the GPU in test_gpu_cli.py)."""
import subprocess
from pathlib import Path

import numpy as np

import cudart_amd as ca
import oracle as O
from tests.helpers import read_png_rgba

ROOT = Path(__file__).resolve().parent.parent
PRT = ROOT / "cudart_amd" / "bin" / "prt"
PRT_CPU = ROOT / "oracle" / "build" / "prt_cpu"
CORNELL = ROOT / "cudart_amd" / "assets" / "cornell.obj"


def run(exe, *args, cwd=None):
    return subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=120, cwd=cwd)


def cornell_reference(w, h, spp, depth=5, seed=0x5EED):
    """The same scene through the Python API and the oracle: lambert4SG emitting Ka = 5."""
    m = ca.Mesh.load_obj(CORNELL)
    for i, (nm, mt) in enumerate(zip(m.material_names(), m.materials)):
        if nm == "lambert4SG":
            mt.bsdf = 5
            mt.ka[:] = (5.0, 5.0, 5.0)
            m.set_material(i, mt)
    cam = ca.camera_from_json(CORNELL.with_name("cornell-camera.json"), w, h)
    orc = O.OracleScene(m.desc())
    acc, _ = orc.render(cam, w, h, 0, spp, O.cfg(max_depth=depth, seed=seed))
    return acc, O.tonemap((acc / np.float32(spp)).reshape(-1, 3)).reshape(h, w, 4)


def test_usage_and_bad_arguments():
    r = run(PRT)
    assert r.returncode == 1 and "Usage" in r.stderr
    assert run(PRT, "-s", "abc").returncode == 2
    assert run(PRT, "-w", "0").returncode == 2
    assert run(PRT, "--integrator", "x").returncode == 2
    assert run(PRT, "--emit", "novalue").returncode == 2
    assert run(PRT, "-s").returncode == 2


def test_cpu_flag_is_refused_by_the_hip_binary():
    r = run(PRT, "--cpu", "-f", CORNELL)
    assert r.returncode == 2 and "prt_cpu" in r.stderr


def test_scene_without_emitter_is_refused():
    r = run(PRT_CPU, "-f", CORNELL, "-s", "1", "-w", "8", "-h", "8", "-o", "/dev/null")
    assert r.returncode == 3 and "--emit" in r.stderr
    r = run(PRT_CPU, "-f", CORNELL, "--emit", "nosuch=5", "-o", "/dev/null")
    assert r.returncode == 2 and "nosuch" in r.stderr


def test_prt_cpu_writes_the_oracle_image(tmp_path):
    out, pfm = tmp_path / "c.png", tmp_path / "c.pfm"
    r = run(PRT_CPU, "-f", CORNELL, "-s", "3", "-w", "40", "-h", "32", "--depth", "5", "--emit", "lambert4SG=5",
            "-o", out, "--pfm", pfm)
    assert r.returncode == 0, r.stderr
    assert "Rendered to" in r.stdout and "for 3 samples per pixel" in r.stdout
    acc, rgba = cornell_reference(40, 32, 3)
    np.testing.assert_array_equal(read_png_rgba(out), rgba)
    # PFM: little-endian, bottom row first
    raw = pfm.read_bytes()
    hdr_end = raw.index(b"\n", raw.index(b"\n", raw.index(b"\n") + 1) + 1) + 1
    assert raw[:hdr_end].split()[:3] == [b"PF", b"40", b"32"]
    img = np.frombuffer(raw[hdr_end:], "<f4").reshape(32, 40, 3)[::-1]
    np.testing.assert_array_equal(img, acc / np.float32(3))


def test_progressive_preview_and_default_output_name(tmp_path):
    r = run(PRT_CPU, "-f", CORNELL, "-s", "2", "-w", "16", "-h", "16", "--emit", "lambert4SG=5", "--x",
            cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "cornell.png").exists() and not (tmp_path / "cornell.png.tmp").exists()


def test_default_camera_without_json(tmp_path):
    obj = tmp_path / "box.obj"
    obj.write_text(CORNELL.read_text())
    (tmp_path / "cornell.mtl").write_text(CORNELL.with_suffix(".mtl").read_text())
    r = run(PRT_CPU, "-f", obj, "-s", "1", "-w", "16", "-h", "16", "--emit", "lambert4SG=5", "-o", tmp_path / "b.png",
            "--env", tmp_path / "none.env")
    assert r.returncode == 0, r.stderr
    assert "default camera" in r.stdout
