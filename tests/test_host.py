"""Host-side behaviour of libprt.so that needs no GPU: loaders, generators, scene
build, error paths and the exported C ABI."""
import ctypes as C
import re
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

import cudart_amd as ca
from cudart_amd import _capi
from cudart_amd import workloads as W
from tests.helpers import OwnedDesc, material, quad

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / "golden"


def test_every_declared_entry_point_is_exported():
    header = (ROOT / "include" / "prt.h").read_text()
    names = sorted(set(re.findall(r"\b(prt_[a-z0-9_]+)\s*\(", header)))
    assert len(names) > 30
    lib = C.CDLL(str(_capi.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert ca.lib().prt_abi_version() == 1


def test_calibration_probe_library_exports_its_kernels():
    """bench.py --issue-probe / --l1-probe load the probe library (no GPU call here)."""
    lib = C.CDLL(str(ROOT / "cudart_amd" / "lib" / "libprt_probe.so"))
    for name in ("prt_probe_valu", "prt_probe_l1", "prt_probe_cu_count"):
        assert hasattr(lib, name), name


def test_cornell_obj_semantics():
    m = ca.Mesh.load_obj(W.ASSETS / "cornell.obj")
    a = m.arrays()
    assert m.num_triangles == 36
    assert m.material_names() == ["initialShadingGroup", "lambert2SG", "lambert3SG", "lambert4SG"]
    mats = m.materials
    assert list(mats[1].kd) == [1.0, 0.0, 0.0] and list(mats[3].ka) == [1.0, 1.0, 1.0]
    assert all(mt.bsdf == ca.DIFFUSE for mt in mats)  # no `type` lines in cornell.mtl
    # first face of the file: f 1/1/1 2/2/2 3/3/3 under usemtl lambert4SG
    np.testing.assert_array_equal(a.positions[0], np.array(
        [[0.267659, 1.074088, 0.267659], [-0.267659, 1.074088, 0.267659], [0.267659, 1.074088, -0.267659]],
        np.float32))
    np.testing.assert_array_equal(a.normals[0], np.tile(np.array([0, -1, 0], np.float32), (3, 1)))
    # second face: f 3/3/3 2/2/2 4/4/4 -> vt3, vt2, vt4
    np.testing.assert_array_equal(a.uvs[1], np.array([[0, 1], [1, 0], [1, 1]], np.float32))
    # the i-th usemtl name goes to the i-th mesh: ceiling light, boxes, green, walls, red
    mats_by_tri = a.material_ids
    assert list(mats_by_tri[:2]) == [3, 3]
    assert set(mats_by_tri[2:26]) == {0} and list(mats_by_tri[26:28]) == [2, 2]
    assert set(mats_by_tri[28:34]) == {0} and list(mats_by_tri[34:36]) == [1, 1]


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_obj_faces_negative_indices_missing_normals_and_polygons(tmp_path):
    _write(tmp_path, "t.mtl", "newmtl A\nKd 0.1 0.2 0.3\ntype SPECULAR\nnewmtl B\nKa 1 1 1\ntype EMISSIVE\n"
                              "diffuse 0.25\nNi 1.4\nnewmtl C\ntype BOGUS\n")
    obj = ("mtllib t.mtl\n"
           "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0\n"
           "vn 0 0 1\n"
           "g one\nusemtl A\n"
           "f -5//1 -4//1 -3//1\n"           # negative indices, v//n
           "f 1 2 3 4\n"                     # quad without normals -> (0,1,3),(1,2,3)
           "g two\nusemtl B\n"
           "f 1 2 3 5 4\n"                   # pentagon -> fan
           "g three\nusemtl C\n"
           "f 1/1 2/1 3/1\n")
    _write(tmp_path, "t.obj", obj.replace("f 1/1 2/1 3/1", "vt 0.5 0.25\nf 1/1 2/1 3/1"))
    m = ca.Mesh.load_obj(tmp_path / "t.obj")
    a = m.arrays()
    assert m.num_triangles == 1 + 2 + 3 + 1
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 1.5, 0]], np.float32)
    np.testing.assert_array_equal(a.positions[0], P[[0, 1, 2]])
    np.testing.assert_array_equal(a.positions[1], P[[0, 1, 3]])
    np.testing.assert_array_equal(a.positions[2], P[[1, 2, 3]])
    np.testing.assert_array_equal(a.positions[3], P[[0, 1, 2]])
    np.testing.assert_array_equal(a.positions[4], P[[0, 2, 4]])
    np.testing.assert_array_equal(a.positions[5], P[[0, 4, 3]])
    # missing normals: cross(p0 - p1, p2 - p1) of the face's first three vertices, unnormalised
    n = np.cross(P[0] - P[1], P[2] - P[1])
    np.testing.assert_array_equal(a.normals[1], np.tile(n, (3, 1)))
    np.testing.assert_array_equal(a.normals[0], np.tile([0, 0, 1], (3, 1)))
    np.testing.assert_array_equal(a.uvs[6], np.tile([0.5, 0.25], (3, 1)))
    names = m.material_names()
    assert names == ["A", "B", "C"]
    mats = m.materials
    assert mats[0].bsdf == ca.SPECULAR and mats[1].bsdf == ca.EMISSIVE and mats[2].bsdf == ca.DIFFUSE
    assert mats[0].ni == 0.0  # objl::Material() default Ni = 0 when the MTL has none (F18)
    assert mats[1].ni == pytest.approx(1.4) and mats[1].diffuse_coef == 0.25
    assert list(a.material_ids) == [0, 0, 0, 1, 1, 1, 2]


def test_obj_errors(tmp_path):
    with pytest.raises(ca.PrtError) as e:
        ca.Mesh.load_obj(tmp_path / "missing.obj")
    assert e.value.code == -6
    _write(tmp_path, "bad.obj", "v 0 0 0\nf 1 2 3\n")
    with pytest.raises(ca.PrtError) as e:
        ca.Mesh.load_obj(tmp_path / "bad.obj")
    assert e.value.code == -7
    with pytest.raises(ca.PrtError):
        ca.Mesh.load_obj(tmp_path / "x.txt")


def test_camera_json_fixture():
    cam = ca.camera_from_json(GOLDEN / "monkeys-camera.json", 640, 480)
    # focus distance is |viewDirection|, which the file itself records as focusDistance
    assert cam.focus_distance == pytest.approx(4.153697583460459, rel=1e-6)
    assert cam.fov == pytest.approx(np.tan(np.radians(58.71550708558254) / 2), rel=1e-6)
    assert cam.aperture_width == pytest.approx(0.032)  # focalLength / fStop
    assert cam.aspect == pytest.approx(640 / 480)
    d, u, r = (np.array(x, np.float64) for x in (cam.dir, cam.up, cam.right))
    for v in (d, u, r):
        assert np.linalg.norm(v) == pytest.approx(1.0, abs=1e-6)
    np.testing.assert_allclose(r, np.cross(d, u) / np.linalg.norm(np.cross(d, u)), atol=1e-6)
    assert (cam.xpixels, cam.ypixels) == (640, 480)


def test_camera_json_errors(tmp_path):
    with pytest.raises(ca.PrtError) as e:
        ca.camera_from_json(tmp_path / "nope.json", 4, 4)
    assert e.value.code == -6
    _write(tmp_path, "bad.json", '{"fieldOfView": 30, "eye": [1, 2]')
    with pytest.raises(ca.PrtError) as e:
        ca.camera_from_json(tmp_path / "bad.json", 4, 4)
    assert e.value.code == -7


def test_env_loader(tmp_path):
    _write(tmp_path, ".env", "# comment\nMESHES_PATH=../meshes\nCAMERA_PATH=../settings\nTEXTURES_PATH=../textures\n")
    bufs = [C.create_string_buffer(256) for _ in range(3)]
    _capi.check(ca.lib().prt_env_load(str(tmp_path / ".env").encode(), *bufs, 256))
    assert [b.value for b in bufs] == [b"../meshes", b"../settings", b"../textures"]
    assert ca.lib().prt_env_load(str(tmp_path / "none").encode(), *bufs, 256) == -6


def test_procedural_scene_sizes_and_determinism():
    a = ca.Mesh.generate("spheres", "", 0x5EED)
    assert a.num_triangles == 200 * 5120 + 2 + 4
    b = ca.Mesh.generate("spheres", "", 0x5EED)
    c = ca.Mesh.generate("spheres", "", 0x5EEE)
    pa, pb, pc = a.arrays().positions, b.arrays().positions, c.arrays().positions
    assert np.array_equal(pa, pb) and not np.array_equal(pa, pc)
    bsdf = [m.bsdf for m in a.materials[:32]]
    assert bsdf.count(ca.DIFFUSE) == 19 and bsdf.count(ca.SPECULAR) == 8 and bsdf.count(ca.REFRACTIVE) == 5
    kd = np.array([list(m.kd) for m in a.materials[:32]])
    assert kd.min() >= 0.1 and kd.max() <= 0.8
    assert a.materials[33].bsdf == ca.EMISSIVE and a.materials[33].ka[0] == 5.0
    # displaced icosphere vertices stay within r +- 0.05 of their centre
    c4 = ca.Mesh.generate("cornell-spheres", "", 1)
    assert 95_000 < c4.num_triangles < 105_000 and len(c4.materials) == 16


def test_scene_validation():
    cam = ca.camera_look_at((0, 1, 3), (0, 0, 0), (0, 1, 0), 40, 8, 8)
    p, n = quad([-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1])
    od = OwnedDesc(p, n, [0, 0], [material(0)], cam)
    s = ca.Scene(od.desc)
    assert s.lights() == (0, 0.0)
    with pytest.raises(ca.PrtError) as e:  # reference divides by a zero light area (F2)
        ca.ParallelRenderer(s, 8, 8, 1)
    assert e.value.code == -4
    bad = OwnedDesc(p, n, [0, 5], [material(0)], cam)
    with pytest.raises(ca.PrtError) as e:
        ca.Scene(bad.desc)
    assert e.value.code == -1


def test_renderer_rejects_camera_size_mismatch():
    wl = W.cornell(32, 32)
    s = ca.Scene(wl.mesh)
    with pytest.raises(ca.PrtError) as e:
        ca.ParallelRenderer(s, 64, 32, 1, camera=wl.camera)
    assert e.value.code == -1


def test_scene_refuses_non_finite_vertex_data():
    cam = ca.camera_look_at((0, 1, 3), (0, 0, 0), (0, 1, 0), 40, 8, 8)
    p, n = quad([-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1])
    for field, value in (("positions", np.nan), ("normals", np.inf), ("uvs", -np.inf)):
        od = OwnedDesc(p, n, [0, 1], [material(0), material(5, ka=(1, 1, 1))], cam)
        getattr(od, field)[1, 2] = value  # the arrays the description points at
        with pytest.raises(ca.PrtError) as e:
            ca.Scene(od.desc)
        assert e.value.code == -1 and "not finite" in str(e.value)


def test_without_a_device_calls_fail_cleanly():
    """No HIP device is visible here: creating a renderer returns PRT_ERR_NO_DEVICE
    with a message; null arguments return PRT_ERR_INVALID_ARG. Nothing crashes."""
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a HIP device is visible")
    wl = W.cornell(16, 16)
    s = ca.Scene(wl.mesh)
    with pytest.raises(ca.PrtError) as e:
        ca.ParallelRenderer(s, 16, 16, 1, camera=wl.camera)
    assert e.value.code == -8 and str(e.value)
    L = _capi.lib()
    h = C.c_void_p()
    assert L.prt_scene_create(None, C.byref(h)) == -1 and L.prt_last_error()
    assert L.prt_adjoint_grad(None, None, 0.0, 1, None, None) == -1
    assert L.prt_render_spp(None, 1) == -1


def test_png_writer_roundtrip(tmp_path):
    img = np.random.default_rng(1).integers(0, 256, (5, 7, 4), dtype=np.uint8)
    ca.write_png(tmp_path / "a.png", img)
    data = (tmp_path / "a.png").read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        ln = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + ln]
        crc = struct.unpack(">I", data[pos + 8 + ln:pos + 12 + ln])[0]
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF
        chunks[typ] = body
        pos += 12 + ln
    w, h = struct.unpack(">II", chunks[b"IHDR"][:8])
    assert (w, h) == (7, 5)
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = np.frombuffer(raw, np.uint8).reshape(5, 1 + 7 * 4)
    assert (rows[:, 0] == 0).all()
    assert np.array_equal(rows[:, 1:].reshape(5, 7, 4), img)


def test_pfm_writer(tmp_path):
    img = np.arange(2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3)
    ca.write_pfm(tmp_path / "a.pfm", img)
    data = (tmp_path / "a.pfm").read_bytes()
    head = b"PF\n3 2\n-1.0\n"
    assert data.startswith(head)
    body = np.frombuffer(data[len(head):], "<f4").reshape(2, 3, 3)
    assert np.array_equal(body[::-1], img)
