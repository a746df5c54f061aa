"""bench.py's command line, without a GPU: the one-GPU refusal, the per-workload
defaults (1080p x 128 spp for the 1M-triangle frame, 4K x 512 spp for the large
scene, 256^2 x 16 spp at depth 5 for the Cornell fixture), the CPU baselines and
the committed counter summaries it reads."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


@pytest.mark.parametrize("workload,expect", [("spheres1m", (1920, 1080, 128, 8)), ("large", (3840, 2160, 512, 8)),
                                             ("cornell", (256, 256, 16, 5))])
def test_workload_defaults(monkeypatch, workload, expect):
    sys.path.insert(0, str(ROOT))
    import bench

    monkeypatch.setattr(sys, "argv", ["bench.py", "--workload", workload])
    a = bench.parse()
    assert (a.width, a.height, a.spp, a.depth) == expect
    monkeypatch.setattr(sys, "argv", ["bench.py", "--workload", workload, "--spp", "4", "--depth", "3"])
    a = bench.parse()
    assert (a.spp, a.depth) == (4, 3)


def test_cpu_baseline_uses_every_allowed_core_and_summaries_parse(tmp_path):
    sys.path.insert(0, str(ROOT))
    import os

    import bench

    assert bench.host_cores() == len(os.sched_getaffinity(0))
    assert isinstance(bench.cpu_model(), str) and bench.cpu_model()
    f = tmp_path / "summary.txt"
    f.write_text("# comment\nalpha = 12.5\nbeta=3e9\n")
    assert bench.summary_value(f, "alpha") == 12.5
    assert bench.summary_value(f, "beta") == 3e9
    assert bench.summary_value(f, "gamma") is None
    assert bench.summary_value(tmp_path / "missing.txt", "alpha") is None
    # the committed summaries the default bench line reads carry their keys
    assert bench.summary_value(bench.TRAFFIC_FILE, "traffic_bytes_per_launch") > 1e9
    assert bench.summary_value(bench.ISSUE_FILE, "wf_extend_valu_lane_ops_per_launch") > 1e10
    assert bench.summary_value(bench.ISSUE_FILE, "issue_probe_valu_lane_ops_per_s") > 1e12


def test_rmse_reference_fixture_and_metric():
    sys.path.insert(0, str(ROOT))
    import numpy as np

    import bench
    from cudart_amd import workloads as W

    ref = np.load(bench.RMSE_FILE, allow_pickle=False)
    assert ref["rgb"].shape == (len(ref["pixels"]), 3) and int(ref["spp"]) == 8192
    assert (int(ref["width"]), int(ref["height"])) == (1920, 1080)

    class Wl:  # the default frame's geometry, without building the scene
        width, height, max_depth = 1920, 1080, 8

    img = np.zeros((1080, 1920, 3), np.float32)
    img.reshape(-1, 3)[ref["pixels"]] = ref["rgb"].astype(np.float32)
    r = bench.rmse_against_reference(img, Wl)
    assert r["value"] < 1e-6 and r["pixels"] == len(ref["pixels"])
    img.reshape(-1, 3)[ref["pixels"]] += np.float32(0.01)
    assert abs(bench.rmse_against_reference(img, Wl)["value"] - 0.01) < 1e-6
    assert W.SEED == int(ref["seed"])


def test_batch_cpu_baseline_covers_every_view():
    """The batched workload's CPU leg: the oracle on pixel-samples of all 64 views,
    each with its own scene, camera and view id (a short sample here)."""
    sys.path.insert(0, str(ROOT))
    import bench
    from cudart_amd import workloads as W

    meshes, views, depth = W.batch()
    r = bench.batch_cpu_baseline(meshes, views, depth, 0.2, 2)
    assert r["kind"] == "port" and r["cores"] == 2 and r["value"] > 0
    assert f"each of the {len(views)} views" in r["sample"]
