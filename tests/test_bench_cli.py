"""bench.py's command line, without a GPU: the one-GPU refusal and the per-workload
defaults (1080p x 128 spp for the 1M-triangle frame, 4K x 512 spp for the large
scene, 256^2 x 16 spp at depth 5 for the Cornell fixture)."""
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
