"""`prt` (HIP) renders the same files as the oracle's `prt_cpu`."""
import numpy as np
import pytest

from tests.helpers import read_png_rgba
from tests.test_cli import CORNELL, PRT, cornell_reference, run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("integrator", ["megakernel", "wavefront"])
def test_prt_writes_the_oracle_image(gpu, tmp_path, integrator):
    out, pfm = tmp_path / "c.png", tmp_path / "c.pfm"
    r = run(PRT, "-f", CORNELL, "-s", "3", "-w", "40", "-h", "32", "--depth", "5", "--emit", "lambert4SG=5",
            "--integrator", integrator, "-o", out, "--pfm", pfm)
    assert r.returncode == 0, r.stdout + r.stderr
    acc, rgba = cornell_reference(40, 32, 3)
    np.testing.assert_array_equal(read_png_rgba(out), rgba)
    raw = pfm.read_bytes()
    img = np.frombuffer(raw[-40 * 32 * 12:], "<f4").reshape(32, 40, 3)[::-1]
    np.testing.assert_array_equal(img, acc / np.float32(3))
