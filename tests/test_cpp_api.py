"""The header-only C++ drop-in API (include/gct/gct.hpp) over the C ABI."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gct_amd as g
import oracle_lib as O

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "_build" / "api_test"


def _exe():
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built: run `python -m gct_amd.build --all`")
    return str(EXE)


def test_cpp_api_geometry_known_answers():
    r = subprocess.run([_exe()], capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "OK" in r.stdout


@pytest.mark.gpu
def test_cpp_api_reconstruct_matches_oracle(torch_dev, tmp_path):
    # the same scan the C++ program builds: flat panel 80 x 64 @ 1 mm, 90 views, SOD 200, SDD 400
    scan = g.circular_scan(g.FLAT, 80, 1.0, 90, np.float32(2.0 * np.pi / 90.0), nv=64, dv=1.0, det_row_center=39.5,
                           det_col_center=31.5, src_det_dist=400.0, src_iso_dist=200.0)
    img = g.Image(48, 48, 0.5, 0.5, 48, 0.5)
    ph = np.array([[1.0, -2.0, 0.5, 7.0, 7.0, 7.0, 0, 0, 0, 0.05]], np.float64)
    proj = g.phantom_project(scan, ph)
    proj.tofile(tmp_path / "proj.raw")
    r = subprocess.run([_exe(), "--gpu", str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = np.fromfile(tmp_path / "vol.raw", np.float32).reshape(img.shape)
    want = O.fdk(scan, img, proj)
    print(O.parity_report(got, want))
    assert O.are_equal(got, want, 1e-5).all()
