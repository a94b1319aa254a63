"""Precision forms at full size: the short-scan C-arm workload (512^3, 248
views over 200 deg, 616 x 480 detector, Parker weighting, Poisson I0 = 1e5)
with FP32, FP16 and BF16 filtered-projection storage, each against the FP64
oracle on sampled z slices; the FP32 form also against the FP32 restatement
of the reference's operation order at the 1e-5 comparator bar."""
import numpy as np
import pytest

import gct_amd as g
import oracle_lib as O
from gct_amd import workloads as W

pytestmark = pytest.mark.gpu

SLICES = (8, 200, 256, 471)


@pytest.fixture(scope="module")
def carm():
    wl = W.carm_short_scan_512()
    proj = wl.projections(16)
    want = {}
    for f64 in (True, False):
        cosw, red = O.weights(wl.scan, f64=f64)
        weighted = (proj * cosw) * red[:, None, :]
        taps = O.ramp_taps(False, wl.scan.nu, wl.scan.du, 1.0, wl.window)[1 if f64 else 0]
        filt = O.filter_rows(taps, weighted, f64=f64, nthreads=16)
        del weighted, cosw
        want[f64] = {z: O.backproject(wl.scan, wl.img, filt, z0=z, z1=z + 1, f64=f64, nthreads=16)[z]
                     for z in SLICES}
        del filt
    return wl, proj, want[True], want[False]


@pytest.mark.parametrize("form", ["f32", "f16", "bf16"])
def test_carm_short_scan_storage_forms_against_fp64_oracle(torch_dev, carm, form):
    torch = torch_dev
    wl, proj, want64, want32 = carm
    st = {"f32": g.STORE_F32, "f16": g.STORE_F16, "bf16": g.STORE_BF16}[form]
    plan = g.Plan(wl.scan, wl.img, storage=st)
    assert plan.info().setup_f64, "512^3 volumes use the FP64 column set-up"
    v = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
    plan.execute(torch.from_numpy(proj).cuda(), v)
    torch.cuda.synchronize()
    got = np.stack([v[z].cpu().numpy() for z in SLICES])
    ref64 = np.stack([want64[z] for z in SLICES])
    err = got.astype(np.float64) - ref64
    rmse = float(np.sqrt(np.mean(err ** 2)))
    mx = float(np.abs(err).max())
    peak = float(np.abs(ref64).max())
    print(f"\n[carm 512^3 {form}] vs FP64 oracle on z={SLICES}: rmse={rmse:.3e} max={mx:.3e} peak={peak:.3e}",
          flush=True)
    assert rmse < 1e-3
    if form == "f32":
        ref32 = np.stack([want32[z] for z in SLICES])
        rep = O.parity_report(got, ref32)
        print(f"[carm 512^3 f32] vs FP32 restatement: {rep}", flush=True)
        assert O.are_equal(got, ref32, 1e-5).all()
