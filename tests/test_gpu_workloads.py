"""The batched and served workloads at their real sizes, against the oracle.

Uneven batch: one full-size cylindrical MSCT slab (672 x 64 detector, 1152
views, 512 x 512 x 64 volume: its rows leave the 64-row detector, so the LDS
kernel keeps its row clamp) and the first 8 seeded fan-arc slices of the
batch (channels U{300..900}, views U{360..1200}, images U{256..768}^2) in one
graph launch.  Each volume equals its own plan's bit for bit; the slab
matches the oracle's FP32 restatement on slices 0, 31 and 63 and each slice
matches it in full, at the reference comparator's 1e-5.

Serving stream: four seeded requests of the serving mix -- its largest flat
short scan (475^3, 661 views of 760 x 570), a cone-parallel volume, a 2-D
parallel volume and a fan-arc slice -- through a 4-stream server with
reprojection.  Each volume equals what the server gives for the request
alone, bit for bit, and matches the oracle on sampled slices at 1e-5; each
reprojection error equals the one recomputed from the returned volume.
"""
import numpy as np
import pytest

import gct_amd as g
import oracle_lib as O
from gct_amd import workloads as W

pytestmark = pytest.mark.gpu

EPS = 1e-5
NTHREADS = 16


def _report(name, got, want):
    rep = O.parity_report(got, want)
    print(f"[{name}] {rep}", flush=True)
    return rep


def test_uneven_batch_full_size_against_oracle(torch_dev):
    torch = torch_dev
    wls = W.uneven_batch()
    slices = [w for w in wls if g.is_1d(w.scan.kind)][:8]
    slab = [w for w in wls if not g.is_1d(w.scan.kind)][0]
    batch_wls = slices + [slab]
    print(f"\n[uneven batch] {len(slices)} fan-arc slices + slab {slab.img.shape} x {slab.scan.num_proj} views",
          flush=True)
    projs = [w.projections_device() for w in batch_wls]
    torch.cuda.synchronize()
    plans = [g.Plan(w.scan, w.img) for w in batch_wls]
    assert plans[-1].info().rows_inside == 0   # the slab's rows leave the detector: clamped LDS kernel
    own = []
    for pl, p, w in zip(plans, projs, batch_wls):
        v = torch.empty(w.img.shape, dtype=torch.float32, device="cuda")
        pl.execute(p, v)
        own.append(v)
    vols = [torch.full(w.img.shape, -2.0, dtype=torch.float32, device="cuda") for w in batch_wls]
    batch = g.Batch(plans)
    for rep in range(2):
        batch.execute(projs, vols, use_graph=True)
        torch.cuda.synchronize()
        for w, v, o in zip(batch_wls, vols, own):
            assert torch.equal(v, o), f"replay {rep}: {w.name} differs from its plan"
    batch.close()

    for w, v, p in zip(slices, vols, projs):
        want = O.fdk(w.scan, w.img, p.cpu().numpy(), nthreads=NTHREADS)
        _report(f"{w.name} {w.img.shape} x {w.scan.num_proj}", v.cpu().numpy(), want)
        assert O.are_equal(v.cpu().numpy(), want, EPS).all(), w.name
    proj = projs[-1].cpu().numpy()
    got = vols[-1].cpu().numpy()
    for z in (0, 31, 63):
        want = O.fdk(slab.scan, slab.img, proj, z0=z, z1=z + 1, nthreads=NTHREADS)[z]
        _report(f"slab z={z}", got[z], want)
        assert O.are_equal(got[z], want, EPS).all(), z


SERVED = (1, 0, 12, 7)   # flat short 475^3, cone-parallel 192^3, 2-D parallel 141^3, fan-arc 565^2


def test_serving_mix_requests_against_oracle(torch_dev):
    torch = torch_dev
    mix = W.serving_mix()
    reqs = [mix[i] for i in SERVED]
    kinds = {g.KIND_NAMES[w.scan.kind] for w in reqs}
    assert kinds == {"flat", "cone_parallel", "parallel_2d", "fan_arc_1d"}
    assert "short" in reqs[0].name and reqs[0].updates == max(w.updates for w in mix)
    projs = []
    for w in reqs:
        d = w.projections_device()
        projs.append(d.cpu().numpy())
        del d
    print("\n[served] requests:", [w.name for w in reqs], flush=True)

    alone = []
    for w, p in zip(reqs, projs):
        sv = g.Server(num_streams=1, max_inflight=1)
        out = np.zeros(w.img.shape, np.float32)
        sv.wait(sv.submit(w.scan, w.img, p, out))
        sv.close()
        alone.append(out)

    sv = g.Server(num_streams=4, max_inflight=4)
    outs = [np.zeros(w.img.shape, np.float32) for w in reqs]
    tickets = [sv.submit(w.scan, w.img, p, o, reproject=True) for w, p, o in zip(reqs, projs, outs)]
    res = [sv.wait(t) for t in tickets]
    st = sv.stats()
    sv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_inflight"] >= 2

    for w, p, o, a, (lat, err) in zip(reqs, projs, outs, alone, res):
        assert np.array_equal(o, a), f"{w.name}: served in the mix differs from served alone"
        # sampled slices against the oracle's FP32 restatement
        zs = (0,) if g.is_1d(w.scan.kind) else (w.img.nz // 3, w.img.nz - 1)
        for z in zs:
            if g.is_1d(w.scan.kind):
                want, got = O.fdk(w.scan, w.img, p, nthreads=NTHREADS), o
            else:
                want = O.fdk(w.scan, w.img, p, z0=z, z1=z + 1, nthreads=NTHREADS)[z]
                got = o[z]
            _report(f"{w.name} z={z}", got, want)
            assert O.are_equal(got, want, EPS).all(), (w.name, z)
        # the server's reprojection error, recomputed from the volume it returned
        plan = g.Plan(w.scan, w.img)
        re = torch.empty(w.scan.proj_shape, dtype=torch.float32, device="cuda")
        plan.forward_project(torch.from_numpy(o).cuda(), re)
        torch.cuda.synchronize()
        r = re.cpu().numpy().astype(np.float64)
        want_err = float(np.sqrt(((r - p) ** 2).sum() / (p.astype(np.float64) ** 2).sum()))
        print(f"[{w.name}] latency {lat:.1f} ms, reprojection error {err:.5f} (recomputed {want_err:.5f})",
              flush=True)
        assert err == pytest.approx(want_err, rel=1e-4)
        assert 0.0 < err < 0.25
        plan.close()
