#!/usr/bin/env python3
"""FDK benchmark on one MI355X.

A step is one plain FDK reconstruction of the flat-cone workload (256^3 at
0.25 mm, 512 views over 360 deg, 512 x 384 detector at 0.4 mm, SOD 500 mm,
SDD 1000 mm, Shepp-Logan-windowed ramp, 12 seeded ellipsoids with Poisson
noise at I0 = 1e5): fused weighting + FFT ramp filter, then backprojection,
with the projections already resident in HBM.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

F_ALG_FLAT = 45          # FP32 operations per voxel x view update of the flat-cone mapping + bilinear + accumulate
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def _percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    lo, hi = int(k), min(int(k) + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def cpu_baseline(wl, proj, gpu_vol, nthreads):
    """The CPU oracle (FP32 restatement of the reference path) timed on this
    host: weighting + ramp filter of every view, backprojection of a sample of
    z slices, scaled to the whole volume.  Also reports how the GPU volume
    agrees with the oracle on those slices."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np
    import oracle_lib as O

    scan, img = wl.scan, wl.img
    nsample = 32
    z0 = img.nz // 2 - nsample // 2
    t = time.perf_counter()
    cosw, red = O.weights(scan)
    weighted = (proj * cosw) * red[:, None, :]
    t32, _ = O.ramp_taps(False, scan.nu, scan.du, 1.0, wl.window)
    filt = O.filter_rows(t32, weighted, nthreads=nthreads)
    t_filter = time.perf_counter() - t
    t = time.perf_counter()
    vol = O.backproject(scan, img, filt, z0=z0, z1=z0 + nsample, nthreads=nthreads)
    t_bp = time.perf_counter() - t
    full = t_filter + t_bp * img.nz / nsample
    rep = O.parity_report(gpu_vol[z0:z0 + nsample], vol[z0:z0 + nsample])
    return {
        "value": round(wl.updates / full / 1e9, 4),
        "unit": "GUPS",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"oracle FP32: weighting+filter of all {scan.num_proj} views ({t_filter:.2f} s) + "
                   f"backprojection of z slices [{z0},{z0 + nsample}) of {img.nz} ({t_bp:.2f} s), "
                   f"scaled to the full volume"),
        "bp_gups": round(nsample * img.nx * img.ny * scan.num_proj / t_bp / 1e9, 4),
        "gpu_vs_oracle_on_sample": rep,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy 5 times (reference kernel for HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (the FDK path here "
              "runs on one device; DESIGN.md describes how it would shard).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gct_amd as g
    from gct_amd import workloads as W

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.copy_calibration:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_launch": {"read": x.numel() * 4, "write": y.numel() * 4}}))
        return

    wl = W.flat_cone_256()
    t = time.perf_counter()
    proj = wl.projections(a.threads)
    t_gen = time.perf_counter() - t
    p = torch.from_numpy(proj).cuda()
    v = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
    plan = g.Plan(wl.scan, wl.img, window=wl.window)
    info = plan.info()
    stream = torch.cuda.current_stream()

    for _ in range(a.warmup):
        plan.execute(p, v, stream=stream)
    torch.cuda.synchronize()

    K = a.steps
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(K):
        ev[i][0].record(stream)
        plan.filter(p, stream=stream)
        ev[i][1].record(stream)
        plan.backproject(v, stream=stream)
        ev[i][2].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()

    ms_step = (t1 - t0) * 1e3 / K
    filt_ms = [ev[i][0].elapsed_time(ev[i][1]) for i in range(K)]
    bp_ms = [ev[i][1].elapsed_time(ev[i][2]) for i in range(K)]
    step_ms = [ev[i][0].elapsed_time(ev[i][2]) for i in range(K)]
    upd = wl.updates
    bp_mean = statistics.mean(bp_ms)
    gups = upd / (ms_step * 1e-3) / 1e9

    achieved_tflops = F_ALG_FLAT * upd / (bp_mean * 1e-3) / 1e12
    alg_bytes = wl.scan.num_proj * wl.scan.nv * wl.scan.nu * 4 + wl.img.nvox * 4
    traffic = None
    tfile = ROOT / "profiles" / "bp_flat_cone_256_traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    vol = v.cpu().numpy()
    truth = g.phantom_rasterize(wl.img, wl.phantom)
    rmse_truth = float(np.sqrt(np.mean((vol - truth) ** 2)))

    out = {
        "metric": "backprojection GUPS + p50/p99 recon latency, served vs plain FDK, at fixed RMSE",
        "value": round(gups, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic ellipsoid line integrals, Poisson I0=1e5)",
        "config": {
            "workload": "flat_cone_256: 256^3 vox @0.25 mm, 512 views/360 deg, 512x384 det @0.4 mm, "
                        "SOD 500 mm, SDD 1000 mm, Shepp-Logan ramp, plain FDK (one scan per launch sequence)",
            "updates_per_step": upd,
            "storage": "f32",
            "z_unroll": info.z_unroll,
            "fft_len": info.fft_len,
        },
        "bp_gups": round(upd / (bp_mean * 1e-3) / 1e9, 3),
        "filter_ms": round(statistics.mean(filt_ms), 4),
        "bp_ms": round(bp_mean, 4),
        "latency_ms": {"p50": round(_percentile(step_ms, 0.5), 4), "p99": round(_percentile(step_ms, 0.99), 4)},
        "rmse_vs_phantom": rmse_truth,
        "roofline": {
            "bound": "valu",
            "kernel": "bp3d_kernel<FLAT, f32>",
            "achieved": round(achieved_tflops, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "alg_flops_per_update": F_ALG_FLAT,
            "alg_hbm_bytes_per_launch": alg_bytes,
            "hbm_frac_of_peak_alg": round(alg_bytes / (bp_mean * 1e-3) / 1e9 / PEAK_HBM_GBS, 5),
        },
        "input_generation_s": round(t_gen, 2),
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, proj, vol, a.threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
