#!/usr/bin/env python3
"""FDK benchmark on one MI355X.  Prints one JSON line.

Sections (all on by default; `--sections flat` runs the first only):

flat      The headline number.  A step is one plain FDK reconstruction of the
          flat-cone workload (256^3 at 0.25 mm, 512 views over 360 deg,
          512 x 384 detector at 0.4 mm, SOD 500 mm, SDD 1000 mm,
          Shepp-Logan-windowed ramp, 12 seeded ellipsoids with Poisson noise
          at I0 = 1e5): fused weighting + FFT ramp filter, then
          backprojection, projections already resident in HBM.  The plan is
          autotuned once before the warm-up.  `value` is voxel x view updates
          per second over the timed steps; the backprojection kernel's
          roofline is measured with HIP events on its stream.
          Beside it: the same reconstruction with the transfers (pinned
          host projections in, volume out), and the kernel's VALU issue
          utilisation from its committed instruction counters.
perview   The same backprojection launched one view at a time with a
          full-volume read-modify-write per view (the reference's structure).
c1        The 2-D parallel-beam workload (256^2, 360 views, 367 bins): the
          GPU plan and the oracle's FP32 restatement on the host cores.
precision The short-scan C-arm workload (512^3, 248 views over 200 deg) with
          FP32, FP16 and BF16 filtered-projection storage.
batched   The uneven batch: 64 fan-arc slices of seeded sizes plus 16
          cylindrical 512 x 512 x 64 slabs (672 x 64 detector, 1152 views).
          Plain = each scan's FDK in turn, waiting for each (one scan per
          call); batched = every filter plus one persistent backprojection
          launch per kernel group, replayed from one HIP graph.
served    The serving stream: 32 seeded requests of mixed geometry and size,
          each FDK + Joseph reprojection + reprojection-error check, host
          buffers in and out.  Plain = the same pipeline one request at a
          time (one stream, one in flight); served = the continuous-admission
          server (4 streams, HBM-footprint admission), first saturated (all
          requests queued at once: capacity), then with Poisson arrivals at
          80 % of that capacity (latency).
joseph    The Joseph forward projector alone (the served requests'
          reprojection) on the largest request of the serving mix and on the
          flat-cone workload's volume and scan: rays per second, device volume
          and projections resident.
next      The rows beside the core path: the one-shot reconstruct's high,
          low-memory and z-slab modes, MetaImage I/O, spiral FDK and a
          projection-matrix (tilted detector) scan.

    python bench.py [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--sections flat,perview,c1,precision,batched,served,joseph,next]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import threading
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

F_ALG_FLAT = 45          # FP32 operations per voxel x view update of the flat-cone mapping + bilinear + accumulate
# the same count for the other geometries (the reference's per-update formula)
F_ALG = {"parallel_1d": 13, "fan_line_1d": 22, "fan_arc_1d": 26, "parallel_2d": 24, "cone_parallel": 25,
         "cylindrical": 56, "flat": 45}
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
N_CU, CLOCK_HZ = 256, 2.4e9   # MI355X: one wave64 VALU instruction per CU per cycle (4 SIMD16, non-packed)
LDS_BYTES_PER_CU_CYCLE = 128  # measured LDS read rate (ds_read2_b64: 1 KiB per wave instruction in 8 cycles)
LDS_READ_BYTES_PER_UPDATE = 16   # the LDS backprojection reads two (value, difference) pairs per update


def variant_is_lds(g, info):
    return info.bp_kernel == g.BP_LDS


def host_cores():
    """Threads the CPU legs use: the cores this process may run on (its
    affinity mask), at most 16 -- a one-GPU box's share of its host."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return min(16, avail), {"os_cpu_count": os.cpu_count(), "affinity": avail}


def valu_roof_gups(kind_name):
    return PEAK_FP32_TFLOPS * 1e12 / F_ALG[kind_name] / 1e9


def roofline_entry(kind_name, updates, bp_ms, alg_bytes):
    """Backprojection kernel against the FP32-VALU roof of its geometry
    (F_ALG operations per update) and the HBM roof of its compulsory bytes."""
    s = bp_ms * 1e-3
    gups = updates / s / 1e9
    return {"bp_ms": round(bp_ms, 4), "bp_gups": round(gups, 2), "valu_roof_gups": round(valu_roof_gups(kind_name), 1),
            "frac_of_valu_roof": round(gups / valu_roof_gups(kind_name), 4),
            "alg_tflops": round(F_ALG[kind_name] * updates / s / 1e12, 2),
            "alg_hbm_gbs": round(alg_bytes / s / 1e9, 1), "hbm_frac": round(alg_bytes / s / 1e9 / PEAK_HBM_GBS, 5)}


def timed_events(torch, stream, fn, reps):
    """Per-repetition milliseconds of fn() on `stream`, by HIP events on that stream."""
    out = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1))
    return out


def _percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    lo, hi = int(k), min(int(k) + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def cpu_baseline(wl, proj, gpu_vol, nthreads, cpu_info):
    """The CPU oracle (FP32 restatement of the reference path; the reference
    has no CPU FDK) timed on this host: weighting + ramp filter of every view
    and backprojection of a 64-slice sample of the volume, scaled to the whole
    volume.  One warm-up run, then the median of 3.  Also reports how the GPU
    volume agrees with the oracle on the sampled slices."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O

    scan, img = wl.scan, wl.img
    nsample = 64
    z0 = img.nz // 2 - nsample // 2
    t32, _ = O.ramp_taps(False, scan.nu, scan.du, 1.0, wl.window)
    runs = []
    for _ in range(4):
        t = time.perf_counter()
        cosw, red = O.weights(scan)
        weighted = (proj * cosw) * red[:, None, :]
        filt = O.filter_rows(t32, weighted, nthreads=nthreads)
        t_filter = time.perf_counter() - t
        del weighted, cosw
        t = time.perf_counter()
        vol = O.backproject(scan, img, filt, z0=z0, z1=z0 + nsample, nthreads=nthreads)
        t_bp = time.perf_counter() - t
        runs.append((t_filter + t_bp * img.nz / nsample, t_filter, t_bp))
        del filt
    runs = sorted(runs[1:])   # the first run is the warm-up
    full, t_filter, t_bp = runs[1]
    rep = O.parity_report(gpu_vol[z0:z0 + nsample], vol[z0:z0 + nsample])
    return {
        "value": round(wl.updates / full / 1e9, 4),
        "unit": "GUPS",
        "cores": nthreads,
        "host": cpu_info,
        "kind": "port",
        "sample": (f"oracle FP32 on {nthreads} threads: weighting+filter of all {scan.num_proj} views "
                   f"({t_filter:.2f} s) + backprojection of z slices [{z0},{z0 + nsample}) of {img.nz} "
                   f"({t_bp:.2f} s), scaled to the full volume; median of 3 after a warm-up"),
        "bp_gups": round(nsample * img.nx * img.ny * scan.num_proj / t_bp / 1e9, 4),
        "runs_s": [round(r[0], 3) for r in runs],
        "gpu_vs_oracle_on_sample": rep,
    }


# ---- config 1: 2-D parallel-beam FDK (the reference's CPU-runnable case) -------

def run_c1(a, g, W, torch, np, nthreads):
    """256^2 image from 360 parallel views of 367 bins: the GPU plan (HIP
    events, projections resident) and the oracle's FP32 restatement on the
    host cores (median of 3 after a warm-up)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O

    wl = W.parallel_1d_256()
    proj = wl.projections()
    p = torch.from_numpy(proj).cuda()
    v = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
    plan = g.Plan(wl.scan, wl.img)
    stream = torch.cuda.current_stream()
    for _ in range(3):
        plan.execute(p, v, stream=stream)
    ms = timed_events(torch, stream, lambda: plan.execute(p, v, stream=stream), 20)
    gpu = v.cpu().numpy()
    cpu_s = []
    for _ in range(4):
        t = time.perf_counter()
        want = O.fdk(wl.scan, wl.img, proj, nthreads=nthreads)
        cpu_s.append(time.perf_counter() - t)
    cpu = statistics.median(cpu_s[1:])
    gms = statistics.median(ms)
    plan.close()
    return {"workload": "parallel_1d_256: 256^2 @0.5 mm, 360 views/180 deg, 367 bins @0.5 mm, Ram-Lak",
            "updates": wl.updates, "gpu_ms": round(gms, 4), "gpu_gups": round(wl.updates / gms / 1e6, 2),
            "cpu_oracle_ms": round(cpu * 1e3, 2), "cpu_oracle_gups": round(wl.updates / cpu / 1e9, 4),
            "cpu_threads": nthreads, "gpu_vs_cpu": round(cpu * 1e3 / gms, 1),
            "gpu_vs_oracle": O.parity_report(gpu, want)}


# ---- flat-cone plain FDK (the headline) -----------------------------------------

def _bp_kernel_name(g, info):
    """rocprofv3's name of the backprojection kernel a flat-detector FP32 plan
    launches (template arguments: kind, storage, FP64 set-up, z unroll, and
    for the LDS kernel staged rows, views per block, pairs per row, spiral
    windows, row clamp -- off when every voxel projects inside the detector)."""
    f64 = "true" if info.setup_f64 else "false"
    if info.bp_kernel == g.BP_LDS:
        clamp = "false" if info.rows_inside else "true"
        return (f"void gct::(anonymous namespace)::bp3d_lds_kernel<6, 0, {f64}, {info.z_unroll}, {info.lds_rows}, "
                f"{info.view_block}, {info.lds_pairs}, false, {clamp}>(gct::BpLaunch)")
    return f"void gct::(anonymous namespace)::bp3d_kernel<6, 0, true, {f64}, {info.z_unroll}>(gct::BpLaunch)"


def run_flat(a, g, W, torch, np, sections):
    wl = W.flat_cone_256()
    t = time.perf_counter()
    proj = wl.projections(a.threads)
    t_gen = time.perf_counter() - t
    p = torch.from_numpy(proj).cuda()
    v = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
    plan = g.Plan(wl.scan, wl.img, window=wl.window)
    stream = torch.cuda.current_stream()
    plan.filter(p, stream=stream)
    torch.cuda.synchronize()
    tuned_zu, tuned_ms = plan.autotune(v)
    info = plan.info()

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
    # the kernel the tuned plan launches, named as rocprofv3 names it; the
    # counter-measured traffic applies only when the profile is of that kernel
    kernel = _bp_kernel_name(g, info)
    traffic, traffic_kernel, valu_insts, lds_busy, dual = None, None, None, None, None
    tfile = ROOT / "profiles" / "bp_flat_cone_256_traffic.json"
    if tfile.exists():
        try:
            tj = json.loads(tfile.read_text())
            traffic_kernel = tj.get("kernel")
            if traffic_kernel == kernel:
                traffic = tj.get("hbm_bytes_per_launch")
                valu_insts = tj.get("valu_insts_per_launch")
                lds_busy = tj.get("lds_busy_frac")
                dual = tj.get("valu_dual_issued_frac")
        except (OSError, ValueError):
            traffic = None
    # LDS read rate of the launch against 128 B per CU-cycle at the nominal clock
    lds_read_frac = round(LDS_READ_BYTES_PER_UPDATE * upd / (bp_mean * 1e-3) /
                          (LDS_BYTES_PER_CU_CYCLE * N_CU * CLOCK_HZ), 4) if variant_is_lds(g, info) else None
    # VALU issue utilisation: the profiled kernel's wave instructions over the
    # issue slots of this run's launch (one per CU per cycle at 2.4 GHz)
    valu_issue = round(valu_insts / (N_CU * CLOCK_HZ * bp_mean * 1e-3), 4) if valu_insts else None

    # End to end with the transfers: pinned host projections in, volume out
    hp = torch.empty(wl.scan.proj_shape, dtype=torch.float32, pin_memory=True)
    hp.copy_(torch.from_numpy(proj))
    hv = torch.empty(wl.img.shape, dtype=torch.float32, pin_memory=True)

    def e2e_serial():
        p.copy_(hp, non_blocking=True)
        plan.execute(p, v, stream=stream)
        hv.copy_(v, non_blocking=True)

    def e2e():
        plan.execute_host(hp, p, v, hv, stream=stream)

    e2e_serial()
    e2e_serial_ms = statistics.median(timed_events(torch, stream, e2e_serial, 5))
    e2e()
    e2e_ms = statistics.median(timed_events(torch, stream, e2e, 5))
    del hp, hv
    # the reference's call shape: one-shot host-to-host reconstruct from a
    # pageable array (plan set-up, allocations and copies included)
    host_vol = np.empty(wl.img.shape, np.float32)
    g.reconstruct(wl.scan, proj, wl.img, window=wl.window, out=host_vol)
    one_shot = []
    for _ in range(3):
        t = time.perf_counter()
        g.reconstruct(wl.scan, proj, wl.img, window=wl.window, out=host_vol)
        one_shot.append((time.perf_counter() - t) * 1e3)
    one_shot_ms = statistics.median(one_shot)
    del host_vol

    vol = v.cpu().numpy()
    truth = g.phantom_rasterize(wl.img, wl.phantom)
    rmse_truth = float(np.sqrt(np.mean((vol - truth) ** 2)))
    variant = {g.BP_LDS: "lds", g.BP_DIRECT: "direct"}.get(info.bp_kernel, str(info.bp_kernel))
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
            "bp_kernel": variant,
            "z_unroll": info.z_unroll,
            "lds_shape": {"view_block": info.view_block, "rows": info.lds_rows, "pairs": info.lds_pairs},
            "tiles_round_robin": bool(info.round_robin),
            "row_clamp": not info.rows_inside,
            "autotune_best_bp_ms": round(tuned_ms, 4),
            "fft_len": info.fft_len,
        },
        "bp_gups": round(upd / (bp_mean * 1e-3) / 1e9, 3),
        "filter_ms": round(statistics.mean(filt_ms), 4),
        "bp_ms": round(bp_mean, 4),
        "latency_ms": {"p50": round(_percentile(step_ms, 0.5), 4), "p99": round(_percentile(step_ms, 0.99), 4)},
        "rmse_vs_phantom": rmse_truth,
        "roofline": {
            "bound": "valu",
            "kernel": kernel,
            "achieved": round(achieved_tflops, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
            "frac_basis": "algorithmic FLOPs of the reference's per-update formula (45 per update) over the "
                          "FP32 peak; the kernel evaluates them strength-reduced along z in fewer instructions, "
                          "so its own limits are valu_issue_frac and lds_read_frac",
            "valu_issue_frac": valu_issue,
            # wave instructions x 64 lanes over the updates (one lane per update)
            "valu_insts_per_update": round(valu_insts * 64 / upd, 2) if valu_insts else None,
            "valu_dual_issued_frac_profiled": dual,
            "lds_read_frac": lds_read_frac,
            "lds_busy_frac_profiled": lds_busy,
            "binding": "LDS bandwidth: 16 B of LDS reads per update at 128 B per CU-cycle (lds_read_frac; the "
                       "profiled LDS busy fraction includes staging writes and bank conflicts); VALU issue has "
                       "headroom because full-rate FP32 instructions dual-issue",
            "traffic": traffic,
            "traffic_profiled_kernel": traffic_kernel,
            "alg_flops_per_update": F_ALG_FLAT,
            "alg_hbm_bytes_per_launch": alg_bytes,
            "hbm_frac_of_peak_alg": round(alg_bytes / (bp_mean * 1e-3) / 1e9 / PEAK_HBM_GBS, 5),
        },
        "end_to_end_with_pcie": {
            "what": "pinned host projections in, host volume out (gct_plan_execute_host): the upload of each "
                    "eighth of the views overlaps the filter and backprojection of the previous one",
            "ms": round(e2e_ms, 3), "gups": round(upd / (e2e_ms * 1e-3) / 1e9, 1),
            "bytes_moved": int(proj.nbytes + wl.img.nvox * 4),
            "serial_ms": round(e2e_serial_ms, 3),
            "serial_what": "the same, upload, FDK and download one after another on one stream",
            "one_shot_reconstruct_ms": round(one_shot_ms, 2),
            "one_shot_what": "gct_reconstruct from a pageable numpy array (the reference's host-to-host call "
                             "shape: plan set-up, device allocations and copies included), median of 3",
        },
        "targets": {
            "served_vs_plain_5x": "measured in 'served'; the plain path already launches one fused kernel per "
                                  "scan at the VALU issue limit, so batching and serving can only recover "
                                  "launch, set-up, transfer and tail time, not multiply kernel throughput",
            "hbm_roofline_60pct": "unattainable for backprojection by construction: its compulsory bytes are "
                                  "0.055 B/update, so at the VALU roof it moves ~2 % of HBM peak "
                                  "(hbm_frac_of_peak_alg); the kernel is bounded by VALU issue instead",
        },
        "input_generation_s": round(t_gen, 2),
    }
    if "perview" in sections:
        # The reference's structure on the same kernels: one view per launch,
        # each launch reading and writing the whole volume (fdk_axial.cuh:189-199).
        V = wl.scan.num_proj
        per_view = []
        for rep in range(2):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for i in range(V):
                plan.backproject(v, first=i, n=1, accumulate=i > 0, finish=i == V - 1, stream=stream)
            e1.record(stream)
            torch.cuda.synchronize()
            per_view.append(e0.elapsed_time(e1))
        per_view_ms = min(per_view)
        out["reference_structured_bp"] = {
            "what": "same kernel launched one view at a time, full-volume read-modify-write per view",
            "bp_ms": round(per_view_ms, 3),
            "bp_gups": round(upd / (per_view_ms * 1e-3) / 1e9, 2),
            "fused_vs_per_view": round(per_view_ms / bp_mean, 3),
        }
    del p, v
    plan.close()
    return out, wl, proj, vol


# ---- precision forms on the C-arm short scan ----------------------------------------

def run_precision(a, g, W, torch, np):
    """FP32 / FP16 / BF16 filtered-projection storage (FP32 accumulation, FP64
    per-view tables and column set-up at 512^3) on the short-scan C-arm
    workload: throughput and error against the rasterised phantom and the
    FP32 form.  (Errors against the FP64 oracle: tests/test_gpu_precision.py.)"""
    wl = W.carm_short_scan_512()
    stream = torch.cuda.current_stream()
    p = wl.projections_device()
    truth = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
    g.phantom_rasterize_device(wl.img, wl.phantom, truth)
    forms = {}
    ref = None
    for name, st in (("f32", g.STORE_F32), ("f16", g.STORE_F16), ("bf16", g.STORE_BF16)):
        plan = g.Plan(wl.scan, wl.img, storage=st)
        v = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
        plan.execute(p, v, stream=stream)
        torch.cuda.synchronize()
        plan.autotune(v)          # each storage form picks its own backprojection variant
        ms = statistics.median(timed_events(torch, stream, lambda: plan.execute(p, v, stream=stream), 5))
        info = plan.info()
        bp_ms = statistics.median(timed_events(torch, stream, lambda: plan.backproject(v, stream=stream), 5))
        plan.execute(p, v, stream=stream)
        sbytes = 4 if st == g.STORE_F32 else 2
        alg_bytes = wl.scan.num_proj * wl.scan.nv * wl.scan.nu * sbytes + wl.img.nvox * 4
        d = v - truth
        f = {"ms": round(ms, 3), "gups": round(wl.updates / (ms * 1e-3) / 1e9, 2),
             "bp_variant": {"kernel": "lds" if info.bp_kernel == g.BP_LDS else "direct",
                            "z_unroll": int(info.z_unroll), "view_block": int(info.view_block),
                            "rows": int(info.lds_rows), "pairs": int(info.lds_pairs)},
             "filtered_bytes": int(info.filt_bytes), "setup_f64": bool(info.setup_f64),
             "roofline": roofline_entry("flat", wl.updates, bp_ms, alg_bytes),
             "rmse_vs_phantom": float(torch.sqrt(torch.mean(d * d)))}
        if ref is None:
            ref = v.clone()
        else:
            e = v - ref
            f["rmse_vs_f32"] = float(torch.sqrt(torch.mean(e * e)))
            f["max_abs_vs_f32"] = float(e.abs().max())
        forms[name] = f
        plan.close()
        del v
    return {"workload": "carm_short_scan_512: SOD 750, SDD 1200 mm, 248 views over 200 deg (Parker), 616x480 det "
                        "@0.616 mm, 512^3 @0.3 mm, Poisson I0=1e5", "updates": wl.updates, "forms": forms}


# ---- uneven batch: plain per-scan calls vs one graph ------------------------------

def run_batched(a, g, W, torch, np):
    wls = W.uneven_batch()
    stream = torch.cuda.current_stream()
    t = time.perf_counter()
    projs = [w.projections_device() for w in wls]
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t
    vols = [torch.empty(w.img.shape, dtype=torch.float32, device="cuda") for w in wls]
    plans = [g.Plan(w.scan, w.img) for w in wls]
    upd = sum(w.updates for w in wls)

    def plain():
        for pl, p, v in zip(plans, projs, vols):
            pl.execute(p, v, stream=stream)
            stream.synchronize()

    def plain_async():
        for pl, p, v in zip(plans, projs, vols):
            pl.execute(p, v, stream=stream)
        stream.synchronize()

    plain()
    want = [v.clone() for v in vols]
    batch = g.Batch(plans)
    batch.execute(projs, vols, use_graph=True, stream=stream)
    stream.synchronize()
    same = all(bool(torch.equal(v, w)) for v, w in zip(vols, want))
    del want

    reps = max(1, min(a.steps, 3))

    def timed(fn):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    t_plain = timed(plain)
    t_async = timed(plain_async)
    t_batch = timed(lambda: batch.execute(projs, vols, use_graph=True, stream=stream))
    n_slices = sum(1 for w in wls if g.is_1d(w.scan.kind))
    # Backprojection kernels alone (each plan's own launch, after its filter),
    # per geometry: the cylindrical slabs (all one shape: the first one) and
    # the sum over the fan-arc slices.
    roof = {}
    slab_i = next(i for i, w in enumerate(wls) if not g.is_1d(w.scan.kind))
    pl, p, v, w = plans[slab_i], projs[slab_i], vols[slab_i], wls[slab_i]
    pl.execute(p, v, stream=stream)
    slab_ms = statistics.median(timed_events(torch, stream, lambda: pl.backproject(v, stream=stream), 3))
    roof["cylindrical_slab"] = roofline_entry("cylindrical", w.updates, slab_ms,
                                              w.scan.num_proj * w.scan.nv * w.scan.nu * 4 + w.img.nvox * 4)
    arc_ms, arc_upd, arc_bytes = 0.0, 0, 0
    for pl, p, v, w in zip(plans, projs, vols, wls):
        if not g.is_1d(w.scan.kind):
            continue
        pl.execute(p, v, stream=stream)
        arc_ms += min(timed_events(torch, stream, lambda: pl.backproject(v, stream=stream), 2))
        arc_upd += w.updates
        arc_bytes += w.scan.num_proj * w.scan.nu * 4 + w.img.nvox * 4
    roof["fan_arc_slices_sum"] = roofline_entry("fan_arc_1d", arc_upd, arc_ms, arc_bytes)
    out = {
        "workload": f"{n_slices} fan-arc slices (channels U{{300..900}}, views U{{360..1200}}, images U{{256..768}}^2) "
                    f"+ {len(wls) - n_slices} cylindrical slabs (672x64 det, 1152 views, 512x512x64)",
        "updates": upd,
        "plain_s": round(t_plain, 4),
        "plain_gups": round(upd / t_plain / 1e9, 2),
        "plain_no_wait_s": round(t_async, 4),
        "batched_graph_s": round(t_batch, 4),
        "batched_gups": round(upd / t_batch / 1e9, 2),
        "batched_vs_plain": round(t_plain / t_batch, 3),
        "batched_identical_to_plain": same,
        "roofline": roof,
        "input_generation_s": round(t_gen, 2),
    }
    batch.close()
    for pl in plans:
        pl.close()
    return out


# ---- serving stream ------------------------------------------------------------

def run_served(a, g, W, torch, np):
    reqs = W.serving_mix()
    t = time.perf_counter()
    host_proj, host_vol = [], []
    for w in reqs:
        d = w.projections_device()
        hp = torch.empty(w.scan.proj_shape, dtype=torch.float32, pin_memory=True)
        hp.copy_(d)
        host_proj.append(hp)
        host_vol.append(torch.empty(w.img.shape, dtype=torch.float32, pin_memory=True))
        del d
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t
    pn = [p.numpy() for p in host_proj]
    vn = [v.numpy() for v in host_vol]
    upd = sum(w.updates for w in reqs)

    def closed_loop(server):
        t0 = time.perf_counter()
        tickets = [server.submit(w.scan, w.img, p, v, reproject=True) for w, p, v in zip(reqs, pn, vn)]
        res = [server.wait(tk) for tk in tickets]
        return time.perf_counter() - t0, res

    plain = g.Server(num_streams=1, max_inflight=1)
    closed_loop(plain)                      # warm-up (pool, kernels)
    t_plain, res_plain = closed_loop(plain)
    plain.close()
    plain_vols = [v.copy() for v in vn[:4]]

    served = g.Server()
    closed_loop(served)
    t_served, res_served = closed_loop(served)
    same = all(np.array_equal(v, w) for v, w in zip(vn[:4], plain_vols))

    # Poisson arrivals at 80 % of the measured served capacity
    qps = 0.8 * len(reqs) / t_served
    rng = np.random.default_rng(1005)
    gaps = rng.exponential(1.0 / qps, len(reqs))
    tickets = [None] * len(reqs)

    def feeder():
        t0 = time.perf_counter()
        due = 0.0
        for i, (w, p, v) in enumerate(zip(reqs, pn, vn)):
            due += gaps[i]
            dt = t0 + due - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets[i] = served.submit(w.scan, w.img, p, v, reproject=True)

    t0 = time.perf_counter()
    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    res_poisson = [served.wait(tk) for tk in tickets]
    t_poisson = time.perf_counter() - t0
    stats = served.stats()
    served.close()

    # The floor of any schedule: the requests' GPU work alone (weighting,
    # filter, backprojection, Joseph, residual-free), device-resident, one
    # after another.  plain / floor bounds what overlapping transfers and set-up
    # with other requests' kernels can gain.
    stream = torch.cuda.current_stream()
    gpu_ms = 0.0
    for w, hp in zip(reqs, host_proj):
        plan = g.Plan(w.scan, w.img)
        p = hp.cuda()
        v = torch.empty(w.img.shape, dtype=torch.float32, device="cuda")
        re = torch.empty_like(p)

        def work():
            plan.execute(p, v, stream=stream)
            plan.forward_project(v, re, stream=stream)

        work()
        gpu_ms += min(timed_events(torch, stream, work, 2))
        plan.close()
        del p, v, re

    # A repeated protocol: one mid-size flat request submitted 8 times in turn,
    # with the plan cache and without it (each submit waits for the previous).
    rep = next(i for i, w in enumerate(reqs) if w.scan.kind == g.FLAT and 300 <= w.img.nx <= 400)
    wr, pr, vr = reqs[rep], pn[rep], vn[rep]
    repeat = {}
    for label, entries in (("plan_cache", 0), ("no_plan_cache", g.NO_PLAN_CACHE)):
        sv = g.Server(plan_cache_entries=entries)
        lat = [sv.wait(sv.submit(wr.scan, wr.img, pr, vr, reproject=True))[0] for _ in range(8)]
        sv.close()
        repeat[label] = {"first_ms": round(lat[0], 2), "p50_rest_ms": round(statistics.median(lat[1:]), 2)}

    lat_plain = [r[0] for r in res_plain]
    lat_sat = [r[0] for r in res_served]
    lat_p = [r[0] for r in res_poisson]
    errs = [r[1] for r in res_poisson]
    kinds = {}
    for w in reqs:
        kinds[g.KIND_NAMES[w.scan.kind]] = kinds.get(g.KIND_NAMES[w.scan.kind], 0) + 1
    return {
        "workload": f"{len(reqs)} requests (seeded mix {kinds}), each FDK + Joseph reprojection + residual, "
                    "pinned host buffers in/out",
        "updates": upd,
        "plain_s": round(t_plain, 4),
        "plain_gups": round(upd / t_plain / 1e9, 2),
        "plain_service_ms": {"p50": round(_percentile(lat_plain, 0.5), 2),
                             "p99": round(_percentile(lat_plain, 0.99), 2)},
        "served_capacity_s": round(t_served, 4),
        "served_gups": round(upd / t_served / 1e9, 2),
        "served_vs_plain": round(t_plain / t_served, 3),
        "served_saturated_latency_ms": {"p50": round(_percentile(lat_sat, 0.5), 2),
                                        "p99": round(_percentile(lat_sat, 0.99), 2)},
        "poisson_qps": round(qps, 3),
        "poisson_latency_ms": {"p50": round(_percentile(lat_p, 0.5), 2), "p99": round(_percentile(lat_p, 0.99), 2)},
        "poisson_gups": round(upd / t_poisson / 1e9, 2),
        "reprojection_rel_err": {"median": round(statistics.median(errs), 5), "max": round(max(errs), 5)},
        "identical_alone_and_served": same,
        "gpu_compute_floor_s": round(gpu_ms * 1e-3, 4),
        "attainable_served_vs_plain": round(t_plain / (gpu_ms * 1e-3), 3),
        "bound": "served runs at the GPU-compute floor of the mix (its requests' kernels one after another); "
                 "what serving overlaps -- PCIe copies, host set-up -- is what plain adds to it, so "
                 "attainable_served_vs_plain bounds served_vs_plain",
        "plan_cache": {"hits": stats["plan_cache_hits"], "misses": stats["plan_cache_misses"],
                       "note": "the mix repeats no geometry within a pass; 'repeat_protocol' shows a repeated one"},
        "repeat_protocol": {"request": wr.name, **repeat},
        "admission": "first fit in arrival order; a request passed over 4 times blocks later ones",
        "peak_device_bytes": stats["peak_device_bytes"],
        "peak_inflight": stats["peak_inflight"],
        "input_generation_s": round(t_gen, 2),
    }


def run_joseph(a, g, W, torch, np):
    """Forward projection time of the largest served request (its volume does
    not fit the caches: the projector re-reads it from beyond L2 view after
    view) and of the flat-cone workload (it does)."""
    reqs = [w for w in W.serving_mix() if not g.is_1d(w.scan.kind)]
    big = max(reqs, key=lambda w: w.img.nvox * w.scan.num_proj)
    res = {}
    for wl in (big, W.flat_cone_256(noise=False)):
        plan = g.Plan(wl.scan, wl.img)
        vol = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
        g.phantom_rasterize_device(wl.img, wl.phantom, vol)
        out = torch.empty(wl.scan.proj_shape, dtype=torch.float32, device="cuda")
        stream = torch.cuda.current_stream()
        plan.forward_project(vol, out, stream=stream)
        torch.cuda.synchronize()
        ts = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            plan.forward_project(vol, out, stream=stream)
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        ms = statistics.median(ts)
        rays = wl.scan.num_proj * wl.scan.nv * wl.scan.nu
        res[wl.name] = {"volume": list(wl.img.shape), "views": wl.scan.num_proj,
                        "detector": [wl.scan.nv, wl.scan.nu], "rays": rays, "ms": round(ms, 3),
                        "grays_per_s": round(rays / ms / 1e6, 3)}
        plan.close()
        del vol, out
    return res


def run_next(a, g, W, torch, np):
    """The rows beside the core path, each timed on a full-size workload:
    f1 streaming and z slabs (the one-shot host-to-host reconstruct of the
    flat-cone workload in its high-memory, low-memory and slab modes, pinned
    host buffers, volumes compared bit for bit), f2 file I/O (the same volume
    written and read back as MetaImage), f3 spiral FDK (a helical
    cylindrical slab, one-shot from a host array) and f4 projection-matrix
    geometry (the flat-cone scan with turned and tilted detectors: per-voxel
    set-up on the direct kernel, device-resident, HIP events)."""
    import tempfile
    res = {}

    def wall(fn, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t) * 1e3)
        return statistics.median(ts)

    # f1: high / low / slab memory paths, host to host
    wl = W.flat_cone_256(noise=False)
    ph = torch.empty(wl.scan.proj_shape, dtype=torch.float32).pin_memory()
    ph.copy_(wl.projections_device().cpu())
    phn = ph.numpy()
    outs = {}
    modes = {"high": dict(mem_mode=g.MEM_HIGH), "low_chunks_64": dict(mem_mode=g.MEM_LOW, view_chunk=64),
             "slab_64_resident": dict(mem_mode=g.MEM_SLAB, z_slab=64),
             "slab_64_streamed": dict(mem_mode=g.MEM_SLAB, z_slab=64, view_chunk=64)}
    f1 = {"workload": wl.name, "what": "gct_reconstruct host to host from pinned buffers, plan set-up and "
                                       "device allocations included; median of 3 after one call"}
    for name, kw in modes.items():
        vh = torch.empty(wl.img.shape, dtype=torch.float32).pin_memory().numpy()
        ms = wall(lambda: g.reconstruct(wl.scan, phn, wl.img, window=wl.window, out=vh, **kw))
        outs[name] = vh
        f1[name] = {"ms": round(ms, 2), "gups": round(wl.updates / ms / 1e6, 1)}
    f1["identical"] = all(np.array_equal(outs[k], outs["high"]) for k in outs)
    res["f1_streaming"] = f1

    # f2: MetaImage write + read of that volume
    vol = outs["high"]
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "vol.mha")
        t = time.perf_counter()
        g.write_meta(fn, vol)
        t_w = time.perf_counter() - t
        t = time.perf_counter()
        back = g.read_array(fn)
        t_r = time.perf_counter() - t
    res["f2_io"] = {"what": f"MetaImage .mha of the {list(vol.shape)} float32 volume", "bytes": int(vol.nbytes),
                    "write_gbs": round(vol.nbytes / t_w / 1e9, 2), "read_gbs": round(vol.nbytes / t_r / 1e9, 2),
                    "round_trip_identical": bool(np.array_equal(back.reshape(vol.shape), vol))}
    del ph, phn, outs, vol, back

    # f3: spiral FDK, one shot
    sw = W.spiral_cylindrical_slab()
    sp = sw.projections_device().cpu().numpy()
    win = g.spiral_windows(sw.scan, sw.img)
    upd = int(sw.img.nx) * sw.img.ny * int(np.sum(win[:, 1] - win[:, 0]))
    sv = np.empty(sw.img.shape, np.float32)
    ms = wall(lambda: sv.__setitem__(Ellipsis, g.reconstruct_spiral(sw.scan, sp, sw.img)))
    res["f3_spiral"] = {"workload": f"{sw.name}: {sw.scan.num_proj} views of {sw.scan.nv} x {sw.scan.nu}, "
                                    f"volume {list(sw.img.shape)}, one rotation of views per slice",
                        "updates": upd, "ms": round(ms, 2), "gups": round(upd / ms / 1e6, 1),
                        "what": "gct_reconstruct_spiral from a pageable host array (upload of "
                                f"{sp.nbytes / 1e6:.0f} MB, plan set-up and allocations included)"}
    del sp, sv

    # f4: projection-matrix (non-circular) geometry, device resident
    tw = W.tilted_flat_cone_256()
    plan = g.Plan(tw.scan, tw.img, window=tw.window)
    info = plan.info()
    p = tw.projections_device()
    v = torch.empty(tw.img.shape, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    plan.filter(p, stream=stream)
    plan.backproject(v, stream=stream)
    torch.cuda.synchronize()
    bp_ms = statistics.median(timed_events(torch, stream, lambda: plan.backproject(v, stream=stream), 5))
    res["f4_projection_matrices"] = {
        "workload": "flat_cone_256 with every detector turned 1.5 deg in plane and tilted 2.5 deg about its "
                    "row axis, views from 3x4 projection matrices",
        "circular": int(info.circular), "setup_f64": int(info.setup_f64),
        "roofline": roofline_entry("flat", tw.updates, bp_ms,
                                   tw.scan.num_proj * tw.scan.nv * tw.scan.nu * 4 + tw.img.nvox * 4)}
    plan.close()
    del p, v
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sections", default="flat,perview,c1,precision,batched,served,joseph,next")
    ap.add_argument("--threads", type=int, default=0, help="host threads of the CPU legs (0: host_cores())")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy 5 times (reference kernel for HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (the FDK path here "
              "runs on one device: independent scans and requests would shard one per GPU, a single "
              "large volume by z slab).", file=sys.stderr)
        sys.exit(2)
    sections = {s.strip() for s in a.sections.split(",") if s.strip()}
    cores, cpu_info = host_cores()
    if not a.threads:
        a.threads = cores

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

    out, wl, proj, vol = run_flat(a, g, W, torch, np, sections)
    if "c1" in sections:
        print("bench: c1 section", file=sys.stderr, flush=True)
        out["c1_parallel_1d"] = run_c1(a, g, W, torch, np, a.threads)
    if "precision" in sections:
        print("bench: precision section", file=sys.stderr, flush=True)
        out["precision"] = run_precision(a, g, W, torch, np)
        torch.cuda.empty_cache()
    if "batched" in sections:
        print("bench: batched section", file=sys.stderr, flush=True)
        out["batched"] = run_batched(a, g, W, torch, np)
        torch.cuda.empty_cache()
    if "served" in sections:
        print("bench: served section", file=sys.stderr, flush=True)
        out["served"] = run_served(a, g, W, torch, np)
        torch.cuda.empty_cache()
    if "joseph" in sections:
        print("bench: joseph section", file=sys.stderr, flush=True)
        out["joseph"] = run_joseph(a, g, W, torch, np)
        torch.cuda.empty_cache()
    if "next" in sections:
        print("bench: next-rows section", file=sys.stderr, flush=True)
        out["next_rows"] = run_next(a, g, W, torch, np)
        torch.cuda.empty_cache()
    if not a.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(wl, proj, vol, a.threads, cpu_info)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
