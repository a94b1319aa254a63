"""GPU tests of the features around the FDK path: precision forms, batched
persistent launches in HIP graphs, the autotuner, serving, the Joseph
projector.  All through the C ABI."""
import math

import numpy as np
import pytest

import gct_amd as g
import oracle_lib as O
from gct_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(torch, wl, proj, **kw):
    plan = g.Plan(wl.scan, wl.img, **kw)
    p = torch.from_numpy(proj).cuda()
    v = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
    plan.execute(p, v)
    torch.cuda.synchronize()
    return v.cpu().numpy(), plan


# ---- precision forms ------------------------------------------------------------------

@pytest.mark.parametrize("storage,bound", [(g.STORE_F16, 1e-3), (g.STORE_BF16, 1e-3)], ids=["f16", "bf16"])
@pytest.mark.parametrize("kind", [g.FAN_LINE_1D, g.CYLINDRICAL, g.FLAT], ids=lambda k: g.KIND_NAMES[k])
def test_half_storage_rmse_against_fp64_oracle(torch_dev, kind, storage, bound):
    """Filtered projections stored in FP16/BF16, FP32 accumulation: RMSE against
    the FP64 oracle within 1e-3 (values are O(0.1) mm^-1)."""
    wl = W.small_case(kind)
    proj = wl.projections()
    got, plan = _run(torch_dev, wl, proj, storage=storage)
    want = O.fdk(wl.scan, wl.img, proj, f64=True)
    rep = O.parity_report(got, want)
    print(f"\n[{g.KIND_NAMES[kind]} storage {storage}] {rep}")
    assert rep["rmse"] < bound
    assert plan.info().storage == storage


def test_fp64_setup_matches_fp64_oracle(torch_dev):
    """FP64 per-thread coordinate setup: at least as close to the FP64 oracle as
    the FP32 setup."""
    wl = W.small_case(g.FLAT)
    proj = wl.projections()
    f32, p32 = _run(torch_dev, wl, proj, setup_precision=g.SETUP_F32)
    f64, p64 = _run(torch_dev, wl, proj, setup_precision=g.SETUP_F64)
    assert p64.info().setup_f64 == 1 and p32.info().setup_f64 == 0
    want = O.fdk(wl.scan, wl.img, proj, f64=True)
    r32, r64 = O.parity_report(f32, want), O.parity_report(f64, want)
    print(f"\n[setup f32] {r32}\n[setup f64] {r64}")
    assert r64["rmse"] <= 1.2 * r32["rmse"] + 1e-9
    assert O.are_equal(f64, O.fdk(wl.scan, wl.img, proj), 1e-5).all()


# ---- batching ---------------------------------------------------------------------------

def test_batch_graph_equals_individual_plans(torch_dev):
    """Uneven scans (fan arc slices of several sizes, cylindrical and flat slabs)
    in one persistent launch inside a HIP graph give each scan's own result bit
    for bit, on every graph replay."""
    torch = torch_dev
    rng = np.random.default_rng(11)
    wls = [W.fan_arc_slice(rng, i) for i in range(3)]
    wls = [W.Workload(w.name, g.circular_scan(g.FAN_ARC_1D, w.scan.nu // 4, w.scan.du * 4, w.scan.num_proj // 4,
                                              np.float32(2 * math.pi / (w.scan.num_proj // 4)),
                                              src_det_dist=1040.0, src_iso_dist=570.0),
                      g.Image(w.img.nx // 4, w.img.ny // 4, w.img.dx * 4, w.img.dy * 4), w.phantom) for w in wls]
    wls += [W.small_case(g.CYLINDRICAL, n=40, views=90), W.small_case(g.FLAT, n=48, views=100),
            W.small_case(g.PARALLEL_2D, n=24, views=60)]
    projs = [torch.from_numpy(w.projections()).cuda() for w in wls]
    plans = [g.Plan(w.scan, w.img, bp_kernel=g.BP_DIRECT, z_unroll=8) for w in wls]
    want = []
    for pl, p, w in zip(plans, projs, wls):
        v = torch.empty(w.img.shape, dtype=torch.float32, device="cuda")
        pl.execute(p, v)
        want.append(v)
    torch.cuda.synchronize()
    want = [v.cpu().numpy() for v in want]
    batch = g.Batch(plans)
    vols = [torch.full(w.img.shape, 7.0, dtype=torch.float32, device="cuda") for w in wls]
    for rep in range(3):
        for v in vols:
            v.fill_(float(rep))
        batch.execute(projs, vols, use_graph=True)
        torch.cuda.synchronize()
        for v, ref in zip(vols, want):
            assert np.array_equal(v.cpu().numpy(), ref)
    batch.execute(projs, vols, use_graph=False)
    torch.cuda.synchronize()
    for v, ref in zip(vols, want):
        assert np.array_equal(v.cpu().numpy(), ref)


def test_batch_graph_and_plain_paths_keep_their_own_tables(torch_dev):
    """Graph replay into volumes A, a plain execute into volumes B, then a
    replay into A again: the replay writes A (its table still names A's
    buffers), and B keeps the plain execute's result."""
    torch = torch_dev
    wls = [W.small_case(g.FLAT, n=32, views=60), W.small_case(g.FAN_ARC_1D, n=40, views=72),
           W.small_case(g.CYLINDRICAL, n=32, views=48)]
    projs = [torch.from_numpy(w.projections()).cuda() for w in wls]
    plans = [g.Plan(w.scan, w.img) for w in wls]
    want = []
    for pl, p, w in zip(plans, projs, wls):
        v = torch.empty(w.img.shape, dtype=torch.float32, device="cuda")
        pl.execute(p, v)
        want.append(v)
    batch = g.Batch(plans)
    A = [torch.zeros(w.img.shape, dtype=torch.float32, device="cuda") for w in wls]
    B = [torch.zeros(w.img.shape, dtype=torch.float32, device="cuda") for w in wls]
    batch.execute(projs, A, use_graph=True)
    batch.execute(projs, B, use_graph=False)
    torch.cuda.synchronize()
    for a, b, ref in zip(A, B, want):
        assert torch.equal(a, ref) and torch.equal(b, ref)
        a.fill_(-1.0)
        b.fill_(-7.0)
    batch.execute(projs, A, use_graph=True)
    torch.cuda.synchronize()
    for a, b, ref in zip(A, B, want):
        assert torch.equal(a, ref)
        assert bool((b == -7.0).all()), "the replay wrote into the plain execute's volumes"
    batch.close()


def test_batch_with_fp64_setup_scan(torch_dev):
    """A 512^3 cone scan (FP64 per-column set-up) batched with small FP32
    scans: every volume equals its own plan's bit for bit (the 512^3 plan's
    parity with the oracle is tested on the same workload in
    test_gpu_precision.py)."""
    torch = torch_dev
    carm = W.carm_short_scan_512()
    wls = [carm, W.small_case(g.FLAT, n=48, views=90), W.small_case(g.FAN_ARC_1D, n=64, views=90),
           W.small_case(g.CYLINDRICAL, n=40, views=60)]
    projs = [carm.projections_device()] + [torch.from_numpy(w.projections()).cuda() for w in wls[1:]]
    plans = [g.Plan(w.scan, w.img) for w in wls]
    assert plans[0].info().setup_f64 == 1 and plans[1].info().setup_f64 == 0
    want = []
    for pl, p, w in zip(plans, projs, wls):
        v = torch.empty(w.img.shape, dtype=torch.float32, device="cuda")
        pl.execute(p, v)
        want.append(v)
    batch = g.Batch(plans)
    for use_graph in (True, False):
        vols = [torch.full(w.img.shape, 9.0, dtype=torch.float32, device="cuda") for w in wls]
        batch.execute(projs, vols, use_graph=use_graph)
        torch.cuda.synchronize()
        for w, v, ref in zip(wls, vols, want):
            assert torch.equal(v, ref), f"{w.name} (graph {use_graph}): {(v != ref).sum().item()} voxels differ"
        del vols
    batch.close()


@pytest.mark.parametrize("storage,size", [(g.STORE_F32, 4), (g.STORE_F16, 2)], ids=["f32", "f16"])
def test_filtered_buffer_spare_elements(torch_dev, storage, size):
    """The filtered buffer holds V views of (Nv + 3) padded rows and then
    2 * pitch + 64 spare zero elements, which the LDS staging loads of the
    last view may reach (tests/test_staging_bounds.py replays them)."""
    wl = W.small_case(g.FLAT, n=32, views=40, nu=45, nv=6)
    plan = g.Plan(wl.scan, wl.img, storage=storage)
    i = plan.info()
    assert i.filt_pitch == 48 and i.filt_rows_per_view == 9
    assert i.filt_bytes == (40 * 9 * 48 + 2 * 48 + 64) * size


def test_plan_before_torch_touches_the_gpu():
    """A fresh process that loads the library and creates a plan before it
    imports torch: the Python layer imports torch before loading the library,
    so both run on torch's HIP runtime (one runtime in the process) and torch
    still finds the GPU afterwards."""
    import subprocess
    import sys
    code = ("import gct_amd as g; from gct_amd import workloads as W\n"
            "wl = W.small_case(g.FLAT, n=16, views=20, nu=24, nv=20)\n"
            "plan = g.Plan(wl.scan, wl.img)\n"
            "import torch\n"
            "v = torch.zeros(wl.img.shape, device='cuda'); p = torch.zeros(wl.scan.proj_shape, device='cuda')\n"
            "plan.execute(p, v); torch.cuda.synchronize(); print('ok', float(v.abs().sum()))\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240,
                       cwd=str(__import__("pathlib").Path(__file__).resolve().parent.parent))
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and r.stdout.startswith("ok 0.0")


def test_autotuner_picks_a_valid_variant_and_keeps_results(torch_dev):
    torch = torch_dev
    wl = W.small_case(g.FLAT, n=64, views=120)
    proj = torch.from_numpy(wl.projections()).cuda()
    plan = g.Plan(wl.scan, wl.img)
    v = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
    plan.execute(proj, v)
    torch.cuda.synchronize()
    ref = v.cpu().numpy()
    zu, ms = plan.autotune(v)
    assert zu in (4, 8, 16, 32) and ms > 0
    plan.execute(proj, v)
    torch.cuda.synchronize()
    assert np.array_equal(v.cpu().numpy(), ref)


# ---- serving ------------------------------------------------------------------------------

def test_server_results_identical_alone_and_in_the_mix(torch_dev):
    """Each request's volume is the same bit for bit when reconstructed alone
    and when it overlaps other requests on the device."""
    mix = [W.small_case(k, n=48, views=90) for k in (g.FLAT, g.FAN_ARC_1D, g.CYLINDRICAL, g.PARALLEL_1D,
                                                     g.CONE_PARALLEL, g.FLAT)]
    projs = [w.projections() for w in mix]
    alone = []
    for w, p in zip(mix, projs):
        sv = g.Server(num_streams=1, max_inflight=1)
        out = np.zeros(w.img.shape, np.float32)
        sv.wait(sv.submit(w.scan, w.img, p, out))
        sv.close()
        alone.append(out)
    sv = g.Server(num_streams=4, max_inflight=4)
    outs = [np.zeros(w.img.shape, np.float32) for w in mix]
    tickets = [sv.submit(w.scan, w.img, p, o, reproject=True) for w, p, o in zip(mix, projs, outs)]
    errs = [sv.wait(t)[1] for t in tickets]
    st = sv.stats()
    sv.close()
    for o, a in zip(outs, alone):
        assert np.array_equal(o, a)
    assert all(0.0 <= e < 0.5 for e in errs), errs
    assert st["completed"] == len(mix) and st["failed"] == 0
    assert st["peak_inflight"] >= 2


def test_server_plan_cache(torch_dev):
    """A repeated request reuses its cached plan (same bits); a budget that
    cannot hold a cached plan beside the next request releases the plan, and
    the server never uses more than its budget."""
    a = W.small_case(g.FLAT, n=40, views=60)
    b = W.small_case(g.CONE_PARALLEL, n=48, views=72)
    pa, pb = a.projections(), b.projections()
    sv = g.Server(num_streams=2)
    out1, out2 = np.zeros(a.img.shape, np.float32), np.zeros(a.img.shape, np.float32)
    sv.wait(sv.submit(a.scan, a.img, pa, out1))
    sv.wait(sv.submit(a.scan, a.img, pa, out2))
    st = sv.stats()
    sv.close()
    assert (st["plan_cache_misses"], st["plan_cache_hits"]) == (1, 1)
    assert np.array_equal(out1, out2)
    assert O.are_equal(out1, O.fdk(a.scan, a.img, pa), 1e-5).all()
    # room for one request and no more: each plan is released for the next request
    fa, fb = g.required_memory(a.scan, a.img), g.required_memory(b.scan, b.img)
    budget = max(fa, fb) + 1024
    sv = g.Server(num_streams=1, hbm_budget_bytes=budget)
    outs = []
    for wl, p in ((a, pa), (b, pb), (a, pa)):
        o = np.zeros(wl.img.shape, np.float32)
        sv.wait(sv.submit(wl.scan, wl.img, p, o))
        outs.append(o)
    st = sv.stats()
    sv.close()
    assert st["peak_device_bytes"] <= budget
    assert st["plan_cache_hits"] == 0 and st["plan_cache_misses"] == 3
    assert np.array_equal(outs[0], out1) and np.array_equal(outs[2], out1)
    assert O.are_equal(outs[1], O.fdk(b.scan, b.img, pb), 1e-5).all()


def test_server_admission_ages_bypassed_requests(torch_dev):
    """First fit with aging: a large request that does not fit while another
    runs is passed over at most max_bypass (here 1) times; later small
    requests then wait for it instead of overtaking it indefinitely.  Without
    the limit every small request finishes before it."""
    x = W.flat_cone_256(noise=False)                     # ~0.9 GB, ~20 ms with its transfers
    big = W.serving_mix()[1]                             # the 475^3 short scan, ~2.8 GB
    small = [W.small_case(g.PARALLEL_1D, n=32, views=40, seed=s) for s in range(5)]
    px, ps = x.projections_device().cpu().numpy(), [w.projections() for w in small]
    pbig = big.projections_device().cpu().numpy()
    fx, fbig = g.required_memory(x.scan, x.img), g.required_memory(big.scan, big.img)
    fs = max(g.required_memory(w.scan, w.img) for w in small)
    assert fbig > fx + fs
    budget = fbig + fs - 1          # the large request only runs alone; x runs beside a small one

    def order(max_bypass):
        sv = g.Server(num_streams=4, hbm_budget_bytes=budget, plan_cache_entries=g.NO_PLAN_CACHE,
                      max_bypass=max_bypass)
        outs = [np.zeros(x.img.shape, np.float32), np.zeros(big.img.shape, np.float32)]
        outs += [np.zeros(w.img.shape, np.float32) for w in small]
        import time
        t0 = time.perf_counter()
        tickets = [sv.submit(x.scan, x.img, px, outs[0]), sv.submit(big.scan, big.img, pbig, outs[1])]
        tickets += [sv.submit(w.scan, w.img, p, o) for w, p, o in zip(small, ps, outs[2:])]
        t_sub = time.perf_counter() - t0
        done = [sv.wait(t)[0] for t in tickets]   # latency from each submit
        sv.close()
        return done, t_sub * 1e3

    done, t_sub = order(1)
    print(f"\n[aging max_bypass=1] latencies {[round(d, 1) for d in done]} ms, submitted in {t_sub:.1f} ms")
    assert done[0] > t_sub, "the long request finished before the others were queued"
    # after one small request passed it, the large one runs before the rest
    assert sum(1 for d in done[2:] if d < done[1]) <= 1
    done, _ = order(1000)
    print(f"[aging max_bypass=1000] latencies {[round(d, 1) for d in done]} ms")
    assert all(d < done[1] for d in done[2:])


def test_server_rejects_requests_over_budget_and_survives_failures(torch_dev):
    w = W.small_case(g.FLAT, n=32, views=60)
    p = w.projections()
    sv = g.Server(hbm_budget_bytes=64 << 20, num_streams=2)
    big = g.Image(512, 512, 0.1, 0.1, 512, 0.1)
    with pytest.raises(g.GctError) as e:
        sv.submit(w.scan, big, p, np.zeros(big.shape, np.float32))
    assert e.value.status == 6
    # a stationary trajectory passes submission but fails in the worker; others complete
    bad_views = w.scan.views.copy()
    bad_views[:, 5:8] = bad_views[0, 5:8]
    bad = g.Scan(w.scan.kind, w.scan.nu, w.scan.du, w.scan.nv, w.scan.dv, bad_views)
    out_ok, out_bad = np.zeros(w.img.shape, np.float32), np.zeros(w.img.shape, np.float32)
    t_bad = sv.submit(bad, w.img, p, out_bad)
    t_ok = sv.submit(w.scan, w.img, p, out_ok)
    with pytest.raises(g.GctError) as e:
        sv.wait(t_bad)
    assert e.value.status == 2
    sv.wait(t_ok)
    ref = O.fdk(w.scan, w.img, p)
    assert O.are_equal(out_ok, ref, 1e-5).all()
    st = sv.stats()
    assert st["rejected"] == 1 and st["failed"] == 1 and st["completed"] == 1
    sv.close()


# ---- Joseph projector ----------------------------------------------------------------------

@pytest.mark.parametrize("kind", range(7), ids=g.KIND_NAMES)
def test_joseph_matches_oracle(torch_dev, kind):
    torch = torch_dev
    wl = W.small_case(kind, n=40, views=30)
    vol = g.phantom_rasterize(wl.img, wl.phantom)
    plan = g.Plan(wl.scan, wl.img)
    out = torch.empty(wl.scan.proj_shape, dtype=torch.float32, device="cuda")
    plan.forward_project(torch.from_numpy(vol).cuda(), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = O.joseph(wl.scan, wl.img, vol)
    rep = O.parity_report(got, want)
    print(f"\n[joseph {g.KIND_NAMES[kind]}] {rep}")
    assert rep["max_rel_to_peak"] < 1e-4


@pytest.mark.parametrize("kind", range(7), ids=g.KIND_NAMES)
def test_joseph_approximates_analytic_line_integrals(torch_dev, kind):
    """Joseph of the rasterised phantom against its exact line integrals."""
    torch = torch_dev
    wl = W.small_case(kind, n=64, views=24)
    vol = g.phantom_rasterize(wl.img, wl.phantom)
    plan = g.Plan(wl.scan, wl.img)
    out = torch.empty(wl.scan.proj_shape, dtype=torch.float32, device="cuda")
    plan.forward_project(torch.from_numpy(vol).cuda(), out)
    torch.cuda.synchronize()
    exact = wl.projections()
    err = np.sqrt(np.mean((out.cpu().numpy() - exact) ** 2)) / np.sqrt(np.mean(exact ** 2))
    print(f"\n[joseph vs exact {g.KIND_NAMES[kind]}] relative rms {err:.4f}")
    assert err < 0.10   # rasterisation error of the phantom; a convention error gives ~1


def _joseph_cases():
    cases = [(k, W.small_case(k, n=40, views=30)) for k in (g.PARALLEL_2D, g.CONE_PARALLEL, g.CYLINDRICAL, g.FLAT)]
    # tiny voxels under wide detector pixels: 32 columns span > 64 voxels, so every
    # segment's brick overflows and its taps come from global memory
    w = W.small_case(g.FLAT, n=48, views=12)
    cases.append(("flat_wide_bricks", W.Workload("wide", w.scan, g.Image(48, 48, 0.02, 0.02, 48, 0.02), w.phantom)))
    # flat voxels (dz << dx): steep rays march along z and take the per-ray path
    cases.append(("flat_z_march", W.Workload("zm", w.scan, g.Image(40, 40, 0.5, 0.5, 40, 0.05), w.phantom)))
    return cases


def test_joseph_lds_bricks_equal_per_ray_kernel(torch_dev, tmp_path):
    """The LDS-staged projector (workgroups of 4 views x 8 rows x 32 columns,
    bricks of 8 planes) gives every ray the same bits as the one-ray-per-thread
    kernel, which a second process selects with GCT_JOSEPH=per_ray."""
    import subprocess
    import sys
    torch = torch_dev
    cases = _joseph_cases()
    got = {}
    rng = np.random.default_rng(5)
    for name, wl in cases:
        vol = rng.random(wl.img.shape, dtype=np.float32)   # every tap matters
        np.save(tmp_path / f"vol_{name}.npy", vol)
        plan = g.Plan(wl.scan, wl.img)
        out = torch.empty(wl.scan.proj_shape, dtype=torch.float32, device="cuda")
        plan.forward_project(torch.from_numpy(vol).cuda(), out)
        torch.cuda.synchronize()
        got[name] = out.cpu().numpy()
        assert np.abs(got[name]).max() > 0, name
    code = ("import sys, numpy as np, torch; sys.path.insert(0, 'tests')\n"
            "import gct_amd as g; from test_gpu_features import _joseph_cases\n"
            f"d = {str(tmp_path)!r}\n"
            "for name, wl in _joseph_cases():\n"
            "    vol = np.load(f'{d}/vol_{name}.npy')\n"
            "    plan = g.Plan(wl.scan, wl.img)\n"
            "    out = torch.empty(wl.scan.proj_shape, dtype=torch.float32, device='cuda')\n"
            "    plan.forward_project(torch.from_numpy(vol).cuda(), out); torch.cuda.synchronize()\n"
            "    np.save(f'{d}/per_ray_{name}.npy', out.cpu().numpy())\n")
    env = dict(__import__("os").environ, GCT_JOSEPH="per_ray")
    root = __import__("pathlib").Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240, cwd=str(root),
                       env=env)
    assert r.returncode == 0, r.stderr[-3000:]
    for name, _ in cases:
        ref = np.load(tmp_path / f"per_ray_{name}.npy")
        diff = int((got[name] != ref).sum())
        print(f"[joseph lds vs per-ray {name}] {diff} of {ref.size} rays differ", flush=True)
        assert diff == 0, name


@pytest.mark.parametrize("kind", [g.PARALLEL_1D, g.PARALLEL_2D], ids=lambda k: g.KIND_NAMES[k])
def test_joseph_adjoint_of_backprojection_for_parallel_beams(torch_dev, kind):
    """<A x, y> du ~ <x, B y> dx dy with B the unweighted backprojector: the
    discrete pair are transposes up to interpolation error on smooth data."""
    torch = torch_dev
    wl = W.small_case(kind, n=48, views=90)
    x = g.phantom_rasterize(wl.img, wl.phantom).astype(np.float64)
    # smooth y: projections of another phantom
    y = g.phantom_project(wl.scan, g.phantom_random(99, 6, 2 if g.is_1d(kind) else 3, 10.0)).astype(np.float64)
    plan = g.Plan(wl.scan, wl.img)
    ax = torch.empty(wl.scan.proj_shape, dtype=torch.float32, device="cuda")
    plan.forward_project(torch.from_numpy(x.astype(np.float32)).cuda(), ax)
    plan.load_filtered(torch.from_numpy(y.astype(np.float32)).cuda())
    by = torch.empty(wl.img.shape, dtype=torch.float32, device="cuda")
    plan.backproject(by, finish=False)
    torch.cuda.synchronize()
    dtheta = float(plan.info().rot_ang_incr)
    lhs = float((ax.cpu().numpy().astype(np.float64) * y).sum()) * wl.scan.du * dtheta
    vox = wl.img.dx * wl.img.dy
    rhs = float((x * by.cpu().numpy().astype(np.float64)).sum()) * vox * dtheta
    if not g.is_1d(kind):
        lhs *= wl.scan.dv
        rhs *= wl.img.dz
    print(f"\n[adjoint {g.KIND_NAMES[kind]}] {lhs} vs {rhs}")
    assert lhs == pytest.approx(rhs, rel=0.02)
