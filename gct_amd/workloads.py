"""Synthetic workloads: the scan geometries, image grids and seeded phantoms
the benchmark and the tests reconstruct.

Every workload is generated here from its stated parameters (no data files):
analytic line integrals of random ellipses/ellipsoids (FP64, cast to FP32),
optionally with Poisson noise.  Densities of 3-D phantoms are drawn in cm^-1
and stored in mm^-1 (x 0.1) so that photon counts at I0 = 1e5 stay meaningful.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import (CONE_PARALLEL, CYLINDRICAL, FAN_ARC_1D, FAN_LINE_1D, FLAT, PARALLEL_1D, PARALLEL_2D, Image, Scan,
               WINDOW_NONE, WINDOW_SHEPP_LOGAN, circular_scan, is_1d, phantom_project, phantom_project_device,
               phantom_random, projection_matrix, projection_matrix_view)


@dataclass
class Workload:
    name: str
    scan: Scan
    img: Image
    phantom: np.ndarray
    i0: float = 0.0
    seed: int = 0
    window: int = WINDOW_NONE
    storage_forms: tuple = ("f32",)
    notes: dict = field(default_factory=dict)

    def projections(self, nthreads: int = 0) -> np.ndarray:
        return phantom_project(self.scan, self.phantom, self.i0, self.seed + 7, nthreads)

    def projections_device(self, out=None, stream=None):
        """The same projections made by the GPU generator into a device
        tensor (allocated on the current device when `out` is None)."""
        import torch
        if out is None:
            out = torch.empty(self.scan.proj_shape, dtype=torch.float32, device="cuda")
        return phantom_project_device(self.scan, self.phantom, out, self.i0, self.seed + 7, stream)

    @property
    def updates(self) -> int:
        return self.img.nvox * self.scan.num_proj


def parallel_1d_256() -> Workload:
    """2-D parallel-beam FDK: 256^2 at 0.5 mm, 360 views over 180 deg, 367 bins
    at 0.5 mm (centre 183), Ram-Lak; 10 seeded ellipses (centres within 0.8 R,
    semi-axes U(0.05, 0.4) R, rotation U(0, pi), attenuation U(-0.3, 1.0))."""
    seed = 1001
    scan = circular_scan(PARALLEL_1D, 367, 0.5, 360, math.pi / 360.0, det_row_center=183.0)
    img = Image(256, 256, 0.5, 0.5)
    ph = phantom_random(seed, 10, 2, 64.0, (0.05, 0.4), (-0.3, 1.0))   # FOV radius 64 mm
    return Workload("parallel_1d_256", scan, img, ph, 0.0, seed, WINDOW_NONE)


def flat_cone_256(window: int = WINDOW_SHEPP_LOGAN, noise: bool = True) -> Workload:
    """Full-scan flat-detector cone beam: SOD 500, SDD 1000 mm, 512 views over
    360 deg, 512 x 384 detector at 0.4 mm (centre 255.5, 191.5), 256^3 volume
    at 0.25 mm; 12 seeded ellipsoids, Poisson noise at I0 = 1e5."""
    seed = 1002
    scan = circular_scan(FLAT, 512, 0.4, 512, 2.0 * math.pi / 512, nv=384, dv=0.4, det_row_center=255.5,
                         det_col_center=191.5, src_det_dist=1000.0, src_iso_dist=500.0)
    img = Image(256, 256, 0.25, 0.25, 256, 0.25)
    ph = phantom_random(seed, 12, 3, 32.0, (0.05, 0.35), (-0.02, 0.1))
    return Workload("flat_cone_256", scan, img, ph, 1e5 if noise else 0.0, seed, window, ("f32",))


def carm_short_scan_512() -> Workload:
    """Short-scan C-arm cone beam with Parker weighting: SOD 750, SDD 1200 mm,
    248 views over 200 deg, 616 x 480 flat detector at 0.616 mm (centre 307.5,
    239.5), 512^3 volume at 0.3 mm; FP32/FP16/BF16 filtered-projection storage."""
    seed = 1003
    scan = circular_scan(FLAT, 616, 0.616, 248, math.radians(200.0) / 248, nv=480, dv=0.616, det_row_center=307.5,
                         det_col_center=239.5, src_det_dist=1200.0, src_iso_dist=750.0)
    img = Image(512, 512, 0.3, 0.3, 512, 0.3)
    ph = phantom_random(seed, 12, 3, 76.8, (0.05, 0.35), (-0.02, 0.1))
    return Workload("carm_short_scan_512", scan, img, ph, 1e5, seed, WINDOW_NONE, ("f32", "f16", "bf16"))


def fan_arc_slice(rng: np.random.Generator, idx: int) -> Workload:
    """One slice of the uneven batch: arc detector, channels U{300..900},
    views U{360..1200} over 360 deg, image U{256..768}^2, SOD 570, SDD 1040 mm.
    The fan spans 0.957 rad (a 500 mm field of view with 5 % margin); the
    pixel is 500 mm / N."""
    nch = int(rng.integers(300, 901))
    nview = int(rng.integers(360, 1201))
    n = int(rng.integers(256, 769))
    dg = 0.957 / nch
    scan = circular_scan(FAN_ARC_1D, nch, dg, nview, 2.0 * math.pi / nview, det_row_center=(nch - 1) / 2.0,
                         src_det_dist=1040.0, src_iso_dist=570.0)
    img = Image(n, n, 500.0 / n, 500.0 / n)
    ph = phantom_random(5000 + idx, 10, 2, 250.0, (0.05, 0.35), (-0.02, 0.1))
    return Workload(f"fan_arc_{idx}", scan, img, ph, 0.0, 5000 + idx)


def cylindrical_slab(idx: int) -> Workload:
    """One MSCT slab of the uneven batch: cylindrical detector 672 channels x 64
    rows (row pitch 1.095 mm, 0.6 mm at iso), 1152 views, 512 x 512 x 64 volume
    at 0.977 x 0.977 x 0.6 mm, SOD 570, SDD 1040 mm."""
    dg = 0.957 / 672
    scan = circular_scan(CYLINDRICAL, 672, dg, 1152, 2.0 * math.pi / 1152, nv=64, dv=1.095,
                         det_row_center=335.5, det_col_center=31.5, src_det_dist=1040.0, src_iso_dist=570.0)
    img = Image(512, 512, 0.977, 0.977, 64, 0.6)
    ph = phantom_random(6000 + idx, 12, 3, 250.0, (0.05, 0.35), (-0.02, 0.1))
    return Workload(f"cyl_slab_{idx}", scan, img, ph, 0.0, 6000 + idx)


def uneven_batch(num_slices: int = 64, num_slabs: int = 16, seed: int = 1004) -> list[Workload]:
    rng = np.random.default_rng(seed)
    out = [fan_arc_slice(rng, i) for i in range(num_slices)]
    out += [cylindrical_slab(i) for i in range(num_slabs)]
    return out


def serving_request(rng: np.random.Generator, idx: int) -> Workload:
    """One request of the serving stream, drawn from the mix:
    40 % parallel / cone-parallel volumes of 128^3..256^3 (1 mm voxels,
    360..720 views over 180 deg), 40 % flat cone beam of 256^3..512^3
    (360..720 views, full 360 deg or short 180 deg + fan scans, detector
    512..1024 x 384..768 at 0.4 mm, SOD 500 mm, SDD 1000 mm, the volume sized
    to the detector's field of view), 20 % fan-arc slices as in the uneven
    batch.  Phantoms: 12 random ellipsoids (10 ellipses for slices)."""
    u = rng.random()
    seed = 7000 + idx
    if u < 0.4:
        kind = PARALLEL_2D if rng.random() < 0.5 else CONE_PARALLEL
        n = int(rng.integers(128, 257))
        views = int(rng.integers(360, 721))
        nu = int(math.ceil(n * math.sqrt(2.0))) + 4
        if kind == PARALLEL_2D:
            scan = circular_scan(kind, nu, 1.0, views, math.pi / views, nv=n + 4, dv=1.0,
                                 det_row_center=(nu - 1) / 2.0, det_col_center=(n + 3) / 2.0)
        else:
            sid, sdd = 500.0, 1000.0
            nv = int(math.ceil(n * 1.1)) + 4
            scan = circular_scan(kind, nu, 1.0, views, math.pi / views, nv=nv, dv=sdd / sid, det_row_center=(nu - 1) / 2.0,
                                 det_col_center=(nv - 1) / 2.0, src_det_dist=sdd, src_iso_dist=sid)
        img = Image(n, n, 1.0, 1.0, n, 1.0)
        ph = phantom_random(seed, 12, 3, 0.5 * n, (0.05, 0.35), (-0.02, 0.1))
        return Workload(f"serve_{idx}_{'par2d' if kind == PARALLEL_2D else 'conepar'}_{n}", scan, img, ph, 0.0, seed)
    if u < 0.8:
        n = int(rng.integers(256, 513))
        views = int(rng.integers(360, 721))
        nu = int(rng.integers(512, 1025))
        nv = int(rng.integers(384, 769))
        sid, sdd, pitch = 500.0, 1000.0, 0.4
        short = rng.random() < 0.5
        half_fan = math.atan(0.5 * nu * pitch / sdd)
        span = (math.pi + 2.0 * half_fan) if short else 2.0 * math.pi
        scan = circular_scan(FLAT, nu, pitch, views, span / views, nv=nv, dv=pitch, det_row_center=(nu - 1) / 2.0,
                             det_col_center=(nv - 1) / 2.0, src_det_dist=sdd, src_iso_dist=sid)
        vox = 0.95 * pitch * (sid / sdd) * min(nu, nv) / n
        img = Image(n, n, vox, vox, n, vox)
        ph = phantom_random(seed, 12, 3, 0.5 * n * vox, (0.05, 0.35), (-0.02, 0.1))
        return Workload(f"serve_{idx}_flat{'_short' if short else ''}_{n}", scan, img, ph, 0.0, seed)
    return fan_arc_slice(rng, idx)


def serving_mix(num_requests: int = 32, seed: int = 1005) -> list[Workload]:
    rng = np.random.default_rng(seed)
    return [serving_request(rng, i) for i in range(num_requests)]


def spiral_case(kind: int = CYLINDRICAL, n: int = 48, rotations: float = 3.0, views_per_rot: int = 64,
                nu: int = 96, nv: int = 16, pitch_mm_per_rot: float = 6.0, seed: int = 77) -> Workload:
    """Small helical scan: source z = -pitch * angle / 2 pi (hel_pitch in mm per
    radian), SOD 200 mm, SDD 400 mm, 0.5 mm voxels; the volume spans the
    middle of the helix so the central slices get a full rotation of views."""
    views = int(rotations * views_per_rot)
    incr = 2.0 * math.pi / views_per_rot
    hel = pitch_mm_per_rot / (2.0 * math.pi)
    sid, sdd = 200.0, 400.0
    du = (0.5 / sid) if kind == CYLINDRICAL else 1.0
    scan = circular_scan(kind, nu, du, views, incr, nv=nv, dv=1.0, det_row_center=(nu - 1) / 2.0,
                         det_col_center=(nv - 1) / 2.0, src_det_dist=sdd, src_iso_dist=sid, hel_pitch=hel)
    zc = -hel * incr * (views - 1) / 2.0   # source z at the middle view
    img = Image(n, n, 0.5, 0.5, 16, 0.5, mid=(0.0, 0.0, zc))
    ph = phantom_random(seed, 8, 3, 0.25 * n, (0.05, 0.35), (-0.02, 0.1))
    ph[:, 2] += zc
    return Workload(f"spiral_{kind}", scan, img, ph, 0.0, seed)


def spiral_cylindrical_slab(seed: int = 78) -> Workload:
    """Helical MSCT scan at the uneven batch's slab geometry: cylindrical
    detector 672 channels x 64 rows (0.6 mm at iso), SOD 570, SDD 1040 mm,
    1160 views per rotation, pitch 38.4 mm per rotation (pitch factor 1),
    2.5 rotations; a 512 x 512 x 64 volume at 0.977 x 0.977 x 0.6 mm in the
    middle of the helix, so every slice gets its full one-rotation window.
    (The slabs' 1152 views per rotation trip the reference's FP32 rotation
    analysis over 2.5 rotations: its accumulated range comes out above 2 pi
    and it counts every view as one rotation, which the spiral path rejects.)"""
    vpr, rotations = 1160, 2.5
    views = int(vpr * rotations)
    incr = 2.0 * math.pi / vpr
    hel = 38.4 / (2.0 * math.pi)
    scan = circular_scan(CYLINDRICAL, 672, 0.957 / 672, views, incr, nv=64, dv=1.095, det_row_center=335.5,
                         det_col_center=31.5, src_det_dist=1040.0, src_iso_dist=570.0, hel_pitch=hel)
    zc = -hel * incr * (views - 1) / 2.0   # source z at the middle view
    img = Image(512, 512, 0.977, 0.977, 64, 0.6, mid=(0.0, 0.0, zc))
    ph = phantom_random(seed, 12, 3, 250.0, (0.05, 0.35), (-0.02, 0.1))
    ph[:, 2] = ph[:, 2] * 0.1 + zc
    return Workload("spiral_cyl_slab", scan, img, ph, 0.0, seed)


def tilt_views(scan: Scan, turn_deg: float = 1.5, tilt_deg: float = 2.5) -> Scan:
    """The flat-panel views of `scan` with every detector turned in its plane
    by turn_deg and tilted about its row axis by tilt_deg (so the row axis is
    no longer along z), each view then passed through its 3 x 4 projection
    matrix and decomposed back (ProjectionMatrix compose / decompose), as a
    calibrated C-arm's geometry arrives."""
    a, b = math.radians(turn_deg), math.radians(tilt_deg)
    sc = scan
    out = np.empty_like(sc.views)
    for i, row in enumerate(sc.views):
        v = row.astype(np.float64)
        eu, ev = v[8:11], v[11:14]
        ew = np.cross(eu, ev)
        eu1 = math.cos(a) * eu + math.sin(a) * ev           # in-plane turn
        ev1 = -math.sin(a) * eu + math.cos(a) * ev
        ev2 = math.cos(b) * ev1 + math.sin(b) * ew          # tilt about the row axis
        t = v.copy()
        t[8:11], t[11:14] = eu1, ev2
        P = projection_matrix(t.astype(np.float32), sc.du, sc.dv)
        out[i] = projection_matrix_view(P, sc.du)
    return Scan(FLAT, sc.nu, sc.du, sc.nv, sc.dv, out)


def tilted_flat_case(n: int = 48, views: int = 90, nu: int = 96, nv: int = 80, seed: int = 43) -> Workload:
    """A flat-panel scan given by projection matrices, as a calibrated C-arm
    is: the circular views of small_case, turned and tilted (tilt_views)."""
    base = small_case(FLAT, n=n, views=views, nu=nu, nv=nv, seed=seed)
    return Workload("tilted_flat", tilt_views(base.scan), base.img, base.phantom, 0.0, seed)


def tilted_flat_cone_256() -> Workload:
    """The flat-cone workload's scan turned and tilted (tilt_views): the same
    volume and phantom, views given by projection matrices."""
    base = flat_cone_256(noise=False)
    return Workload("tilted_flat_cone_256", tilt_views(base.scan), base.img, base.phantom, 0.0, base.seed,
                    base.window)


# The reference's end-to-end FDK regression geometries (test/cuda/recon/
# fdk_axial_tests.cu:15-537): 360 views at 1 deg (the flat short scan at
# 0.0105434 rad), 256 x 256 (x 5) images at 1 mm, the detector pitches and
# source distances of each case.  The reference reads its projections and
# reference volumes from data files that are not in its repository; the
# detector widths here are chosen to cover the image (the reference's are
# unknown) and the projections are analytic line integrals of seeded
# ellipses / ellipsoids.  name: (kind, columns, column pitch, rows, row pitch,
# SID, SDD, angular increment)
REGRESSION_CASES = {
    "1dpdct": (PARALLEL_1D, 410, 0.8961, 1, 1.0, 0.0, 0.0, None),
    "1dadct": (FAN_ARC_1D, 420, 0.0017922, 1, 1.0, 500.0, 1000.0, None),
    "1dldct": (FAN_LINE_1D, 290, 2.0, 1, 1.0, 1000.0, 1536.0, None),
    "2dcdct": (CYLINDRICAL, 420, 0.0017922, 16, 1.14, 500.0, 1000.0, None),
    "2dfdct": (FLAT, 290, 2.0, 12, 2.0, 1000.0, 1536.0, None),
    "2dfdct_short": (FLAT, 290, 2.0, 12, 2.0, 1000.0, 1536.0, 0.0105434),
    "2dpdct": (PARALLEL_2D, 410, 0.8961, 8, 1.14, 0.0, 0.0, None),
    "2drdct": (CONE_PARALLEL, 410, 0.8961, 12, 1.14, 500.0, 1000.0, None),
}


def reference_regression(name: str, seed: int = 4242) -> Workload:
    kind, nu, du, nv, dv, sid, sdd, incr = REGRESSION_CASES[name]
    if incr is None:
        incr = float(np.float32(1.0 * math.pi / 180.0))   # static_cast<float>(1.0 * pi / 180.0)
    scan = circular_scan(kind, nu, du, 360, incr, nv=nv, dv=dv, det_row_center=(nu - 1) / 2.0,
                         det_col_center=(nv - 1) / 2.0, src_det_dist=sdd, src_iso_dist=sid)
    # densities in mm^-1, as for every 256 mm field of view here (water is ~0.02)
    if is_1d(kind):
        img = Image(256, 256, 1.0, 1.0)
        ph = phantom_random(seed, 10, 2, 128.0, (0.05, 0.4), (-0.03, 0.1))
    else:
        img = Image(256, 256, 1.0, 1.0, 5, 1.0)
        ph = phantom_random(seed, 16, 3, 128.0, (0.1, 0.45), (-0.02, 0.1))
        ph[:, 2] *= 0.05   # centres near the 5-slice slab, so that most ellipsoids cross it
    return Workload(f"regression_{name}", scan, img, ph, 0.0, seed)


def small_case(kind: int, n: int = 64, views: int = 120, nu: int = 96, nv: int = 80, short_scan: bool = False,
               direction: int = 1, seed: int = 42) -> Workload:
    """Small instance of every geometry for parity tests against the oracle."""
    span = math.radians(200.0) if short_scan else 2.0 * math.pi
    if kind in (PARALLEL_1D, PARALLEL_2D, CONE_PARALLEL) and not short_scan:
        span = math.pi
    incr = span / views
    fov = 0.5 * n * 0.5   # radius in mm with 0.5 mm pixels
    sid, sdd = 200.0, 400.0
    if kind == PARALLEL_1D:
        scan = circular_scan(kind, nu, 0.6, views, incr, det_row_center=(nu - 1) / 2.0, direction=direction)
    elif kind == FAN_LINE_1D:
        scan = circular_scan(kind, nu, 1.0, views, incr, det_row_center=(nu - 1) / 2.0 + 0.25, src_det_dist=sdd,
                             src_iso_dist=sid, direction=direction)
    elif kind == FAN_ARC_1D:
        scan = circular_scan(kind, nu, 0.5 / sid, views, incr, det_row_center=(nu - 1) / 2.0 - 0.3,
                             src_det_dist=sdd, src_iso_dist=sid, direction=direction)
    elif kind == PARALLEL_2D:
        scan = circular_scan(kind, nu, 0.6, views, incr, nv=nv, dv=0.6, det_row_center=(nu - 1) / 2.0,
                             det_col_center=(nv - 1) / 2.0 + 0.5, direction=direction)
    elif kind == CONE_PARALLEL:
        scan = circular_scan(kind, nu, 0.6, views, incr, nv=nv, dv=0.9, det_row_center=(nu - 1) / 2.0,
                             det_col_center=(nv - 1) / 2.0, src_det_dist=sdd, src_iso_dist=sid, direction=direction)
    elif kind == CYLINDRICAL:
        scan = circular_scan(kind, nu, 0.5 / sid, views, incr, nv=nv, dv=1.0, det_row_center=(nu - 1) / 2.0,
                             det_col_center=(nv - 1) / 2.0, src_det_dist=sdd, src_iso_dist=sid, direction=direction)
    else:
        scan = circular_scan(kind, nu, 1.0, views, incr, nv=nv, dv=1.0, det_row_center=(nu - 1) / 2.0 + 0.5,
                             det_col_center=(nv - 1) / 2.0 - 0.25, src_det_dist=sdd, src_iso_dist=sid,
                             direction=direction)
    if is_1d(kind):
        img = Image(n, n, 0.5, 0.5, mid=(0.3, -0.2))
        ph = phantom_random(seed, 8, 2, fov, (0.05, 0.35), (-0.02, 0.1))
    else:
        img = Image(n, n, 0.5, 0.5, n, 0.5, mid=(0.3, -0.2, 0.1))
        ph = phantom_random(seed, 8, 3, fov, (0.05, 0.35), (-0.02, 0.1))
    return Workload(f"small_{kind}", scan, img, ph, 0.0, seed)
