"""ctypes bindings of the CPU oracle (oracle/_build/liboracle.so) and of the
reference-built checker (oracle/_ref/libgctref.so) -- test infrastructure.

Only tests, smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_PATH = ROOT / "oracle" / "_build" / "liboracle.so"
REF_PATH = ROOT / "oracle" / "_ref" / "libgctref.so"


class OrScan(C.Structure):
    _fields_ = [("kind", C.c_int), ("nu", C.c_uint), ("du", C.c_float), ("nv", C.c_uint), ("dv", C.c_float),
                ("nviews", C.c_uint), ("views", C.c_void_p)]


class OrImage(C.Structure):
    _fields_ = [("dim", C.c_uint), ("nx", C.c_uint), ("ny", C.c_uint), ("nz", C.c_uint), ("dx", C.c_float),
                ("dy", C.c_float), ("dz", C.c_float), ("mx", C.c_float), ("my", C.c_float), ("mz", C.c_float)]


class OrRot(C.Structure):
    _fields_ = [("dir", C.c_int), ("incr", C.c_float), ("range", C.c_float), ("nppr", C.c_uint)]


_o = None
_r = None


def oracle() -> C.CDLL:
    global _o
    if _o is None:
        if not ORACLE_PATH.exists():
            raise ImportError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        _o = C.CDLL(str(ORACLE_PATH))
        _o.or_wrap_0_2pi.restype = C.c_float
        _o.or_wrap_0_2pi.argtypes = [C.c_float]
        _o.or_quadrant.restype = C.c_uint
        _o.or_quadrant.argtypes = [C.c_float]
        P, U, I, F = C.c_void_p, C.c_uint, C.c_int, C.c_float
        sigs = {
            "or_rotation_geom": [P, P, P, P],
            "or_rotation_geom_f64": [P, P, P, P, P],
            "or_ramp_taps": [I, U, F, F, I, P, P],
            "or_weights_f32": [P, P, P],
            "or_weights_f64": [P, P, P],
            "or_filter_rows_f32": [U, U, P, P, P, I],
            "or_filter_rows_f64": [U, U, P, P, P, I],
            "or_filter_rows_direct_f64": [U, U, P, P, P],
            "or_bp_f32": [P, P, P, P, U, U, I, I],
            "or_bp_f64": [P, P, P, P, U, U, I, I],
            "or_fdk_f32": [P, P, F, I, P, P, U, U, I],
            "or_fdk_f64": [P, P, F, I, P, P, U, U, I],
            "or_joseph_f64": [P, P, P, P, I],
            "or_spiral_windows": [P, P, P],
            "or_fdk_spiral_f32": [P, P, P, P, U, U, I],
            "or_fdk_spiral_f64": [P, P, P, P, U, U, I],
        }
        for name, args in sigs.items():
            f = getattr(_o, name)
            f.argtypes = args
            f.restype = C.c_int
    return _o


def ref_available() -> bool:
    return REF_PATH.exists()


def ref() -> C.CDLL:
    global _r
    if _r is None:
        _r = C.CDLL(str(REF_PATH))
        _r.ref_wrap.restype = C.c_float
        _r.ref_wrap.argtypes = [C.c_float]
        _r.ref_quadrant.restype = C.c_uint
        _r.ref_quadrant.argtypes = [C.c_float]
        P = C.c_void_p
        _r.ref_ramp_taps.argtypes = [C.c_int, C.c_uint, C.c_float, C.c_float, C.c_int, P]
        _r.ref_ramp_taps.restype = C.c_int
        _r.ref_rotation_geom.argtypes = [C.c_int, C.c_uint, P, P, P, P, P, P, P, P]
        _r.ref_rotation_geom.restype = C.c_int
    return _r


def _scan(scan) -> OrScan:
    s = OrScan()
    s.kind, s.nu, s.du, s.nv, s.dv = scan.kind, scan.nu, scan.du, scan.nv, scan.dv
    s.nviews = scan.num_proj
    s.views = scan.views.ctypes.data
    return s


def _img(img) -> OrImage:
    i = OrImage()
    i.dim, i.nx, i.ny, i.nz = img.dim, img.nx, img.ny, (img.nz if img.dim == 3 else 1)
    i.dx, i.dy, i.dz = img.dx, img.dy, img.dz
    i.mx, i.my, i.mz = img.mid
    return i


_keep = []


def _sref(scan):
    s = _scan(scan)
    _keep.append(s)
    if len(_keep) > 64:
        del _keep[:32]
    return C.addressof(s)


def _iref(img):
    i = _img(img)
    _keep.append(i)
    return C.addressof(i)


def rotation(scan):
    r = OrRot()
    rot = np.zeros(scan.num_proj, np.float32)
    rel = np.zeros(scan.num_proj, np.float32)
    rc = oracle().or_rotation_geom(_sref(scan), C.addressof(r), rot.ctypes.data, rel.ctypes.data)
    return rc, r, rot, rel


def ramp_taps(arc: bool, n: int, width: float, cutoff: float = 1.0, window: int = 0):
    t32 = np.zeros(n, np.float32)
    t64 = np.zeros(n, np.float64)
    oracle().or_ramp_taps(int(arc), n, width, cutoff, window, t32.ctypes.data, t64.ctypes.data)
    return t32, t64


def weights(scan, f64=False):
    dt = np.float64 if f64 else np.float32
    cosw = np.zeros(scan.proj_shape, dt)
    red = np.zeros((scan.num_proj, scan.nu), dt)
    fn = oracle().or_weights_f64 if f64 else oracle().or_weights_f32
    fn(_sref(scan), cosw.ctypes.data, red.ctypes.data)
    return cosw, red


def filter_rows(taps, rows, f64=False, nthreads=8):
    rows = np.ascontiguousarray(rows, np.float64 if f64 else np.float32)
    taps = np.ascontiguousarray(taps, np.float64 if f64 else np.float32)
    n = rows.shape[-1]
    nr = rows.size // n
    out = np.zeros_like(rows)
    fn = oracle().or_filter_rows_f64 if f64 else oracle().or_filter_rows_f32
    fn(nr, n, taps.ctypes.data, rows.ctypes.data, out.ctypes.data, nthreads)
    return out


def filter_rows_direct(taps, rows):
    rows = np.ascontiguousarray(rows, np.float64)
    taps = np.ascontiguousarray(taps, np.float64)
    n = rows.shape[-1]
    out = np.zeros_like(rows)
    oracle().or_filter_rows_direct_f64(rows.size // n, n, taps.ctypes.data, rows.ctypes.data,
                                       out.ctypes.data)
    return out


def backproject(scan, img, filt, z0=0, z1=None, scale=True, f64=False, nthreads=8):
    dt = np.float64 if f64 else np.float32
    filt = np.ascontiguousarray(filt, dt)
    vol = np.zeros(img.shape, dt)
    z1 = (img.nz if img.dim == 3 else 1) if z1 is None else z1
    fn = oracle().or_bp_f64 if f64 else oracle().or_bp_f32
    rc = fn(_sref(scan), _iref(img), filt.ctypes.data, vol.ctypes.data, z0, z1, int(scale), nthreads)
    if rc != 0:
        raise RuntimeError("oracle backprojection failed")
    return vol


def fdk(scan, img, proj, cutoff=1.0, window=0, z0=0, z1=None, f64=False, nthreads=8):
    dt = np.float64 if f64 else np.float32
    proj = np.ascontiguousarray(proj, np.float32)
    vol = np.zeros(img.shape, dt)
    z1 = (img.nz if img.dim == 3 else 1) if z1 is None else z1
    fn = oracle().or_fdk_f64 if f64 else oracle().or_fdk_f32
    rc = fn(_sref(scan), _iref(img), cutoff, window, proj.ctypes.data, vol.ctypes.data,
            z0, z1, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle FDK failed ({rc})")
    return vol


def spiral_windows(scan, img):
    win = np.zeros((img.nz, 2), np.uint32)
    if oracle().or_spiral_windows(_sref(scan), _iref(img), win.ctypes.data) != 0:
        raise RuntimeError("invalid spiral geometry")
    return win


def fdk_spiral(scan, img, proj, z0=0, z1=None, f64=False, nthreads=8):
    dt = np.float64 if f64 else np.float32
    proj = np.ascontiguousarray(proj, np.float32)
    vol = np.zeros(img.shape, dt)
    z1 = img.nz if z1 is None else z1
    fn = oracle().or_fdk_spiral_f64 if f64 else oracle().or_fdk_spiral_f32
    rc = fn(_sref(scan), _iref(img), proj.ctypes.data, vol.ctypes.data, z0, z1, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle spiral FDK failed ({rc})")
    return vol


def joseph(scan, img, vol, nthreads=8):
    vol = np.ascontiguousarray(vol, np.float64)
    proj = np.zeros(scan.proj_shape, np.float64)
    rc = oracle().or_joseph_f64(_sref(scan), _iref(img), vol.ctypes.data, proj.ctypes.data, nthreads)
    if rc != 0:
        raise RuntimeError("oracle Joseph projection failed")
    return proj


# ---- the reference's comparator (src/cpp/numeric/comparator.h:48-59) --------

def are_equal(a, b, eps):
    """Elementwise abs-or-relative equality of the reference's areEqual."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    return (d <= eps) | (d <= eps * np.maximum(np.abs(a), np.abs(b)))


def parity_report(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    d = np.abs(got - want)
    scale = max(np.abs(want).max(), 1e-30)
    return {
        "max_abs": float(d.max()),
        "max_rel_to_peak": float(d.max() / scale),
        "rmse": float(np.sqrt(np.mean((got - want) ** 2))),
        "pass_1e-5": float(np.mean(are_equal(got, want, 1e-5))),
    }
