"""Python host layer over libgct_amd.so, the MI355X FDK reconstruction library.

The library's C ABI (include/gct_amd.h) is the product boundary; this module
binds it with ctypes so tests, the benchmark and Python callers drive the same
entry points a C or C++ caller would.  Device buffers are passed as raw
pointers (torch tensors' data_ptr() on ROCm); host buffers as numpy arrays.

Importing fails loudly when the library has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
# GCT_AMD_LIB selects another build of the library (default: the in-tree one).
LIB_PATH = Path(os.environ["GCT_AMD_LIB"]) if os.environ.get("GCT_AMD_LIB") else _PKG / "libgct_amd.so"

# gct_geom_kind
PARALLEL_1D, FAN_ARC_1D, FAN_LINE_1D, PARALLEL_2D, CONE_PARALLEL, CYLINDRICAL, FLAT = range(7)
KIND_NAMES = ["parallel_1d", "fan_arc_1d", "fan_line_1d", "parallel_2d", "cone_parallel", "cylindrical", "flat"]
# gct_window
WINDOW_NONE, WINDOW_HANN, WINDOW_HAMMING, WINDOW_SHEPP_LOGAN = range(4)
# gct_storage
STORE_F32, STORE_F16, STORE_BF16 = range(3)
# gct_setup_precision
SETUP_AUTO, SETUP_F32, SETUP_F64 = range(3)
# gct_mem_mode
MEM_AUTO, MEM_HIGH, MEM_LOW, MEM_SLAB = range(4)
# gct_bp_kernel
BP_AUTO, BP_DIRECT, BP_LDS = range(3)

STATUS = {0: "ok", 1: "invalid argument", 2: "runtime error", 3: "out of memory", 4: "no device",
          5: "unsupported", 6: "rejected", 7: "not ready"}


def is_1d(kind: int) -> bool:
    return kind <= FAN_LINE_1D


class Sampler(C.Structure):
    _fields_ = [("num_samples", C.c_uint32), ("sampling_width", C.c_float)]


class ProjGeom(C.Structure):
    _fields_ = [("rot_ang", C.c_float), ("src_det_dist", C.c_float), ("src_iso_dist", C.c_float),
                ("det_row_center", C.c_float), ("det_col_center", C.c_float), ("src_pos", C.c_float * 3),
                ("det_row_dir", C.c_float * 3), ("det_col_dir", C.c_float * 3), ("iso_ray_dir", C.c_float * 3)]


PROJ_GEOM_FLOATS = C.sizeof(ProjGeom) // 4   # 17


class CTGeom(C.Structure):
    _fields_ = [("kind", C.c_int32), ("det_row", Sampler), ("det_col", Sampler), ("num_proj", C.c_uint32),
                ("proj_geom", C.POINTER(ProjGeom))]


class ImageGeom(C.Structure):
    _fields_ = [("dim", C.c_uint32), ("x", Sampler), ("y", Sampler), ("z", Sampler), ("mid_point", C.c_float * 3)]


class RotationGeom(C.Structure):
    _fields_ = [("dir", C.c_int32), ("rot_ang_incr", C.c_float), ("rot_ang_range", C.c_float),
                ("num_proj_per_rot", C.c_uint32)]


class ReconOpts(C.Structure):
    _fields_ = [("cut_off_freq", C.c_float), ("window", C.c_int32), ("storage", C.c_int32),
                ("setup_precision", C.c_int32), ("mem_mode", C.c_int32), ("device", C.c_int32),
                ("z_unroll", C.c_uint32), ("view_chunk", C.c_uint32), ("bp_kernel", C.c_int32),
                ("view_block", C.c_uint32), ("lds_rows", C.c_uint32), ("lds_pairs", C.c_uint32),
                ("z_slab", C.c_uint32)]


class CircularParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("dir", C.c_int32), ("init_ang", C.c_float), ("ang_incr", C.c_float),
                ("num_proj", C.c_uint32), ("hel_pitch", C.c_float), ("det_row_center", C.c_float),
                ("det_col_center", C.c_float), ("src_det_dist", C.c_float), ("src_iso_dist", C.c_float)]


class PlanInfo(C.Structure):
    _fields_ = [("fft_len", C.c_uint32), ("filt_pitch", C.c_uint32), ("filt_rows_per_view", C.c_uint32),
                ("z_unroll", C.c_uint32), ("filt_bytes", C.c_uint64), ("device_bytes", C.c_uint64),
                ("updates", C.c_uint64), ("rot_ang_incr", C.c_float), ("rot_ang_range", C.c_float),
                ("rot_dir", C.c_int32), ("setup_f64", C.c_int32), ("circular", C.c_int32), ("storage", C.c_int32),
                ("bp_kernel", C.c_int32), ("view_block", C.c_uint32), ("lds_rows", C.c_uint32),
                ("lds_pairs", C.c_uint32), ("rows_inside", C.c_int32), ("round_robin", C.c_int32),
                ("narrow_fan", C.c_int32), ("box_inner", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("hbm_budget_bytes", C.c_uint64), ("max_inflight", C.c_uint32),
                ("num_streams", C.c_uint32), ("plan_cache_entries", C.c_uint32), ("max_bypass", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("rejected", C.c_uint64), ("updates", C.c_uint64), ("peak_device_bytes", C.c_uint64),
                ("peak_inflight", C.c_uint32), ("busy_seconds", C.c_double), ("plan_cache_hits", C.c_uint64),
                ("plan_cache_misses", C.c_uint64)]


NO_PLAN_CACHE = 0xFFFFFFFF   # gct_server_opts.plan_cache_entries: keep no plans


class ArrayMetadata(C.Structure):
    """ArrayMetadata (metadata.h:23-29): rows x cols per slice, slices per file."""
    _fields_ = [("num_rows", C.c_uint32), ("num_cols", C.c_uint32), ("num_slices", C.c_uint32), ("type", C.c_int32)]

    def __repr__(self):
        return (f"ArrayMetadata(rows={self.num_rows}, cols={self.num_cols}, slices={self.num_slices}, "
                f"type={ELEM_NAMES[self.type]})")


class GctError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built.

    The library binds the HIP runtime already loaded in the process.  PyTorch
    for ROCm ships its own HIP and HSA runtimes; loaded after this library,
    they would sit beside the system ones it pulled in, and two runtimes in
    one process cannot share the GPU (the second finds no device).  So when
    torch is installed it is imported first, and the library then runs on
    torch's runtime: one runtime, whose device pointers and streams both
    sides use."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m gct_amd.build`")
        if "torch" not in sys.modules:
            import importlib.util
            if importlib.util.find_spec("torch") is not None:
                import torch  # noqa: F401  (its HIP runtime must be the one the library binds)
        L = C.CDLL(str(LIB_PATH))
        _declare(L)
        _lib = L
    return _lib


def _declare(L: C.CDLL) -> None:
    P = C.c_void_p
    st = C.c_int
    sig = {
        "gct_abi_version": (C.c_uint32, []),
        "gct_last_error": (C.c_char_p, []),
        "gct_status_string": (C.c_char_p, [C.c_int]),
        "gct_recon_opts_default": (None, [C.POINTER(ReconOpts)]),
        "gct_device_count": (C.c_int, []),
        "gct_geom_create_circular": (st, [C.POINTER(CircularParams), P]),
        "gct_rotation_geom_compute": (st, [C.POINTER(CTGeom), C.POINTER(RotationGeom), P, P]),
        "gct_check_axial_recon": (st, [C.POINTER(CTGeom), C.c_uint32, C.c_uint32, C.c_uint32]),
        "gct_ramp_kernel": (st, [C.POINTER(CTGeom), C.c_float, C.c_int32, P]),
        "gct_projection_matrix": (st, [C.POINTER(ProjGeom), C.c_float, C.c_float, P]),
        "gct_projection_matrix_decompose": (st, [P, C.c_float, C.POINTER(ProjGeom), P]),
        "gct_reconstruct": (st, [C.POINTER(CTGeom), P, C.POINTER(ImageGeom), P, C.POINTER(ReconOpts)]),
        "gct_required_memory": (st, [C.POINTER(CTGeom), C.POINTER(ImageGeom), C.POINTER(ReconOpts),
                                     C.POINTER(C.c_uint64)]),
        "gct_plan_create": (st, [C.POINTER(CTGeom), C.POINTER(ImageGeom), C.POINTER(ReconOpts), C.POINTER(P)]),
        "gct_plan_destroy": (None, [P]),
        "gct_plan_get_info": (st, [P, C.POINTER(PlanInfo)]),
        "gct_plan_filter": (st, [P, P, C.c_uint32, C.c_uint32, P]),
        "gct_plan_backproject": (st, [P, P, C.c_uint32, C.c_uint32, C.c_int, C.c_int, P]),
        "gct_plan_execute": (st, [P, P, P, P]),
        "gct_plan_execute_host": (st, [P, P, P, P, P, C.c_uint32, P]),
        "gct_plan_load_filtered": (st, [P, P, P]),
        "gct_plan_store_filtered": (st, [P, P, P]),
        "gct_plan_forward_project": (st, [P, P, P, P]),
        "gct_plan_autotune": (st, [P, P, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
        "gct_batch_create": (st, [C.POINTER(P), C.c_uint32, C.POINTER(P)]),
        "gct_batch_destroy": (None, [P]),
        "gct_batch_execute": (st, [P, C.POINTER(P), C.POINTER(P), C.c_int, P]),
        "gct_batch_execute_noalloc": (st, [C.POINTER(P), C.c_uint32, C.POINTER(P), C.POINTER(P), P]),
        "gct_server_create": (st, [C.POINTER(ServerOpts), C.POINTER(P)]),
        "gct_server_destroy": (None, [P]),
        "gct_server_submit": (st, [P, C.POINTER(CTGeom), C.POINTER(ImageGeom), C.POINTER(ReconOpts), P, P, C.c_int,
                                   C.POINTER(C.c_uint64)]),
        "gct_server_wait": (st, [P, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "gct_server_poll": (st, [P, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "gct_server_get_stats": (st, [P, C.POINTER(ServerStats)]),
        "gct_phantom_random": (st, [C.c_uint64, C.c_uint32, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, P]),
        "gct_phantom_project": (st, [C.POINTER(CTGeom), C.c_int, P, C.c_uint32, C.c_double, C.c_uint64, P, C.c_int]),
        "gct_phantom_rasterize": (st, [C.POINTER(ImageGeom), P, C.c_uint32, P]),
        "gct_spiral_windows": (st, [C.POINTER(CTGeom), C.POINTER(ImageGeom), P]),
        "gct_reconstruct_spiral": (st, [C.POINTER(CTGeom), P, C.POINTER(ImageGeom), P, C.POINTER(ReconOpts)]),
        "gct_file_read_metadata": (st, [C.c_char_p, C.c_int32, C.POINTER(ArrayMetadata), C.POINTER(ArrayMetadata)]),
        "gct_file_read": (st, [C.c_char_p, C.c_int32, C.c_int, C.POINTER(ArrayMetadata), C.c_int32, P, C.c_uint64,
                               C.POINTER(C.c_uint64)]),
        "gct_meta_type_name": (C.c_char_p, [C.c_int32]),
        "gct_meta_write": (st, [C.c_char_p, C.c_uint32, P, P, P, C.c_int32, P, C.c_uint64]),
        "gct_den_write": (st, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, P, C.c_uint64]),
        "gct_raw_write": (st, [C.c_char_p, C.c_int32, P, C.c_uint64]),
        "gct_phantom_project_device": (st, [C.POINTER(CTGeom), C.c_int, P, C.c_uint32, C.c_double, C.c_uint64, P, P]),
        "gct_phantom_rasterize_device": (st, [C.POINTER(ImageGeom), P, C.c_uint32, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args


def check(status: int) -> None:
    if status != 0:
        raise GctError(status, lib().gct_last_error().decode())


def _ptr(a) -> int:
    """Address of a device tensor already checked by _dev."""
    return a.data_ptr()


def _dev(t, shape, what: str) -> int:
    """Address of a device buffer of float32 values of `shape`: a contiguous
    float32 torch tensor on a GPU with exactly that many elements.  Anything
    else -- a numpy array, a host tensor, an integer, a wrong size -- raises
    before any library call: a kernel handed a host address or a short buffer
    faults the GPU."""
    torch = sys.modules.get("torch")
    if torch is None or not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch tensor on the GPU, not {type(t).__name__}")
    if t.device.type != "cuda":
        raise ValueError(f"{what} must be on the GPU, not on {t.device}")
    if t.dtype != torch.float32:
        raise ValueError(f"{what} must be float32, not {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    n = int(np.prod(shape))
    if t.numel() != n:
        raise ValueError(f"{what} holds {t.numel()} values; {tuple(shape)} needs {n}")
    return t.data_ptr()


def _host(t, shape, what: str) -> int:
    """Address of a host buffer of float32 values of `shape`: a C-contiguous
    float32 numpy array or a contiguous float32 torch tensor in host memory
    (pinned or not) with exactly that many elements."""
    torch = sys.modules.get("torch")
    n = int(np.prod(shape))
    if isinstance(t, np.ndarray):
        if t.dtype != np.float32 or not t.flags["C_CONTIGUOUS"]:
            raise ValueError(f"{what} must be a C-contiguous float32 array")
        if t.size != n:
            raise ValueError(f"{what} holds {t.size} values; {tuple(shape)} needs {n}")
        return t.ctypes.data
    if torch is not None and isinstance(t, torch.Tensor):
        if t.device.type != "cpu":
            raise ValueError(f"{what} must be in host memory, not on {t.device}")
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError(f"{what} must be a contiguous float32 tensor")
        if t.numel() != n:
            raise ValueError(f"{what} holds {t.numel()} values; {tuple(shape)} needs {n}")
        return t.data_ptr()
    raise TypeError(f"{what} must be a numpy array or a host tensor, not {type(t).__name__}")


# ---- geometry --------------------------------------------------------------

class Scan:
    """An AxialCTGeom: detector samplers plus one ProjGeom per view.

    views is a float32 array [V][17] with gct_proj_geom's layout."""

    def __init__(self, kind: int, nu: int, du: float, nv: int = 1, dv: float = 1.0, views=None):
        self.kind = int(kind)
        self.nu, self.du = int(nu), float(du)
        self.nv, self.dv = (1, 1.0) if is_1d(kind) else (int(nv), float(dv))
        self.views = np.ascontiguousarray(views, dtype=np.float32).reshape(-1, PROJ_GEOM_FLOATS)
        self._c = None

    @property
    def num_proj(self) -> int:
        return self.views.shape[0]

    def c(self) -> CTGeom:
        g = CTGeom()
        g.kind = self.kind
        g.det_row = Sampler(self.nu, self.du)
        g.det_col = Sampler(self.nv, self.dv)
        g.num_proj = self.num_proj
        g.proj_geom = C.cast(self.views.ctypes.data, C.POINTER(ProjGeom))
        self._c = g
        return g

    @property
    def proj_shape(self):
        return (self.num_proj, self.nv, self.nu)


def circular_scan(kind: int, nu: int, du: float, num_proj: int, ang_incr: float, init_ang: float = 0.0,
                  nv: int = 1, dv: float = 1.0, det_row_center: float | None = None,
                  det_col_center: float | None = None, src_det_dist: float = 0.0, src_iso_dist: float = 0.0,
                  direction: int = 1, hel_pitch: float = 0.0) -> Scan:
    """ProjGeomFactory<...>::create for every view of an ideal circular (or,
    with hel_pitch != 0, helical: source z = -hel_pitch * angle) scan."""
    p = CircularParams()
    p.kind, p.dir = kind, direction
    p.init_ang, p.ang_incr, p.num_proj, p.hel_pitch = init_ang, ang_incr, num_proj, hel_pitch
    p.det_row_center = (nu - 1) / 2.0 if det_row_center is None else det_row_center
    p.det_col_center = (nv - 1) / 2.0 if det_col_center is None else det_col_center
    p.src_det_dist, p.src_iso_dist = src_det_dist, src_iso_dist
    views = np.zeros((num_proj, PROJ_GEOM_FLOATS), np.float32)
    check(lib().gct_geom_create_circular(C.byref(p), views.ctypes.data))
    return Scan(kind, nu, du, nv, dv, views)


class Image:
    def __init__(self, nx: int, ny: int, dx: float, dy: float, nz: int = 0, dz: float = 1.0, mid=(0.0, 0.0, 0.0)):
        self.dim = 3 if nz else 2
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz) if nz else 1
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.mid = tuple(float(m) for m in mid) + (0.0,) * (3 - len(mid))

    def c(self) -> ImageGeom:
        g = ImageGeom()
        g.dim = self.dim
        g.x = Sampler(self.nx, self.dx)
        g.y = Sampler(self.ny, self.dy)
        g.z = Sampler(self.nz if self.dim == 3 else 1, self.dz)
        g.mid_point[:] = self.mid
        return g

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx) if self.dim == 3 else (self.ny, self.nx)

    @property
    def nvox(self) -> int:
        return self.nx * self.ny * (self.nz if self.dim == 3 else 1)


def opts(**kw) -> ReconOpts:
    """gct_recon_opts with the library's defaults and the given fields set;
    an unknown field name is an error (ctypes would otherwise accept it as a
    plain attribute that never reaches the library)."""
    o = ReconOpts()
    lib().gct_recon_opts_default(C.byref(o))
    names = {f[0] for f in ReconOpts._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"gct_recon_opts has no field {k!r}")
        setattr(o, k, v)
    return o


def rotation_geom(scan: Scan):
    r = RotationGeom()
    rot = np.zeros(scan.num_proj, np.float32)
    rel = np.zeros(scan.num_proj, np.float32)
    check(lib().gct_rotation_geom_compute(C.byref(scan.c()), C.byref(r), rot.ctypes.data, rel.ctypes.data))
    return r, rot, rel


def spiral_windows(scan: Scan, img: Image) -> np.ndarray:
    """[first, last) views of every slice of a spiral reconstruction ([nz][2])."""
    win = np.zeros((img.nz, 2), np.uint32)
    check(lib().gct_spiral_windows(C.byref(scan.c()), C.byref(img.c()), win.ctypes.data))
    return win


def reconstruct_spiral(scan: Scan, proj: np.ndarray, img: Image, **kw) -> np.ndarray:
    """Spiral FDK, host to host (Recon::CUDA::reconstruct for SpiralCTGeom)."""
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    if proj.size != scan.num_proj * scan.nv * scan.nu:
        raise ValueError("projection stack does not match the scan")
    vol = np.zeros(img.shape, np.float32)
    o = opts(**kw)
    check(lib().gct_reconstruct_spiral(C.byref(scan.c()), proj.ctypes.data, C.byref(img.c()), vol.ctypes.data,
                                       C.byref(o)))
    return vol


def ramp_kernel(scan: Scan, cutoff: float = 1.0, window: int = WINDOW_NONE) -> np.ndarray:
    taps = np.zeros(scan.nu, np.float32)
    check(lib().gct_ramp_kernel(C.byref(scan.c()), cutoff, window, taps.ctypes.data))
    return taps


def projection_matrix(view, pix_width: float, pix_height: float) -> np.ndarray:
    """The 3 x 4 projection matrix of a flat-panel view (a row of Scan.views)
    normalised by P[2, 3] (ProjectionMatrix(pg, du, dv))."""
    pg = ProjGeom.from_buffer_copy(np.ascontiguousarray(view, np.float32).tobytes())
    P = np.zeros((3, 4), np.float64)
    check(lib().gct_projection_matrix(C.byref(pg), pix_width, pix_height, P.ctypes.data))
    return P


def projection_matrix_view(P: np.ndarray, pix_width: float) -> np.ndarray:
    """The flat-panel view of a 3 x 4 projection matrix, as a row of
    Scan.views, with the source-detector distance in mm (the focal length in
    pixels times pix_width)."""
    P = np.ascontiguousarray(P, np.float64)
    pg = ProjGeom()
    check(lib().gct_projection_matrix_decompose(P.ctypes.data, pix_width, C.byref(pg), None))
    return np.frombuffer(bytes(pg), np.float32).copy()


def required_memory(scan: Scan, img: Image, **kw) -> int:
    """Device bytes of a reconstruction with every view resident
    (gct_required_memory): projections, volume and plan -- also a served
    request's footprint when it asks for no reprojection."""
    n = C.c_uint64()
    check(lib().gct_required_memory(C.byref(scan.c()), C.byref(img.c()), C.byref(opts(**kw)), C.byref(n)))
    return n.value


def reconstruct(scan: Scan, proj: np.ndarray, img: Image, out: np.ndarray | None = None, **kw) -> np.ndarray:
    """One-shot host-to-host reconstruction (Recon::CUDA::reconstruct); the
    volume goes to `out` when given (a float32 array of the image's shape)."""
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    if proj.size != scan.num_proj * scan.nv * scan.nu:
        raise ValueError("projection buffer does not match the scan")
    vol = np.zeros(img.shape, np.float32) if out is None else out
    _host(vol, img.shape, "out")
    o = opts(**kw)
    check(lib().gct_reconstruct(C.byref(scan.c()), proj.ctypes.data, C.byref(img.c()), vol.ctypes.data, C.byref(o)))
    return vol


class Plan:
    """Device-resident FDK work for one scan shape (gct_plan)."""

    def __init__(self, scan: Scan, img: Image, **kw):
        self.scan, self.img = scan, img
        self._o = opts(**kw)
        h = C.c_void_p()
        check(lib().gct_plan_create(C.byref(scan.c()), C.byref(img.c()), C.byref(self._o), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().gct_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> PlanInfo:
        i = PlanInfo()
        check(lib().gct_plan_get_info(self.h, C.byref(i)))
        return i

    def filter(self, proj, first: int = 0, n: int | None = None, stream=None):
        """Weights and filters views [first, first + n) of the device stack
        `proj` (all views, [V][Nv][Nu])."""
        n = self.scan.num_proj - first if n is None else n
        p = _dev(proj, self.scan.proj_shape, "proj")
        check(lib().gct_plan_filter(self.h, p, first, n, _stream(stream)))

    def backproject(self, vol, first: int = 0, n: int | None = None, accumulate=False, finish=True, stream=None):
        n = self.scan.num_proj - first if n is None else n
        v = _dev(vol, self.img.shape, "vol")
        check(lib().gct_plan_backproject(self.h, v, first, n, int(accumulate), int(finish), _stream(stream)))

    def execute(self, proj, vol, stream=None):
        p, v = _dev(proj, self.scan.proj_shape, "proj"), _dev(vol, self.img.shape, "vol")
        check(lib().gct_plan_execute(self.h, p, v, _stream(stream)))

    def execute_host(self, proj_host, proj_dev, vol_dev, vol_host, chunk_views: int = 0, stream=None):
        """Host projections in, host volume out, the upload pipelined with the
        filter and backprojection (gct_plan_execute_host); proj_dev and
        vol_dev are the device buffers it works in."""
        ph, vh = _host(proj_host, self.scan.proj_shape, "proj_host"), _host(vol_host, self.img.shape, "vol_host")
        p, v = _dev(proj_dev, self.scan.proj_shape, "proj_dev"), _dev(vol_dev, self.img.shape, "vol_dev")
        check(lib().gct_plan_execute_host(self.h, ph, p, v, vh, int(chunk_views), _stream(stream)))

    def load_filtered(self, filt, stream=None):
        f = _dev(filt, self.scan.proj_shape, "filt")
        check(lib().gct_plan_load_filtered(self.h, f, _stream(stream)))

    def store_filtered(self, filt, stream=None):
        f = _dev(filt, self.scan.proj_shape, "filt")
        check(lib().gct_plan_store_filtered(self.h, f, _stream(stream)))

    def forward_project(self, vol, proj, stream=None):
        v, p = _dev(vol, self.img.shape, "vol"), _dev(proj, self.scan.proj_shape, "proj")
        check(lib().gct_plan_forward_project(self.h, v, p, _stream(stream)))

    def autotune(self, vol):
        """Times the backprojection variants (vol is scratch: overwritten)."""
        v = _dev(vol, self.img.shape, "vol")
        zu, ms = C.c_uint32(), C.c_float()
        check(lib().gct_plan_autotune(self.h, v, C.byref(zu), C.byref(ms)))
        return zu.value, ms.value


def _batch_args(plans, projs, vols):
    """Device addresses of one stack and one volume per plan, checked."""
    if len(projs) != len(plans) or len(vols) != len(plans):
        raise ValueError(f"{len(plans)} plans need as many projection stacks and volumes "
                         f"(got {len(projs)} and {len(vols)})")
    pa = (C.c_void_p * len(plans))(*[_dev(p, pl.scan.proj_shape, f"projs[{i}]")
                                      for i, (p, pl) in enumerate(zip(projs, plans))])
    va = (C.c_void_p * len(plans))(*[_dev(v, pl.img.shape, f"vols[{i}]")
                                      for i, (v, pl) in enumerate(zip(vols, plans))])
    return pa, va


class Batch:
    def __init__(self, plans):
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p.h.value for p in self.plans])
        h = C.c_void_p()
        check(lib().gct_batch_create(arr, len(self.plans), C.byref(h)))
        self.h = h

    def execute(self, projs, vols, use_graph=True, stream=None):
        pa, va = _batch_args(self.plans, projs, vols)
        check(lib().gct_batch_execute(self.h, pa, va, int(use_graph), _stream(stream)))

    def close(self):
        if getattr(self, "h", None):
            lib().gct_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NOALLOC_MAX_SCANS = 1024   # GCT_NOALLOC_MAX_SCANS


def batch_execute_noalloc(plans, projs, vols, stream=None):
    """gct_batch_execute_noalloc: the plans' scans backprojected by the
    persistent batch kernels with the work table and counters in the
    library's fixed device array -- no device allocation."""
    plans = list(plans)
    pa, va = _batch_args(plans, projs, vols)
    arr = (C.c_void_p * len(plans))(*[p.h.value for p in plans])
    check(lib().gct_batch_execute_noalloc(arr, len(plans), pa, va, _stream(stream)))


def _stream(s):
    if s is None:
        return None
    if isinstance(s, int):
        return s
    return s.cuda_stream


# ---- phantoms ----------------------------------------------------------------

def phantom_random(seed: int, count: int, dim: int, radius: float, axis=(0.05, 0.35), dens=(-0.2, 1.0)) -> np.ndarray:
    p = np.zeros((count, 10), np.float64)
    check(lib().gct_phantom_random(seed, count, dim, radius, axis[0], axis[1], dens[0], dens[1], p.ctypes.data))
    return p


def phantom_project(scan: Scan, params: np.ndarray, i0: float = 0.0, noise_seed: int = 0, nthreads: int = 0):
    params = np.ascontiguousarray(params, np.float64)
    proj = np.zeros(scan.proj_shape, np.float32)
    dim = 2 if is_1d(scan.kind) else 3
    check(lib().gct_phantom_project(C.byref(scan.c()), dim, params.ctypes.data, params.shape[0], i0, noise_seed,
                                    proj.ctypes.data, nthreads))
    return proj


def phantom_rasterize(img: Image, params: np.ndarray) -> np.ndarray:
    params = np.ascontiguousarray(params, np.float64)
    vol = np.zeros(img.shape, np.float32)
    check(lib().gct_phantom_rasterize(C.byref(img.c()), params.ctypes.data, params.shape[0], vol.ctypes.data))
    return vol


def phantom_project_device(scan: Scan, params: np.ndarray, out, i0: float = 0.0, noise_seed: int = 0, stream=None):
    """Projections of a phantom written into the device tensor `out`
    (shape scan.proj_shape, float32) by the GPU generator."""
    params = np.ascontiguousarray(params, np.float64)
    if tuple(out.shape) != tuple(scan.proj_shape) or out.dtype != _torch().float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape {scan.proj_shape}")
    dim = 2 if is_1d(scan.kind) else 3
    check(lib().gct_phantom_project_device(C.byref(scan.c()), dim, params.ctypes.data, params.shape[0], i0,
                                           noise_seed, _dev(out, scan.proj_shape, "out"), _stream(stream)))
    return out


def phantom_rasterize_device(img: Image, params: np.ndarray, out, stream=None):
    params = np.ascontiguousarray(params, np.float64)
    if tuple(out.shape) != tuple(img.shape) or out.dtype != _torch().float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape {img.shape}")
    check(lib().gct_phantom_rasterize_device(C.byref(img.c()), params.ctypes.data, params.shape[0],
                                             _dev(out, img.shape, "out"), _stream(stream)))
    return out


# ---- array files (MetaImage / Dennerlein / raw) ---------------------------------

FILE_AUTO, FILE_META, FILE_DEN, FILE_RAW = range(4)
_ELEM_DTYPES = [np.uint8, np.int8, np.uint16, np.int16, np.uint32, np.int32, np.uint64, np.int64, np.float32,
                np.float64]
ELEM_NAMES = ["uint8", "int8", "uint16", "int16", "uint32", "int32", "uint64", "int64", "float32", "float64"]


def elem_type(dtype) -> int:
    dt = np.dtype(dtype)
    for i, d in enumerate(_ELEM_DTYPES):
        if np.dtype(d) == dt:
            return i
    raise ValueError(f"unsupported element type {dt}")


def meta_type_name(dtype) -> str:
    return lib().gct_meta_type_name(elem_type(dtype)).decode()


def read_metadata(filename, fmt: int = FILE_AUTO, raw_md: ArrayMetadata | None = None) -> ArrayMetadata:
    md = ArrayMetadata()
    check(lib().gct_file_read_metadata(os.fsencode(filename), fmt, C.byref(raw_md) if raw_md else None, C.byref(md)))
    return md


def read_array(filename, dtype=np.float32, series: bool = False, fmt: int = FILE_AUTO,
               raw_md: ArrayMetadata | None = None) -> np.ndarray:
    """Elements of an array file (or of its series) converted to `dtype`, flat."""
    t = elem_type(dtype)
    n = C.c_uint64()
    raw = C.byref(raw_md) if raw_md else None
    check(lib().gct_file_read(os.fsencode(filename), fmt, int(series), raw, t, None, 0, C.byref(n)))
    out = np.empty(n.value, dtype)
    check(lib().gct_file_read(os.fsencode(filename), fmt, int(series), raw, t, out.ctypes.data, n.value,
                              C.byref(n)))
    return out


def write_meta(filename, data: np.ndarray, dim_size=None, spacing=None, offset=None):
    """MetaImage writer; dim_size defaults to the array's shape reversed (x fastest)."""
    data = np.ascontiguousarray(data)
    dims = list(dim_size) if dim_size is not None else list(reversed(data.shape))
    nd = len(dims)
    sp = np.asarray(spacing if spacing is not None else [0.0] * nd, np.float32)
    of = np.asarray(offset if offset is not None else [0.0] * nd, np.float32)
    d = np.asarray(dims, np.uint32)
    check(lib().gct_meta_write(os.fsencode(filename), nd, d.ctypes.data, sp.ctypes.data, of.ctypes.data,
                               elem_type(data.dtype), data.ctypes.data, data.size))


def write_den(filename, data: np.ndarray, rows: int, cols: int, slices: int):
    data = np.ascontiguousarray(data)
    check(lib().gct_den_write(os.fsencode(filename), rows, cols, slices, elem_type(data.dtype), data.ctypes.data,
                              data.size))


def write_raw(filename, data: np.ndarray):
    data = np.ascontiguousarray(data)
    check(lib().gct_raw_write(os.fsencode(filename), elem_type(data.dtype), data.ctypes.data, data.size))


def read_projections(filename, proj_dim: int, series: bool = False, dtype=np.float32) -> np.ndarray:
    """ProjectionReader::read (projection_reader.h:18-36): [views][rows][cols]
    for 2-D detectors, [views][cols] for 1-D ones."""
    md = read_metadata(filename)
    pix = read_array(filename, dtype, series)
    if proj_dim == 1:
        return pix.reshape(-1, md.num_cols)
    return pix.reshape(-1, md.num_rows, md.num_cols)


def write_image(img: "Image", vol: np.ndarray, filename):
    """IO::write(ReconImg, filename) (img_writer.h:20-42): .den, or MetaImage
    with the voxel spacing and the first voxel centre as the offset."""
    ext = os.path.splitext(str(filename))[1].lower()
    vol = np.ascontiguousarray(vol)
    if ext == ".den":
        write_den(filename, vol, img.ny, img.nx, img.nz if img.dim == 3 else 1)
    elif ext in (".mha", ".mhd"):
        n = [img.nx, img.ny] + ([img.nz] if img.dim == 3 else [])
        w = [img.dx, img.dy] + ([img.dz] if img.dim == 3 else [])
        mid = list(img.mid)[:img.dim]
        first = [np.float32(np.float32(m) - (np.float32(k) - np.float32(1.0)) * np.float32(d) / np.float32(2.0))
                 for m, k, d in zip(mid, n, w)]
        write_meta(filename, vol, n, w, first)
    else:
        raise ValueError("Currently images can be saved to den or meta image file.")


def _torch():
    import torch
    return torch


def device_count() -> int:
    return int(lib().gct_device_count())


class Server:
    """Continuous-batching reconstruction server on one device (gct_server)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, max_inflight: int = 0, num_streams: int = 0,
                 plan_cache_entries: int = 0, max_bypass: int = 0):
        o = ServerOpts(device, hbm_budget_bytes, max_inflight, num_streams, plan_cache_entries, max_bypass)
        h = C.c_void_p()
        check(lib().gct_server_create(C.byref(o), C.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, scan: Scan, img: Image, proj: np.ndarray, out: np.ndarray, reproject: bool = False, **kw) -> int:
        """Queues one request; proj and out must stay alive until wait() returns."""
        proj = np.ascontiguousarray(proj, np.float32)
        assert out.dtype == np.float32 and out.flags.c_contiguous and out.size == img.nvox
        o = opts(**kw)
        t = C.c_uint64()
        check(lib().gct_server_submit(self.h, C.byref(scan.c()), C.byref(img.c()), C.byref(o), proj.ctypes.data,
                                      out.ctypes.data, int(reproject), C.byref(t)))
        self._keep[t.value] = (proj, out, scan)
        return t.value

    def wait(self, ticket: int):
        """(latency_ms, reprojection error or -1); raises GctError if the request failed."""
        lat, err = C.c_double(), C.c_double()
        st = lib().gct_server_wait(self.h, ticket, C.byref(lat), C.byref(err))
        self._keep.pop(ticket, None)
        check(st)
        return lat.value, err.value

    def poll(self, ticket: int):
        lat, err = C.c_double(), C.c_double()
        st = lib().gct_server_poll(self.h, ticket, C.byref(lat), C.byref(err))
        if st == 7:
            return None
        self._keep.pop(ticket, None)
        check(st)
        return lat.value, err.value

    def stats(self) -> dict:
        s = ServerStats()
        check(lib().gct_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().gct_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
