"""gdiamond_amd -- MI355X (gfx950) 3D Yee FDTD time-stepping library.

Python binding of the C ABI in include/gdiamond/gd_capi.h (library
``gdiamond_amd/libgdiamond_hip.so``, built in-tree by ``gdiamond_amd.build``).
There is no CPU fallback: every stepping call runs the HIP kernels, and
importing a missing library raises.

The reference's GPU path is the member functions of ``gdiamond::gDiamond``
(gDiamond/gdiamond.hpp:76-177); its C++ drop-in is include/gdiamond/gdiamond.hpp.
This module exposes the same device state to Python as :class:`Context`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgdiamond_hip.so")

EX, EY, EZ, HX, HY, HZ = range(6)
COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
COEFF_NAMES = ("Cax", "Cbx", "Cay", "Cby", "Caz", "Cbz",
               "Dax", "Dbx", "Day", "Dby", "Daz", "Dbz")
PLAIN, TILED, FUSED = 0, 1, 2
F32, F64, F16, BF16 = 0, 1, 2, 3  # storage formats
MAT_AUTO, MAT_DENSE, MAT_INDEXED, MAT_FULL = 0, 1, 2, 3  # material formats
BATCH_PERSISTENT, BATCH_GRAPH, BATCH_PER_INSTANCE = 0, 1, 2  # batch run modes

GD_OK = 0
ERRORS = {-1: "GD_ERR_ARG", -2: "GD_ERR_STATE", -3: "GD_ERR_HIP", -4: "GD_ERR_NOMEM",
          -5: "GD_ERR_UNSUPPORTED", -6: "GD_ERR_NODEVICE"}


class GdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Tiling(C.Structure):
    _fields_ = [("blt", C.c_int), ("rows", C.c_int), ("waves", C.c_int), ("zchunk", C.c_int)]

    def as_dict(self):
        return {"blt": self.blt, "rows": self.rows, "waves": self.waves, "zchunk": self.zchunk}


_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_lib = None


class PointSource(C.Structure):
    """gd_point_source: a point source whose value table is a device array."""
    _fields_ = [("comp", C.c_int), ("i", C.c_int), ("j", C.c_int), ("k", C.c_int),
                ("table", C.c_void_p)]


class NoallocLaunch(C.Structure):
    """gd_noalloc_launch: one TILED launch on caller-owned device buffers
    (gd_tiled_launch_noalloc). Pointer fields other than `sources` and
    `f_cycles_per_step` are device addresses."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_float),
                ("fields_in", C.c_void_p * 6), ("fields_out", C.c_void_p * 6),
                ("cb", C.c_void_p), ("material_id", C.c_void_p),
                ("material_table", C.c_void_p), ("db", C.c_float), ("nsrc", C.c_int),
                ("sources", C.POINTER(PointSource)), ("n0", C.c_int), ("nsteps", C.c_int),
                ("dft_mask", C.c_uint), ("dft_axis", C.c_int), ("dft_index", C.c_int),
                ("nf", C.c_int), ("f_cycles_per_step", _DP), ("dft_scale", C.c_double),
                ("dft_re", C.c_void_p), ("dft_im", C.c_void_p), ("stream", C.c_void_p)]


NOALLOC_MAX_STEPS = 4
NOALLOC_MAX_FREQS = 16
NOALLOC_MAX_SNAP_CELLS = 1 << 21

# name -> (restype, argtypes)
_SIGS = {
    "gd_version": (C.c_char_p, []),
    "gd_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "gd_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, C.c_size_t, C.c_size_t,
                            C.c_float]),
    "gd_destroy": (C.c_int, [C.c_void_p]),
    "gd_last_error": (C.c_char_p, [C.c_void_p]),
    "gd_set_storage": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "gd_get_storage": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gd_get_fields_f64": (C.c_int, [C.c_void_p, _DP * 6]),
    "gd_get_field_f64": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "gd_set_coefficients": (C.c_int, [C.c_void_p, _FP * 12]),
    "gd_set_isotropic": (C.c_int, [C.c_void_p, _FP, C.c_float]),
    "gd_set_isotropic_eps": (C.c_int, [C.c_void_p, _FP, C.c_float]),
    "gd_set_isotropic_indexed": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8), _FP, C.c_int,
                                           C.c_float]),
    "gd_set_material_format": (C.c_int, [C.c_void_p, C.c_int]),
    "gd_get_material_format": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gd_add_source": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, _FP,
                                C.c_size_t]),
    "gd_clear_sources": (C.c_int, [C.c_void_p]),
    "gd_set_fields": (C.c_int, [C.c_void_p, _FP * 6]),
    "gd_get_fields": (C.c_int, [C.c_void_p, _FP * 6]),
    "gd_get_field": (C.c_int, [C.c_void_p, C.c_int, _FP]),
    "gd_get_plane_xy": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, _FP]),
    "gd_reset": (C.c_int, [C.c_void_p]),
    "gd_get_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "gd_dft_config": (C.c_int, [C.c_void_p, C.c_uint, C.c_int, C.c_size_t, _DP, C.c_size_t,
                                C.c_double]),
    "gd_dft_get": (C.c_int, [C.c_void_p, C.c_int, _DP, _DP]),
    "gd_dft_disable": (C.c_int, [C.c_void_p]),
    "gd_run": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.POINTER(Tiling)]),
    "gd_run_async": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.POINTER(Tiling)]),
    "gd_sync": (C.c_int, [C.c_void_p]),
    "gd_plan_tiling": (C.c_int, [C.c_void_p, C.POINTER(Tiling)]),
    "gd_plan_tiling_for": (C.c_int, [C.c_size_t, C.c_size_t, C.c_size_t, C.c_int,
                                     C.POINTER(Tiling)]),
    "gd_last_run_stats": (C.c_int, [C.c_void_p, _DP, _DP, C.POINTER(C.c_size_t)]),
    "gd_set_launch_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "gd_device_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "gd_calibration_copy": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, _DP]),
    "gd_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_size_t),
                                  C.c_float]),
    "gd_batch_destroy": (C.c_int, [C.c_void_p]),
    "gd_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "gd_batch_set_isotropic": (C.c_int, [C.c_void_p, C.c_int, _FP, C.c_float]),
    "gd_batch_set_isotropic_eps": (C.c_int, [C.c_void_p, C.c_int, _FP, C.c_float]),
    "gd_batch_add_source": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_size_t, C.c_size_t,
                                      C.c_size_t, _FP, C.c_size_t]),
    "gd_batch_set_fields": (C.c_int, [C.c_void_p, C.c_int, _FP * 6]),
    "gd_batch_get_fields": (C.c_int, [C.c_void_p, C.c_int, _FP * 6]),
    "gd_batch_reset": (C.c_int, [C.c_void_p]),
    "gd_batch_get_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "gd_batch_num_classes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "gd_batch_class_info": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_int)]),
    "gd_batch_set_class_tiling": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Tiling)]),
    "gd_batch_get_class_tiling": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Tiling)]),
    "gd_batch_autotune": (C.c_int, [C.c_void_p, C.c_int]),
    "gd_batch_run": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int]),
    "gd_batch_last_run_stats": (C.c_int, [C.c_void_p, _DP, C.POINTER(C.c_size_t)]),
    "gd_batch_device_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "gd_nn_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, C.c_size_t, C.c_size_t]),
    "gd_nn_destroy": (C.c_int, [C.c_void_p]),
    "gd_nn_last_error": (C.c_char_p, [C.c_void_p]),
    "gd_nn_num_params": (C.c_int, [C.POINTER(C.c_size_t)]),
    "gd_nn_set_params": (C.c_int, [C.c_void_p, _FP, C.c_size_t]),
    "gd_nn_get_params": (C.c_int, [C.c_void_p, _FP, C.c_size_t]),
    "gd_nn_set_input": (C.c_int, [C.c_void_p, _FP]),
    "gd_nn_set_target": (C.c_int, [C.c_void_p, _FP, C.POINTER(C.c_uint8)]),
    "gd_nn_forward": (C.c_int, [C.c_void_p, _FP]),
    "gd_nn_train_step": (C.c_int, [C.c_void_p, C.c_float, _FP, _FP]),
    "gd_nn_mfma_selftest": (C.c_int, [_FP, _FP, _FP]),
    "gd_restrict_fields": (C.c_int, [C.c_void_p, _FP * 6]),
    "gd_run_corrected": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, _FP,
                                   C.POINTER(Tiling)]),
    "gd_field_rgb": (C.c_int, [_FP, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_uint8)]),
    "gd_save_field_png": (C.c_int, [_FP, C.c_char_p, C.c_int, C.c_int, C.c_float]),
    "gd_save_mask_png": (C.c_int, [C.POINTER(C.c_uint8), C.c_char_p, C.c_int, C.c_int]),
    "gd_write_png_rgb": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_uint8)]),
    "gd_default_problem_coefficients": (C.c_int, [C.c_size_t, C.c_size_t, C.c_size_t, _FP * 12]),
    "gd_default_mz_table": (C.c_int, [C.c_size_t, _FP]),
    "gd_sphere_medium": (C.c_int, [C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint64, C.c_int,
                                   C.c_double, C.c_double, C.c_double, C.c_double, _FP]),
    "gd_gaussian_pulse_table": (C.c_int, [C.c_size_t, C.c_double, C.c_double, C.c_double, _FP]),
    "gd_tiled_launch_noalloc": (C.c_int, [C.POINTER(NoallocLaunch)]),
}


def exported_symbols():
    """Names of every C-ABI entry point this binding expects in the library."""
    return sorted(_SIGS)


def lib():
    """Load libgdiamond_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m gdiamond_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def version() -> str:
    return lib().gd_version().decode()


def device_count() -> int:
    n = C.c_int(0)
    lib().gd_device_count(C.byref(n))
    return n.value


def plan_tiling_for(nx, ny, nz, storage=F32):
    """The planner's TILED tiling for a grid and storage format (no device
    needed)."""
    t = Tiling()
    rc = lib().gd_plan_tiling_for(int(nx), int(ny), int(nz), int(storage), C.byref(t))
    if rc != 0:
        raise GdError(rc, "bad grid or storage format")
    return t.as_dict()


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} elements, got {a.size}")
    return a


def _fp(a):
    return a.ctypes.data_as(_FP)


def _dp(a):
    return a.ctypes.data_as(_DP)


def _check_host(rc, what="host call"):
    if rc != GD_OK:
        raise GdError(rc, what)


# ---------------------------------------------------------------- host builders


def default_problem_coefficients(nx, ny, nz):
    """Reference default problem (gdiamond.hpp:20-73, utils.h:542-690): 12 arrays (nz,ny,nx)."""
    out = [np.empty((nz, ny, nx), np.float32) for _ in range(12)]
    _check_host(lib().gd_default_problem_coefficients(nx, ny, nz,
                                                      (_FP * 12)(*[_fp(a) for a in out])))
    return out


def default_mz_table(nsteps):
    """Reference default source Mz(t) (gdiamond.hpp:1189-1191), t = 0..nsteps-1."""
    out = np.empty(nsteps, np.float32)
    _check_host(lib().gd_default_mz_table(nsteps, _fp(out)))
    return out


def sphere_medium(nx, ny, nz, seed, nspheres=200, rmin=4.0, rmax=20.0, emin=1.0, emax=12.0):
    """Seeded random dielectric spheres in vacuum; eps_r array shaped (nz, ny, nx)."""
    out = np.empty((nz, ny, nx), np.float32)
    _check_host(lib().gd_sphere_medium(nx, ny, nz, seed, nspheres, rmin, rmax, emin, emax,
                                       _fp(out)))
    return out


def gaussian_pulse_table(nsteps, t0=30.0, tau=10.0, f_cps=0.05):
    """J(n) = exp(-((n-t0)/tau)^2) sin(2 pi f n), double then float."""
    out = np.empty(nsteps, np.float32)
    _check_host(lib().gd_gaussian_pulse_table(nsteps, t0, tau, f_cps, _fp(out)))
    return out


def field_rgb(u, vmax):
    """The reference's field colormap: (ny, nx) array -> (ny, nx, 3) uint8."""
    a = np.ascontiguousarray(u, np.float32)
    ny, nx = a.shape
    out = np.empty((ny, nx, 3), np.uint8)
    _check_host(lib().gd_field_rgb(_fp(a), nx, ny, float(vmax),
                                   out.ctypes.data_as(C.POINTER(C.c_uint8))), "gd_field_rgb")
    return out


def save_field_png(u, filename, vmax):
    a = np.ascontiguousarray(u, np.float32)
    ny, nx = a.shape
    _check_host(lib().gd_save_field_png(_fp(a), str(filename).encode(), nx, ny, float(vmax)),
                "gd_save_field_png")


def save_mask_png(mask, filename):
    m = np.ascontiguousarray(mask, np.uint8)
    ny, nx = m.shape
    _check_host(lib().gd_save_mask_png(m.ctypes.data_as(C.POINTER(C.c_uint8)),
                                       str(filename).encode(), nx, ny), "gd_save_mask_png")


def iso_cb_from_eps(eps_r, dt, dx):
    """The library's isotropic coefficients: cb = dt/(eps*dx), db = dt/dx (double, then float)."""
    eps = np.asarray(eps_r, np.float64)
    cb = (np.float64(dt) / (eps * np.float64(dx))).astype(np.float32)
    db = np.float32(np.float64(dt) / np.float64(dx))
    return cb, db


# ---------------------------------------------------------------- device context


class Context:
    """One device-resident FDTD problem (a ``gd_ctx``)."""

    def __init__(self, nx, ny, nz, dx=1.0):
        self.shape = (int(nx), int(ny), int(nz))
        self.N = int(nx) * int(ny) * int(nz)
        h = C.c_void_p()
        rc = lib().gd_create(C.byref(h), nx, ny, nz, dx)
        if rc != GD_OK:
            raise GdError(rc, lib().gd_last_error(None).decode())
        self._h = h
        self._dft = None

    # -- plumbing
    def _ok(self, rc):
        if rc != GD_OK:
            raise GdError(rc, lib().gd_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().gd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def zyx(self):
        nx, ny, nz = self.shape
        return (nz, ny, nx)

    # -- storage
    def set_storage(self, storage, scale_exp=None):
        se = None
        if scale_exp is not None:
            se = (C.c_int * 6)(*[int(v) for v in scale_exp])
        self._ok(lib().gd_set_storage(self._h, int(storage), se))

    def get_storage(self):
        st, se = C.c_int(), (C.c_int * 6)()
        self._ok(lib().gd_get_storage(self._h, C.byref(st), se))
        return st.value, list(se)

    def get_fields_f64(self):
        out = [np.empty(self.zyx, np.float64) for _ in range(6)]
        self._ok(lib().gd_get_fields_f64(self._h, (_DP * 6)(*[_dp(a) for a in out])))
        return out

    # -- material
    def set_coefficients(self, coeffs):
        arrs = [_f32(c, self.N) for c in coeffs]
        if len(arrs) != 12:
            raise ValueError("need 12 coefficient arrays")
        self._ok(lib().gd_set_coefficients(self._h, (_FP * 12)(*[_fp(a) for a in arrs])))

    def set_isotropic(self, cb, db):
        a = _f32(cb, self.N)
        self._ok(lib().gd_set_isotropic(self._h, _fp(a), float(db)))

    def set_isotropic_eps(self, eps_r, dt):
        a = _f32(eps_r, self.N)
        self._ok(lib().gd_set_isotropic_eps(self._h, _fp(a), float(dt)))

    def set_isotropic_indexed(self, ids, table, db):
        i = np.ascontiguousarray(ids, np.uint8).ravel()
        if i.size != self.N:
            raise ValueError(f"expected {self.N} material ids, got {i.size}")
        t = np.ascontiguousarray(table, np.float32).ravel()
        self._ok(lib().gd_set_isotropic_indexed(self._h, i.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                _fp(t), t.size, float(db)))

    def set_material_format(self, fmt):
        self._ok(lib().gd_set_material_format(self._h, int(fmt)))

    def get_material_format(self):
        f, n = C.c_int(), C.c_int()
        self._ok(lib().gd_get_material_format(self._h, C.byref(f), C.byref(n)))
        return f.value, n.value

    # -- sources
    def add_source(self, comp, i, j, k, table):
        t = _f32(table)
        self._ok(lib().gd_add_source(self._h, int(comp), i, j, k, _fp(t), t.size))

    def clear_sources(self):
        self._ok(lib().gd_clear_sources(self._h))

    # -- fields
    def set_fields(self, fields):
        arrs = [_f32(f, self.N) for f in fields]
        self._ok(lib().gd_set_fields(self._h, (_FP * 6)(*[_fp(a) for a in arrs])))

    def get_fields(self):
        out = [np.empty(self.zyx, np.float32) for _ in range(6)]
        self._ok(lib().gd_get_fields(self._h, (_FP * 6)(*[_fp(a) for a in out])))
        return out

    def get_field(self, comp):
        out = np.empty(self.zyx, np.float32)
        self._ok(lib().gd_get_field(self._h, int(comp), _fp(out)))
        return out

    def get_plane_xy(self, comp, k):
        nx, ny, _ = self.shape
        out = np.empty((ny, nx), np.float32)
        self._ok(lib().gd_get_plane_xy(self._h, int(comp), k, _fp(out)))
        return out

    def reset(self):
        self._ok(lib().gd_reset(self._h))

    @property
    def step(self):
        s = C.c_size_t()
        self._ok(lib().gd_get_step(self._h, C.byref(s)))
        return s.value

    # -- DFT
    def dft_config(self, comps, axis, index, f_cps, scale=1.0):
        mask = 0
        for c in comps:
            mask |= 1 << int(c)
        f = np.ascontiguousarray(f_cps, np.float64)
        self._ok(lib().gd_dft_config(self._h, mask, int(axis), int(index), _dp(f), f.size,
                                     float(scale)))
        nx, ny, nz = self.shape
        plane = {0: ny * nz, 1: nx * nz, 2: nx * ny}[int(axis)]
        self._dft = (sorted(int(c) for c in comps), f.size, plane)

    def dft_get(self, comp):
        _, nf, plane = self._dft
        re = np.empty((nf, plane), np.float64)
        im = np.empty((nf, plane), np.float64)
        self._ok(lib().gd_dft_get(self._h, int(comp), _dp(re), _dp(im)))
        return re, im

    def dft_disable(self):
        self._ok(lib().gd_dft_disable(self._h))
        self._dft = None

    # -- stepping
    def run(self, nsteps, variant=TILED, tiling=None, wait=True):
        """Advance nsteps steps. wait=False enqueues the run and returns at
        once (gd_run_async); sync() or the next call that reads or changes the
        context completes it."""
        t = None
        if tiling is not None:
            t = Tiling(**tiling) if isinstance(tiling, dict) else tiling
        fn = lib().gd_run if wait else lib().gd_run_async
        self._ok(fn(self._h, int(variant), int(nsteps), C.byref(t) if t is not None else None))

    def sync(self):
        """Waits for a run enqueued with run(..., wait=False)."""
        self._ok(lib().gd_sync(self._h))

    def plan_tiling(self):
        t = Tiling()
        self._ok(lib().gd_plan_tiling(self._h, C.byref(t)))
        return t.as_dict()

    def set_launch_timing(self, enable=True):
        self._ok(lib().gd_set_launch_timing(self._h, 1 if enable else 0))

    def last_run_stats(self):
        tot, ker, nl = C.c_double(), C.c_double(), C.c_size_t()
        self._ok(lib().gd_last_run_stats(self._h, C.byref(tot), C.byref(ker), C.byref(nl)))
        return {"total_ms": tot.value, "kernel_ms": ker.value, "nlaunch": nl.value}

    def device_bytes(self):
        b = C.c_size_t()
        self._ok(lib().gd_device_bytes(self._h, C.byref(b)))
        return b.value

    def restrict_fields(self):
        """Fields restricted to the grid of twice the cell size (Yee-consistent
        edge / face means), 6 arrays shaped (nz//2, ny//2, nx//2)."""
        nx, ny, nz = self.shape
        out = [np.empty((nz // 2, ny // 2, nx // 2), np.float32) for _ in range(6)]
        self._ok(lib().gd_restrict_fields(self._h, (_FP * 6)(*[_fp(a) for a in out])))
        return out

    def run_corrected(self, net, nsteps, interval, cfg, tiling=None):
        """TILED run with the corrector `net` applied every `interval` steps;
        cfg = (se, sh, sd, re, rh, dx/dt)."""
        c = _f32(cfg, 6)
        t = None
        if tiling is not None:
            t = Tiling(**tiling) if isinstance(tiling, dict) else tiling
        self._ok(lib().gd_run_corrected(self._h, net._h, int(nsteps), int(interval), _fp(c),
                                        C.byref(t) if t is not None else None))

    def calibration_copy(self, n, reps=10):
        ms = C.c_double()
        self._ok(lib().gd_calibration_copy(self._h, n, reps, C.byref(ms)))
        return ms.value


# ---------------------------------------------------------------- batches


class Batch:
    """Many independent FP32 problems advanced together (a ``gd_batch``)."""

    def __init__(self, dims, dx=1.0):
        self.dims = [tuple(int(v) for v in d) for d in dims]
        arr = (C.c_size_t * (3 * len(self.dims)))(*[v for d in self.dims for v in d])
        h = C.c_void_p()
        rc = lib().gd_batch_create(C.byref(h), len(self.dims), arr, float(dx))
        if rc != GD_OK:
            raise GdError(rc, lib().gd_batch_last_error(None).decode())
        self._h = h

    def _ok(self, rc):
        if rc != GD_OK:
            raise GdError(rc, lib().gd_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().gd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.dims)

    def _n(self, i):
        nx, ny, nz = self.dims[i]
        return nx * ny * nz

    def zyx(self, i):
        nx, ny, nz = self.dims[i]
        return (nz, ny, nx)

    def set_isotropic(self, i, cb, db):
        a = _f32(cb, self._n(i))
        self._ok(lib().gd_batch_set_isotropic(self._h, int(i), _fp(a), float(db)))

    def set_isotropic_eps(self, i, eps_r, dt):
        a = _f32(eps_r, self._n(i))
        self._ok(lib().gd_batch_set_isotropic_eps(self._h, int(i), _fp(a), float(dt)))

    def add_source(self, i, comp, si, sj, sk, table):
        t = _f32(table)
        self._ok(lib().gd_batch_add_source(self._h, int(i), int(comp), si, sj, sk, _fp(t), t.size))

    def set_fields(self, i, fields):
        arrs = [_f32(f, self._n(i)) for f in fields]
        self._ok(lib().gd_batch_set_fields(self._h, int(i), (_FP * 6)(*[_fp(a) for a in arrs])))

    def get_fields(self, i):
        out = [np.empty(self.zyx(i), np.float32) for _ in range(6)]
        self._ok(lib().gd_batch_get_fields(self._h, int(i), (_FP * 6)(*[_fp(a) for a in out])))
        return out

    def reset(self):
        self._ok(lib().gd_batch_reset(self._h))

    @property
    def step(self):
        s = C.c_size_t()
        self._ok(lib().gd_batch_get_step(self._h, C.byref(s)))
        return s.value

    def classes(self):
        n = C.c_int()
        self._ok(lib().gd_batch_num_classes(self._h, C.byref(n)))
        out = []
        for c in range(n.value):
            d = (C.c_size_t * 3)()
            k = C.c_int()
            self._ok(lib().gd_batch_class_info(self._h, c, d, C.byref(k)))
            t = Tiling()
            self._ok(lib().gd_batch_get_class_tiling(self._h, c, C.byref(t)))
            out.append({"dims": tuple(d), "ninst": k.value, "tiling": t.as_dict()})
        return out

    def set_class_tiling(self, cls, tiling):
        t = Tiling(**tiling) if isinstance(tiling, dict) else tiling
        self._ok(lib().gd_batch_set_class_tiling(self._h, int(cls), C.byref(t)))

    def autotune(self, reps=2):
        self._ok(lib().gd_batch_autotune(self._h, int(reps)))
        return self.classes()

    def run(self, nsteps, mode=BATCH_GRAPH):
        self._ok(lib().gd_batch_run(self._h, int(nsteps), int(mode)))

    def last_run_stats(self):
        ms, nl = C.c_double(), C.c_size_t()
        self._ok(lib().gd_batch_last_run_stats(self._h, C.byref(ms), C.byref(nl)))
        return {"total_ms": ms.value, "nlaunch": nl.value}

    def device_bytes(self):
        b = C.c_size_t()
        self._ok(lib().gd_batch_device_bytes(self._h, C.byref(b)))
        return b.value


# ---------------------------------------------------------------- corrector network


NN_LAYER_PARAMS = 27 * 16 * 16 + 16
NN_PARAMS = 3 * NN_LAYER_PARAMS
NN_CIN = (12, 16, 16)
NN_COUT = (16, 16, 12)


def mfma_selftest(a, b):
    """D = A B (16x16) through the MFMA operand layout of the network kernels."""
    A = _f32(a, 256)
    Bm = _f32(b, 256)
    D = np.empty(256, np.float32)
    _check_host(lib().gd_nn_mfma_selftest(_fp(A), _fp(Bm), _fp(D)), "gd_nn_mfma_selftest")
    return D.reshape(16, 16)


class Net:
    """The corrector network on one grid (a ``gd_nn``): tensors are
    channels-last arrays shaped (nz, ny, nx, 16)."""

    def __init__(self, nx, ny, nz):
        self.shape = (int(nx), int(ny), int(nz))
        self.N = int(nx) * int(ny) * int(nz)
        h = C.c_void_p()
        rc = lib().gd_nn_create(C.byref(h), nx, ny, nz)
        if rc != GD_OK:
            raise GdError(rc, lib().gd_nn_last_error(None).decode())
        self._h = h

    def _ok(self, rc):
        if rc != GD_OK:
            raise GdError(rc, lib().gd_nn_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().gd_nn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, p):
        a = _f32(p, NN_PARAMS)
        self._ok(lib().gd_nn_set_params(self._h, _fp(a), a.size))

    def get_params(self):
        out = np.empty(NN_PARAMS, np.float32)
        self._ok(lib().gd_nn_get_params(self._h, _fp(out), out.size))
        return out

    def set_input(self, x):
        a = _f32(x, self.N * 16)
        self._ok(lib().gd_nn_set_input(self._h, _fp(a)))

    def set_target(self, t, mask):
        a = _f32(t, self.N * 16)
        m = np.ascontiguousarray(mask, np.uint8).ravel()
        if m.size != self.N:
            raise ValueError("mask must have one byte per cell")
        self._ok(lib().gd_nn_set_target(self._h, _fp(a), m.ctypes.data_as(C.POINTER(C.c_uint8))))

    def forward(self):
        nx, ny, nz = self.shape
        y = np.empty((nz, ny, nx, 16), np.float32)
        self._ok(lib().gd_nn_forward(self._h, _fp(y)))
        return y

    def train_step(self, lr, want_grads=False):
        loss = C.c_float()
        g = np.empty(NN_PARAMS, np.float32) if want_grads else None
        self._ok(lib().gd_nn_train_step(self._h, float(lr), C.byref(loss),
                                        _fp(g) if g is not None else None))
        return (loss.value, g) if want_grads else loss.value
