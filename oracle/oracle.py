"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Imported only by tests/, by ``__graft_entry__.smoke()`` and by the CPU-baseline
leg of ``bench.py``, always as the reference the GPU results are compared
against. The product package never imports
this module.

Functions mirror the reference CPU path:
  * :func:`run` -- ``update_FDTD_seq`` (gdiamond.hpp:1164-1263) and the OpenMP
    twin ``update_FDTD_omp`` (gdiamond_omp.hpp:110-172), FP32 or FP64.
  * :func:`default_problem_coefficients` -- ``gDiamond::gDiamond``
    (gdiamond.hpp:20-73) + ``set_FDTD_matrices_3D_structure`` (utils.h:542-690).
  * DFT accumulation in the convention of ``update_field_FFT_xy``
    (kernels.cuh:87-113) with FP64 sums.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libgdref.so")

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
COEFF_NAMES = ("Cax", "Cbx", "Cay", "Cby", "Caz", "Cbz",
               "Dax", "Dbx", "Day", "Dby", "Daz", "Dbz")


class _Source(C.Structure):
    _fields_ = [("comp", C.c_int), ("i", C.c_int64), ("j", C.c_int64), ("k", C.c_int64),
                ("table", C.POINTER(C.c_float))]


class _Dft(C.Structure):
    _fields_ = [("ncomp", C.c_int), ("comps", C.c_int * 6), ("axis", C.c_int),
                ("index", C.c_int64), ("nf", C.c_int), ("f_cps", C.POINTER(C.c_double)),
                ("scale", C.c_double), ("re", C.POINTER(C.c_double)),
                ("im", C.POINTER(C.c_double))]


class _Material(C.Structure):
    _fields_ = [("full", C.c_int), ("c", C.POINTER(C.c_float) * 12),
                ("cb", C.POINTER(C.c_float)), ("db", C.c_float)]


_lib = None


def build() -> None:
    """Compile liboracle.so (and the reference driver when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/gDiamond"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_run_f32.restype = C.c_int
        L.or_run_f32.argtypes = [C.c_int64] * 3 + [C.c_float, C.POINTER(C.c_float) * 6,
                                                   C.POINTER(_Material), C.c_int,
                                                   C.POINTER(_Source), C.c_int64, C.c_int64,
                                                   C.POINTER(_Dft), C.c_int]
        L.or_run_f64.restype = C.c_int
        L.or_run_f64.argtypes = [C.c_int64] * 3 + [C.c_double, C.POINTER(C.c_double) * 6,
                                                   C.POINTER(_Material), C.c_int,
                                                   C.POINTER(_Source), C.c_int64, C.c_int64,
                                                   C.POINTER(_Dft), C.c_int]
        L.or_run_f32_jm_arrays.restype = C.c_int
        L.or_run_f32_jm_arrays.argtypes = [C.c_int64] * 3 + [C.c_float,
                                                             C.POINTER(C.c_float) * 6,
                                                             C.POINTER(C.c_float) * 12, C.c_int,
                                                             C.POINTER(_Source), C.c_int64,
                                                             C.c_int64]
        L.or_run_f16.restype = C.c_int
        L.or_run_f16.argtypes = [C.c_int64] * 3 + [C.c_float, C.POINTER(C.c_uint16) * 6,
                                                   C.POINTER(C.c_int), C.POINTER(C.c_float),
                                                   C.c_float, C.c_int, C.POINTER(_Source),
                                                   C.c_int64, C.c_int64, C.POINTER(_Dft),
                                                   C.c_int64, C.c_int]
        L.or_run_h16.restype = C.c_int
        L.or_run_h16.argtypes = [C.c_int64] * 3 + [C.c_float, C.POINTER(C.c_uint16) * 6,
                                                   C.POINTER(C.c_int), C.POINTER(C.c_float),
                                                   C.c_float, C.c_int, C.POINTER(_Source),
                                                   C.c_int64, C.c_int64, C.POINTER(_Dft),
                                                   C.c_int64, C.c_int, C.c_int]
        L.or_f32_to_bf16_n.restype = None
        L.or_f32_to_bf16_n.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_uint16), C.c_int64]
        L.or_bf16_to_f32_n.restype = None
        L.or_bf16_to_f32_n.argtypes = [C.POINTER(C.c_uint16), C.POINTER(C.c_float), C.c_int64]
        L.or_f32_to_f16.restype = C.c_uint16
        L.or_f32_to_f16.argtypes = [C.c_float]
        L.or_f16_to_f32.restype = C.c_float
        L.or_f16_to_f32.argtypes = [C.c_uint16]
        L.or_f32_to_f16_n.restype = None
        L.or_f32_to_f16_n.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_uint16), C.c_int64]
        L.or_f16_to_f32_n.restype = None
        L.or_f16_to_f32_n.argtypes = [C.POINTER(C.c_uint16), C.POINTER(C.c_float), C.c_int64]
        L.or_default_problem_coefficients.restype = None
        L.or_default_problem_coefficients.argtypes = [C.c_int64] * 3 + [C.POINTER(C.c_float) * 12]
        L.or_default_mz.restype = C.c_float
        L.or_default_mz.argtypes = [C.c_int64]
        L.or_default_source_idx.restype = C.c_int64
        L.or_default_source_idx.argtypes = [C.c_int64] * 3
        L.or_dft_weights.restype = None
        L.or_dft_weights.argtypes = [C.c_double, C.c_int64, C.c_double,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check_arr(a, shape, dtype):
    if a.dtype != dtype or a.shape != shape or not a.flags.c_contiguous:
        raise ValueError(f"expected contiguous {dtype} array of shape {shape}, got "
                         f"{a.dtype} {a.shape}")


def run(shape, fields, *, coeffs=None, cb=None, db=None, dx=1.0, sources=(), n0=0,
        nsteps=1, dft=None, threads=1, dtype=np.float32):
    """Advance ``fields`` (list of 6 arrays shaped (nz, ny, nx)) in place.

    Material: ``coeffs`` (12 float32 arrays, reference layout) or the isotropic
    form ``cb`` (float32 array) + scalar ``db``. ``sources`` is a list of
    (comp, i, j, k, table) with ``table`` float32 indexed by absolute step.
    ``dft``: dict(comps=[...], axis=0|1|2, index=int, f_cps=array, scale=float)
    -- returns (re, im) arrays shaped (ncomp, nf, plane) when given.
    """
    nx, ny, nz = shape
    sh = (nz, ny, nx)
    for f in fields:
        _check_arr(f, sh, dtype)
    mat = _Material()
    keep = []
    if coeffs is not None:
        mat.full = 1
        for a in range(12):
            c = np.ascontiguousarray(coeffs[a], dtype=np.float32).reshape(sh)
            keep.append(c)
            mat.c[a] = _fp(c)
    else:
        mat.full = 0
        c = np.ascontiguousarray(cb, dtype=np.float32).reshape(sh)
        keep.append(c)
        mat.cb = _fp(c)
        mat.db = float(db)
    nsrc = len(sources)
    src_arr = (_Source * max(nsrc, 1))()
    for s, (comp, i, j, k, table) in enumerate(sources):
        t = np.ascontiguousarray(table, dtype=np.float32)
        if t.shape[0] < n0 + nsteps:
            raise ValueError("source table shorter than the run")
        keep.append(t)
        src_arr[s].comp, src_arr[s].i, src_arr[s].j, src_arr[s].k = comp, i, j, k
        src_arr[s].table = _fp(t)
    dptr = None
    out = None
    if dft is not None:
        comps = list(dft["comps"])
        f_cps = np.ascontiguousarray(dft["f_cps"], dtype=np.float64)
        axis = int(dft["axis"])
        plane = {0: ny * nz, 1: nx * nz, 2: nx * ny}[axis]
        re = dft.get("re")
        im = dft.get("im")
        if re is None:
            re = np.zeros((len(comps), len(f_cps), plane), np.float64)
            im = np.zeros_like(re)
        d = _Dft()
        d.ncomp = len(comps)
        for a, c in enumerate(comps):
            d.comps[a] = c
        d.axis = axis
        d.index = int(dft["index"])
        d.nf = len(f_cps)
        d.f_cps = _dp(f_cps)
        d.scale = float(dft.get("scale", 1.0))
        d.re = _dp(re)
        d.im = _dp(im)
        keep += [f_cps, re, im]
        dptr = C.pointer(d)
        out = (re, im)
    if dtype == np.float32:
        fa = (C.POINTER(C.c_float) * 6)(*[_fp(f) for f in fields])
        rc = lib().or_run_f32(nx, ny, nz, float(dx), fa, C.byref(mat), nsrc, src_arr,
                              n0, nsteps, dptr, threads)
    else:
        fa = (C.POINTER(C.c_double) * 6)(*[_dp(f) for f in fields])
        rc = lib().or_run_f64(nx, ny, nz, float(dx), fa, C.byref(mat), nsrc, src_arr,
                              n0, nsteps, dptr, threads)
    if rc != 0:
        raise RuntimeError(f"oracle run failed: {rc}")
    return out


def _dft_struct(dft, shape, keep):
    nx, ny, nz = shape
    comps = list(dft["comps"])
    f_cps = np.ascontiguousarray(dft["f_cps"], dtype=np.float64)
    axis = int(dft["axis"])
    plane = {0: ny * nz, 1: nx * nz, 2: nx * ny}[axis]
    re = dft.get("re")
    im = dft.get("im")
    if re is None:
        re = np.zeros((len(comps), len(f_cps), plane), np.float64)
        im = np.zeros_like(re)
    d = _Dft()
    d.ncomp = len(comps)
    for a, c in enumerate(comps):
        d.comps[a] = c
    d.axis = axis
    d.index = int(dft["index"])
    d.nf = len(f_cps)
    d.f_cps = _dp(f_cps)
    d.scale = float(dft.get("scale", 1.0))
    d.re = _dp(re)
    d.im = _dp(im)
    keep += [f_cps, re, im, d]
    return C.pointer(d), (re, im)


def run_f16(shape, fields16, scale_exp, cb, db, dx=1.0, sources=(), n0=0, nsteps=1,
            dft=None, block=0, threads=1, fmt="f16"):
    """FP16-storage twin: fields16 are 6 uint16 (binary16 bit) arrays shaped
    (nz, ny, nx), advanced in place; cb the float32 Cb array. ``block`` 0
    rounds after every half step (untiled update), block B after every B steps
    (temporally blocked update with B levels per launch); ``threads`` host
    threads run the block mode, with results independent of the count.
    ``fmt`` "bf16": the fields hold bfloat16 bit patterns instead (same rules)."""
    nx, ny, nz = shape
    sh = (nz, ny, nx)
    for f in fields16:
        _check_arr(f, sh, np.uint16)
    keep = []
    cb = np.ascontiguousarray(cb, dtype=np.float32).reshape(sh)
    keep.append(cb)
    se = (C.c_int * 6)(*[int(v) for v in scale_exp])
    nsrc = len(sources)
    src_arr = (_Source * max(nsrc, 1))()
    for s, (comp, i, j, k, table) in enumerate(sources):
        t = np.ascontiguousarray(table, dtype=np.float32)
        keep.append(t)
        src_arr[s].comp, src_arr[s].i, src_arr[s].j, src_arr[s].k = comp, i, j, k
        src_arr[s].table = _fp(t)
    dptr, out = (None, None)
    if dft is not None:
        dptr, out = _dft_struct(dft, shape, keep)
    P16 = C.POINTER(C.c_uint16)
    fa = (P16 * 6)(*[f.ctypes.data_as(P16) for f in fields16])
    if fmt not in ("f16", "bf16"):
        raise ValueError("fmt must be 'f16' or 'bf16'")
    rc = lib().or_run_h16(nx, ny, nz, float(dx), fa, se, _fp(cb), float(db),
                          nsrc, src_arr, n0, nsteps, dptr, int(block), int(threads),
                          1 if fmt == "bf16" else 0)
    if rc != 0:
        raise RuntimeError("oracle run failed")
    return out


def f32_to_f16_bits(values):
    v = np.ascontiguousarray(values, np.float32).ravel()
    out = np.empty(v.size, np.uint16)
    lib().or_f32_to_f16_n(_fp(v), out.ctypes.data_as(C.POINTER(C.c_uint16)), v.size)
    return out


def encode_f16(values, scale_exp=0):
    """Storage encoding of float values: binary16 of value * 2^-scale_exp."""
    v = np.ascontiguousarray(values, np.float32) * np.float32(2.0 ** -scale_exp)
    return f32_to_f16_bits(v).reshape(np.shape(values))


def decode_f16(bits, scale_exp=0):
    b = np.ascontiguousarray(bits, np.uint16)
    flat = np.ascontiguousarray(b.ravel())
    out = np.empty(flat.size, np.float32)
    lib().or_f16_to_f32_n(flat.ctypes.data_as(C.POINTER(C.c_uint16)), _fp(out), flat.size)
    return (out * np.float32(2.0 ** scale_exp)).reshape(b.shape)


def f32_to_bf16_bits(values):
    v = np.ascontiguousarray(values, np.float32).ravel()
    out = np.empty(v.size, np.uint16)
    lib().or_f32_to_bf16_n(_fp(v), out.ctypes.data_as(C.POINTER(C.c_uint16)), v.size)
    return out


def encode_bf16(values, scale_exp=0):
    """Storage encoding of float values: bfloat16 of value * 2^-scale_exp."""
    v = np.ascontiguousarray(values, np.float32) * np.float32(2.0 ** -scale_exp)
    return f32_to_bf16_bits(v).reshape(np.shape(values))


def decode_bf16(bits, scale_exp=0):
    b = np.ascontiguousarray(bits, np.uint16)
    flat = np.ascontiguousarray(b.ravel())
    out = np.empty(flat.size, np.float32)
    lib().or_bf16_to_f32_n(flat.ctypes.data_as(C.POINTER(C.c_uint16)), _fp(out), flat.size)
    return (out * np.float32(2.0 ** scale_exp)).reshape(b.shape)


def run_jm_arrays(shape, fields, coeffs, dx, sources, n0, nsteps):
    """Literal reference loop with full J/M arrays (gdiamond.hpp:1164-1221)."""
    nx, ny, nz = shape
    sh = (nz, ny, nx)
    keep = []
    ca = []
    for a in range(12):
        c = np.ascontiguousarray(coeffs[a], dtype=np.float32).reshape(sh)
        keep.append(c)
        ca.append(_fp(c))
    nsrc = len(sources)
    src_arr = (_Source * max(nsrc, 1))()
    for s, (comp, i, j, k, table) in enumerate(sources):
        t = np.ascontiguousarray(table, dtype=np.float32)
        keep.append(t)
        src_arr[s].comp, src_arr[s].i, src_arr[s].j, src_arr[s].k = comp, i, j, k
        src_arr[s].table = _fp(t)
    fa = (C.POINTER(C.c_float) * 6)(*[_fp(f) for f in fields])
    rc = lib().or_run_f32_jm_arrays(nx, ny, nz, float(dx), fa, (C.POINTER(C.c_float) * 12)(*ca),
                                    nsrc, src_arr, n0, nsteps)
    if rc != 0:
        raise RuntimeError("oracle run failed")


def zero_fields(shape, dtype=np.float32):
    nx, ny, nz = shape
    return [np.zeros((nz, ny, nx), dtype) for _ in range(6)]


def default_problem_coefficients(shape):
    nx, ny, nz = shape
    out = [np.empty((nz, ny, nx), np.float32) for _ in range(12)]
    lib().or_default_problem_coefficients(nx, ny, nz,
                                          (C.POINTER(C.c_float) * 12)(*[_fp(a) for a in out]))
    return out


def default_mz_table(nsteps):
    return np.array([lib().or_default_mz(t) for t in range(nsteps)], np.float32)


def default_source_ijk(shape):
    nx, ny, nz = shape
    return nx // 2, ny // 2, nz // 2


def dft_weights(f_cps, n, scale=1.0):
    wc, ws = C.c_double(), C.c_double()
    lib().or_dft_weights(f_cps, n, scale, C.byref(wc), C.byref(ws))
    return wc.value, ws.value


# ---------------------------------------------------------------- reference build


def ref_lib():
    """The reference's own CPU path compiled from /root/reference (oracle/_ref)."""
    if not os.path.exists(REF_LIB_PATH):
        raise FileNotFoundError("oracle/_ref/libgdref.so not built (needs /root/reference)")
    L = C.CDLL(REF_LIB_PATH)
    L.ref_default_coefficients.argtypes = [C.c_int] * 3 + [C.POINTER(C.c_float)]
    L.ref_seq.argtypes = [C.c_int] * 4 + [C.POINTER(C.c_float)]
    L.ref_seq_rate.argtypes = [C.c_int] * 4 + [C.POINTER(C.c_double)]
    L.ref_default_mz.argtypes = [C.c_int]
    L.ref_default_mz.restype = C.c_float
    L.ref_save_field_png.argtypes = [C.POINTER(C.c_float), C.c_char_p, C.c_int, C.c_int,
                                     C.c_float]
    L.ref_save_mask_png.argtypes = [C.POINTER(C.c_uint8), C.c_char_p, C.c_int, C.c_int]
    return L
