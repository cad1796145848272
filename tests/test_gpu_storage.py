"""Field-storage and material formats on the GPU: binary16 fields with a
per-component power-of-two scale, FP64 fields (the error reference of the
fixed-error comparison), and the dense / indexed isotropic material.

* FP16 PLAIN and TILED == oracle.run_f16 bit for bit (fields and DFT), for
  several scales, tilings, run splits and both material formats. The semantic
  (decode exactly, compute in float in the reference operand order, round to
  binary16 once per step) is pinned on the CPU against numpy's float16 in
  test_oracle_f16.py.
* FP32 with the dense (FP32 Cb per cell, what the other parity tests run by
  default) and the indexed material == the oracle on every tiling.
* FP64 PLAIN == oracle.run(dtype=float64) bit for bit (fields and DFT).
* Material-format selection and errors.
* bfloat16 PLAIN and TILED == oracle.run_f16(fmt="bf16") bit for bit.
* Error of FP32, FP16 and BF16 storage against FP64 on the same problem.
"""
import numpy as np
import pytest

from cases import (DT_C1, bits_equal, centre, describe_diff, iso_problem, random_fields,
                   spectral_error)
from test_gpu_parity import TILINGS, say

pytestmark = pytest.mark.gpu

SCALES = [[0, 0, 0, 0, 0, 0], [-4, -4, -3, 2, 3, 1]]


def _f16_problem(gd, shape, T, seed=11):
    eps, cb, db = iso_problem(gd, shape, seed=seed, nspheres=10, rmin=2.0, rmax=5.0)
    ci, cj, ck = centre(shape)
    src = [(2, ci, cj, ck, gd.gaussian_pulse_table(T, t0=12.0, tau=5.0, f_cps=0.06) * 4.0),
           (3, ci - 3, cj + 2, ck - 1, gd.gaussian_pulse_table(T, t0=9.0, tau=4.0, f_cps=0.08))]
    dft = dict(comps=[0, 2, 4], axis=2, index=ck, f_cps=np.array([0.03, 0.06]), scale=0.5)
    return cb, db, src, dft


def _codec(oracle, fmt):
    if fmt == "bf16":
        return oracle.encode_bf16, oracle.decode_bf16
    return oracle.encode_f16, oracle.decode_f16


def _oracle_f16(oracle, shape, init, scale_exp, cb, db, src, T, dft, chunks=None, block=0,
                fmt="f16"):
    """Oracle twin of a 16-bit run; Cb is exact FP32 in both material forms."""
    enc, dec = _codec(oracle, fmt)
    f16 = [enc(init[c], scale_exp[c]) for c in range(6)]
    dft = dict(dft)
    n0 = 0
    for n in (chunks or [T]):
        dft["re"], dft["im"] = oracle.run_f16(shape, f16, scale_exp, cb, db, sources=src,
                                              n0=n0, nsteps=n, dft=dft, block=block, fmt=fmt)
        n0 += n
    re, im = dft["re"], dft["im"]
    vals = [dec(f16[c], scale_exp[c]) for c in range(6)]
    return vals, (re, im)


MATS = {"auto": 0, "dense": 1, "indexed": 2}


def _gpu(gd, shape, storage, scale_exp, cb, db, src, T, dft, init=None, variant=None,
         tiling=None, chunks=None, f64=False, mat="auto"):
    ctx = gd.Context(*shape)
    try:
        ctx.set_storage(storage, scale_exp)
        ctx.set_material_format(MATS[mat])
        ctx.set_isotropic(cb, db)
        if mat != "auto":
            assert ctx.get_material_format()[0] == MATS[mat]
        for s in src:
            ctx.add_source(*s)
        if init is not None:
            ctx.set_fields(init)
        ctx.dft_config(dft["comps"], dft["axis"], dft["index"], dft["f_cps"], dft["scale"])
        for n in (chunks or [T]):
            ctx.run(n, gd.TILED if variant is None else variant, tiling)
        fields = ctx.get_fields_f64() if f64 else ctx.get_fields()
        spec = [ctx.dft_get(c) for c in dft["comps"]]
    finally:
        ctx.close()
    re = np.stack([s[0] for s in spec])
    im = np.stack([s[1] for s in spec])
    return fields, (re, im)


def _assert_same(got, want, label):
    fields, (re, im) = got
    wf, (wre, wim) = want
    for c in range(6):
        assert bits_equal(fields[c], wf[c]), f"{label} comp {c}: {describe_diff(fields[c], wf[c])}"
    assert bits_equal(re, wre), f"{label} DFT re: {describe_diff(re, wre)}"
    assert bits_equal(im, wim), f"{label} DFT im: {describe_diff(im, wim)}"


@pytest.mark.parametrize("mat", ["dense", "indexed"])
@pytest.mark.parametrize("scale_exp", SCALES)
def test_f16_plain_matches_oracle(gd, oracle, scale_exp, mat):
    shape = (41, 30, 23)
    T = 40
    cb, db, src, dft = _f16_problem(gd, shape, T)
    init = random_fields(shape, 5, scale=0.05, boundary_zero=True)
    want = _oracle_f16(oracle, shape, init, scale_exp, cb, db, src, T, dft)
    say("f16 plain", shape, scale_exp, mat)
    got = _gpu(gd, shape, gd.F16, scale_exp, cb, db, src, T, dft, init=init, variant=gd.PLAIN,
               mat=mat)
    _assert_same(got, want, f"f16 plain {mat}")


@pytest.mark.parametrize("mat", ["dense", "indexed"])
@pytest.mark.parametrize("tiling", TILINGS, ids=lambda t: "b{blt}r{rows}w{waves}z{zchunk}".format(**t))
def test_f16_tiled_matches_oracle(gd, oracle, tiling, mat):
    shape = (70, 45, 31)
    T = 24
    scale_exp = SCALES[1]
    cb, db, src, dft = _f16_problem(gd, shape, T, seed=3)
    init = random_fields(shape, 9, scale=0.05, boundary_zero=True)
    want = _oracle_f16(oracle, shape, init, scale_exp, cb, db, src, T, dft,
                       block=tiling["blt"])
    say("f16 tiled", shape, tiling, mat)
    got = _gpu(gd, shape, gd.F16, scale_exp, cb, db, src, T, dft, init=init, variant=gd.TILED,
               tiling=tiling, mat=mat)
    _assert_same(got, want, f"f16 tiled {tiling} {mat}")


@pytest.mark.parametrize("mat", ["dense", "indexed"])
@pytest.mark.parametrize("tiling", [None] + TILINGS,
                         ids=lambda t: "plain" if t is None else
                         "b{blt}r{rows}w{waves}z{zchunk}".format(**t))
def test_f32_material_matches_oracle(gd, oracle, tiling, mat):
    shape = (70, 45, 31)
    T = 24
    cb, db, src, dft = _f16_problem(gd, shape, T, seed=6)
    init = random_fields(shape, 4, scale=0.05, boundary_zero=True)
    wf = [a.copy() for a in init]
    wre, wim = oracle.run(shape, wf, cb=cb, db=db, sources=src, nsteps=T, dft=dft)
    say("f32", mat, shape, tiling)
    got = _gpu(gd, shape, gd.F32, None, cb, db, src, T, dft, init=init,
               variant=gd.PLAIN if tiling is None else gd.TILED, tiling=tiling, mat=mat)
    _assert_same(got, (wf, (wre, wim)), f"f32 {mat} {tiling}")


def test_material_formats(gd, oracle):
    shape = (34, 29, 21)
    nz, ny, nx = shape[2], shape[1], shape[0]
    T = 20
    cb, db, src, dft = _f16_problem(gd, shape, T, seed=12)
    distinct = np.unique(cb.view(np.uint32)).size
    assert 2 <= distinct <= 256
    # AUTO picks the indexed form for a sphere medium, with one table entry per
    # distinct value, for FP32 and 16-bit fields; DENSE on request
    with gd.Context(*shape) as ctx:
        ctx.set_storage(gd.F16, [0] * 6)
        ctx.set_isotropic(cb, db)
        assert ctx.get_material_format() == (gd.MAT_INDEXED, distinct)
    with gd.Context(*shape) as ctx:
        ctx.set_storage(gd.BF16, [0] * 6)
        ctx.set_material_format(gd.MAT_DENSE)
        ctx.set_isotropic(cb, db)
        assert ctx.get_material_format() == (gd.MAT_DENSE, 0)
    with gd.Context(*shape) as ctx:
        ctx.set_isotropic(cb, db)
        assert ctx.get_material_format() == (gd.MAT_INDEXED, distinct)
        with pytest.raises(gd.GdError):
            ctx.set_material_format(gd.MAT_DENSE)  # after the material
    # a continuous medium has too many values: AUTO -> dense, INDEXED -> error
    rng = np.random.default_rng(1)
    cbr = (0.5 / rng.uniform(1.0, 6.0, (nz, ny, nx))).astype(np.float32)
    with gd.Context(*shape) as ctx:
        ctx.set_storage(gd.F16, [0] * 6)
        ctx.set_isotropic(cbr, db)
        assert ctx.get_material_format() == (gd.MAT_DENSE, 0)
    with gd.Context(*shape) as ctx:
        ctx.set_material_format(gd.MAT_INDEXED)
        with pytest.raises(gd.GdError):
            ctx.set_isotropic(cbr, db)
    # FP64 storage has no indexed form
    with gd.Context(*shape) as ctx:
        ctx.set_storage(gd.F64)
        ctx.set_isotropic(cb, db)
        assert ctx.get_material_format()[0] == gd.MAT_DENSE
    with gd.Context(*shape) as ctx:
        ctx.set_storage(gd.F64)
        ctx.set_material_format(gd.MAT_INDEXED)
        with pytest.raises(gd.GdError):
            ctx.set_isotropic(cb, db)
    # explicit ids + table == the expanded array, bit for bit
    table, ids = np.unique(cb, return_inverse=True)
    ids = ids.reshape(cb.shape).astype(np.uint8)
    with gd.Context(*shape) as ctx:
        with pytest.raises(gd.GdError):
            ctx.set_isotropic_indexed(ids, table[:-1], db)  # an id past the table
        ctx.set_isotropic_indexed(ids, table, db)
        assert ctx.get_material_format() == (gd.MAT_INDEXED, table.size)
        for s in src:
            ctx.add_source(*s)
        ctx.run(T, gd.TILED)
        got = ctx.get_fields()
    want = [np.zeros((nz, ny, nx), np.float32) for _ in range(6)]
    oracle.run(shape, want, cb=cb, db=db, sources=src, nsteps=T)
    for c in range(6):
        assert bits_equal(got[c], want[c]), c


def test_f16_runs_continue(gd, oracle):
    shape = (52, 37, 29)
    T = 23
    chunks = [10, 13]
    scale_exp = SCALES[0]
    cb, db, src, dft = _f16_problem(gd, shape, T, seed=8)
    for variant, tiling in ((gd.PLAIN, None), (gd.TILED, dict(blt=3, rows=2, waves=8, zchunk=7))):
        want = _oracle_f16(oracle, shape, [np.zeros((29, 37, 52), np.float32)] * 6, scale_exp, cb,
                           db, src, T, dft, chunks=chunks,
                           block=0 if tiling is None else tiling["blt"])
        say("f16 continue", variant, tiling)
        got = _gpu(gd, shape, gd.F16, scale_exp, cb, db, src, T, dft, variant=variant,
                   tiling=tiling, chunks=chunks)
        _assert_same(got, want, f"f16 continue {variant}")


@pytest.mark.parametrize("mat", ["dense", "indexed"])
@pytest.mark.parametrize("tiling", [None, TILINGS[2], TILINGS[5], TILINGS[4]],
                         ids=lambda t: "plain" if t is None else
                         "b{blt}r{rows}w{waves}z{zchunk}".format(**t))
def test_bf16_matches_oracle(gd, oracle, tiling, mat):
    """bfloat16 storage, PLAIN (rounding every half step) and TILED (once per
    launch) == the oracle's bfloat16 twin, fields and DFT, over a run split in
    two calls."""
    shape = (57, 38, 27)
    T = 21
    scale_exp = SCALES[1]
    cb, db, src, dft = _f16_problem(gd, shape, T, seed=13)
    init = random_fields(shape, 4, scale=0.05, boundary_zero=True)
    chunks = [9, 12]
    want = _oracle_f16(oracle, shape, init, scale_exp, cb, db, src, T, dft, chunks=chunks,
                       block=0 if tiling is None else tiling["blt"], fmt="bf16")
    say("bf16", shape, tiling, mat)
    got = _gpu(gd, shape, gd.BF16, scale_exp, cb, db, src, T, dft, init=init,
               variant=gd.PLAIN if tiling is None else gd.TILED, tiling=tiling, mat=mat,
               chunks=chunks)
    _assert_same(got, want, f"bf16 {tiling} {mat}")


def test_f64_plain_matches_oracle(gd, oracle):
    shape = (39, 31, 26)
    T = 35
    cb, db, src, dft = _f16_problem(gd, shape, T, seed=21)
    init = random_fields(shape, 2, scale=0.1, boundary_zero=True)
    wf = [a.astype(np.float64) for a in init]
    wre, wim = oracle.run(shape, wf, cb=cb, db=db, sources=src, nsteps=T, dft=dft,
                          dtype=np.float64)
    say("f64 plain", shape)
    got = _gpu(gd, shape, gd.F64, None, cb, db, src, T, dft, init=init, variant=gd.PLAIN,
               f64=True)
    _assert_same(got, (wf, (wre, wim)), "f64 plain")


def test_storage_errors(gd):
    shape = (20, 18, 16)
    ctx = gd.Context(*shape)
    try:
        ctx.set_storage(gd.F64)
        cb = np.full((16, 18, 20), 0.3, np.float32)
        ctx.set_isotropic(cb, 0.5)
        with pytest.raises(gd.GdError):
            ctx.run(4, gd.TILED)
        ctx.run(2, gd.PLAIN)
        assert ctx.get_storage()[0] == gd.F64
        with pytest.raises(gd.GdError):
            ctx.set_storage(gd.F16)  # storage is fixed once material is set
    finally:
        ctx.close()
    ctx = gd.Context(*shape)
    try:
        ctx.set_storage(gd.F16, [0] * 6)
        coeffs = gd.default_problem_coefficients(*shape)
        with pytest.raises(gd.GdError):
            ctx.set_coefficients(coeffs)  # binary16 storage is isotropic-only
        with pytest.raises(gd.GdError):
            ctx.set_storage(gd.F16, [0, 0, 0, 0, 0, 101])
    finally:
        ctx.close()


def test_storage_error_vs_f64(gd):
    """FP32 and FP16 storage against the FP64 run of the same problem, by the
    relative DFT-spectrum error ||X - X_ref||_2 / ||X_ref||_2 per component and
    frequency. With the indexed material Cb is exact and only the per-step
    binary16 rounding of the fields remains."""
    shape = (96, 80, 72)
    T = 300
    eps, cb, db = iso_problem(gd, shape, seed=5, nspheres=30, rmin=3.0, rmax=9.0)
    ci, cj, ck = centre(shape)
    src = [(2, ci, cj, ck, gd.gaussian_pulse_table(T, t0=30.0, tau=10.0, f_cps=0.05))]
    dft = dict(comps=[0, 2], axis=2, index=ck + 5, f_cps=np.array([0.03, 0.05, 0.07]),
               scale=1.0)
    _, (ref_re, ref_im) = _gpu(gd, shape, gd.F64, None, cb, db, src, T, dft, variant=gd.PLAIN,
                               f64=True)
    err = {}
    for name, st, se, mat, var in (("f32", gd.F32, None, "indexed", gd.TILED),
                                   ("f16", gd.F16, [-2] * 6, "indexed", gd.TILED),
                                   ("f16-plain", gd.F16, [-2] * 6, "indexed", gd.PLAIN),
                                   ("bf16", gd.BF16, [0] * 6, "indexed", gd.TILED)):
        _, (re, im) = _gpu(gd, shape, st, se, cb, db, src, T, dft, variant=var, mat=mat)
        err[name] = spectral_error(re, im, ref_re, ref_im)
        say(f"{name}: spectrum error per (component, frequency) {np.array2string(err[name], formatter={'float_kind': lambda x: f'{x:.2e}'})}")
    assert err["f32"].max() < 1e-5
    # rounding once per launch of 3 steps (TILED) instead of every half step
    # (PLAIN) lowers the error
    assert err["f16"].max() < 1e-3
    assert err["f16-plain"].max() < 3e-3
    assert err["f16-plain"].max() > err["f16"].max() > err["f32"].max()
    # bfloat16 keeps 8 significand bits against binary16's 11: about 2^3 times
    # the per-rounding error, with the same rounding schedule
    assert 2 * err["f16"].max() < err["bf16"].max() < 16 * err["f16"].max()
