"""HIP path vs the oracle and the reference's golden vectors (GPU).

The reference's own known-answer relation is seq == GPU to |a-b| < 1e-8 on
every element (unittests/checker_gpu.cu:9-22, gdiamond.hpp:1330-1366). Both
kernels here use the reference operand order with one rounding per float
operation, so the bar is bit equality.
"""
import sys

import numpy as np
import pytest

from cases import (DT_C1, bits_equal, centre, describe_diff, first_diff, iso_problem,
                   random_fields, rel_err)
from conftest import golden_cases, load_golden

pytestmark = pytest.mark.gpu

TILINGS = [
    dict(blt=1, rows=2, waves=4, zchunk=7),
    dict(blt=2, rows=2, waves=8, zchunk=5),
    dict(blt=3, rows=4, waves=8, zchunk=64),
    dict(blt=4, rows=4, waves=4, zchunk=9),
    dict(blt=3, rows=2, waves=4, zchunk=3),
    dict(blt=4, rows=2, waves=8, zchunk=16),
    dict(blt=4, rows=4, waves=8, zchunk=11),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_fields_equal(got, want, label):
    for c in range(6):
        if not bits_equal(got[c], want[c]):
            fd = first_diff(got[c], want[c])
            raise AssertionError(f"{label}: component {c} differs: first (i,j,k)={fd[0]} "
                                 f"got={fd[1]} want={fd[2]} ndiff={fd[3]} "
                                 f"rel={rel_err(got[c], want[c]):.3e}")


def run_variant(gd, shape, setup, nsteps, variant, tiling=None, chunks=None):
    ctx = gd.Context(*shape, dx=setup.get("dx", 1.0))
    try:
        if "coeffs" in setup:
            ctx.set_coefficients(setup["coeffs"])
        else:
            ctx.set_isotropic(setup["cb"], setup["db"])
        for s in setup.get("sources", ()):
            ctx.add_source(*s)
        if "fields" in setup:
            ctx.set_fields(setup["fields"])
        if "dft" in setup:
            d = setup["dft"]
            ctx.dft_config(d["comps"], d["axis"], d["index"], d["f_cps"], d.get("scale", 1.0))
        for n in (chunks or [nsteps]):
            ctx.run(n, variant, tiling)
        out = ctx.get_fields()
        dft = None
        if "dft" in setup:
            dft = {c: ctx.dft_get(c) for c in setup["dft"]["comps"]}
        return out, dft
    finally:
        ctx.close()


def oracle_run(O, shape, setup, nsteps, threads=8):
    f = [a.copy() for a in setup["fields"]] if "fields" in setup else O.zero_fields(shape)
    kw = {}
    if "coeffs" in setup:
        kw["coeffs"] = setup["coeffs"]
    else:
        kw["cb"], kw["db"] = setup["cb"], setup["db"]
    dft = None
    if "dft" in setup:
        dft = dict(setup["dft"])
    res = O.run(shape, f, dx=np.float32(setup.get("dx", 1.0)), sources=setup.get("sources", ()),
                nsteps=nsteps, dft=dft, threads=threads, **kw)
    return f, res


# ------------------------------------------------------------------ golden vectors


@pytest.mark.parametrize("name", golden_cases())
@pytest.mark.parametrize("variant", ["plain", "tiled3", "tiled4"])
def test_reference_default_problem_golden(gpu, name, variant):
    """gDiamond(nx,ny,nz).update_FDTD_seq(T) of the reference, bit for bit."""
    g = load_golden(name)
    nx, ny, nz = (int(v) for v in g["shape_xyz"])
    T = int(g["steps"])
    coeffs = [g[n] for n in gpu.COEFF_NAMES]
    i, j, k = nx // 2, ny // 2, nz // 2
    setup = dict(coeffs=coeffs, dx=np.float32(0.1), sources=[(gpu.HZ, i, j, k, g["Mz_table"])])
    say(f"[golden] {name} {variant}")
    if variant == "plain":
        got, _ = run_variant(gpu, (nx, ny, nz), setup, T, gpu.PLAIN)
    else:
        t = (dict(blt=3, rows=2, waves=4, zchunk=5) if variant == "tiled3"
             else dict(blt=4, rows=4, waves=8, zchunk=6))
        got, _ = run_variant(gpu, (nx, ny, nz), setup, T, gpu.TILED, t)
    assert_fields_equal(got, [g[c] for c in gpu.COMPONENTS], f"{name}/{variant}")


# ------------------------------------------------------------------ oracle parity


@pytest.mark.parametrize("shape", [(37, 29, 23), (70, 45, 33), (64, 32, 20), (5, 4, 3)])
def test_iso_sphere_medium_plain_and_tiled(gpu, oracle, shape):
    """Isotropic sphere medium + Gaussian Jz point source, every tiling, bit-exact."""
    eps, cb, db = iso_problem(gpu, shape, seed=11, nspheres=6, rmin=2, rmax=5)
    T = 23
    src = gpu.gaussian_pulse_table(T, 8.0, 3.0, 0.07)
    i, j, k = centre(shape)
    setup = dict(cb=cb, db=db, sources=[(gpu.EZ, i, j, k, src), (gpu.HX, 1, 1, 1, src * 0.5)])
    want, _ = oracle_run(oracle, shape, setup, T)
    say(f"[iso] {shape} plain")
    got, _ = run_variant(gpu, shape, setup, T, gpu.PLAIN)
    assert_fields_equal(got, want, f"plain {shape}")
    for t in TILINGS:
        say(f"[iso] {shape} tiled {t}")
        got, _ = run_variant(gpu, shape, setup, T, gpu.TILED, t)
        assert_fields_equal(got, want, f"tiled {t} {shape}")


@pytest.mark.parametrize("shape", [(1300, 30, 9), (616, 50, 12)])
def test_wide_grids_column_tile_order(gpu, oracle, shape):
    """Rows of more than 10 tiles: FP32 launches walk the tiles in columns
    10 tiles wide (the last one narrower: 24 tiles = 10+10+4, 11 = 10+1 at
    B=4). The order changes which workgroups run together, never a value."""
    eps, cb, db = iso_problem(gpu, shape, seed=21, nspheres=8, rmin=2, rmax=6)
    T = 9
    src = gpu.gaussian_pulse_table(T, 4.0, 2.0, 0.07)
    i, j, k = centre(shape)
    setup = dict(cb=cb, db=db, fields=random_fields(shape, 31, scale=0.05),
                 sources=[(gpu.EZ, i, j, k, src), (gpu.HY, shape[0] - 40, 3, 2, src * 0.5)],
                 dft=dict(comps=[0, 2, 4], axis=0, index=shape[0] - 30,
                          f_cps=np.array([0.05, 0.09])))
    want, (wre, wim) = oracle_run(oracle, shape, setup, T)
    for t in [None] + TILINGS:
        say(f"[wide] {shape} tiling {t}")
        got, dft = run_variant(gpu, shape, setup, T, gpu.TILED, t)
        assert_fields_equal(got, want, f"wide {shape} {t}")
        for n, c in enumerate(setup["dft"]["comps"]):
            assert bits_equal(dft[c][0], wre[n]) and bits_equal(dft[c][1], wim[n]), \
                f"wide {shape} {t} DFT comp {c}"


@pytest.mark.parametrize("case", ["no-extras", "all-extras", "split", "split-remainder"])
def test_split_launch_forms(gpu, oracle, case):
    """The planner's 4 x 8 tile runs a launch as the lean kernel alone (no
    sources, no DFT), the full kernel alone (every tile holds the source or
    the monitored plane), or both on two streams (the usual case, also with a
    shorter last launch); every form == the oracle."""
    shape = {"no-extras": (150, 70, 40), "all-extras": (40, 20, 12),
             "split": (150, 70, 40), "split-remainder": (131, 61, 45)}[case]
    T = {"split-remainder": 10}.get(case, 8)
    eps, cb, db = iso_problem(gpu, shape, seed=5, nspheres=6, rmin=2, rmax=6)
    setup = dict(cb=cb, db=db, fields=random_fields(shape, 17, scale=0.05))
    if case != "no-extras":
        src = gpu.gaussian_pulse_table(T, 3.0, 2.0, 0.07)
        setup["sources"] = [(gpu.EZ, *centre(shape), src), (gpu.HX, 3, shape[1] - 4, 5, src)]
        setup["dft"] = dict(comps=[1, 5], axis=1, index=shape[1] // 3,
                            f_cps=np.array([0.04, 0.1]))
    want, res = oracle_run(oracle, shape, setup, T)
    say(f"[split] {case} {shape}")
    got, dft = run_variant(gpu, shape, setup, T, gpu.TILED, None)
    assert_fields_equal(got, want, f"split {case}")
    if res is not None:
        wre, wim = res
        for n, c in enumerate(setup["dft"]["comps"]):
            assert bits_equal(dft[c][0], wre[n]) and bits_equal(dft[c][1], wim[n]), \
                f"split {case} DFT comp {c}"


def test_split_sources_at_tile_window_edges(gpu, oracle):
    """Sources on and just outside the edges of the planner's tile windows (x
    windows [56t-4, 56t+60), y [24t-4, 24t+28), z chunks of 16 planes with a
    B-plane halo) and the monitored yz plane on a tile boundary: the host's
    full/lean classification must hand every source to a full-kernel tile.
    Nine steps: launches of 4, 4 and 1 levels (the last one classified anew)."""
    shape = (150, 90, 70)
    T = 9
    eps, cb, db = iso_problem(gpu, shape, seed=9, nspheres=10, rmin=3, rmax=8)
    cells = [(51, 19, 15), (52, 20, 16), (59, 27, 24), (60, 28, 23), (111, 43, 31),
             (112, 44, 32), (3, 3, 8), (140, 80, 60)]
    src = [((q % 6), *c, gpu.gaussian_pulse_table(T, 3.0, 2.0, 0.05 + 0.01 * q))
           for q, c in enumerate(cells)]
    setup = dict(cb=cb, db=db, sources=src, fields=random_fields(shape, 41, scale=0.05),
                 dft=dict(comps=[0, 4], axis=0, index=56, f_cps=np.array([0.06])))
    assert gpu.plan_tiling_for(*shape) == dict(blt=4, rows=4, waves=8, zchunk=16)
    want, (wre, wim) = oracle_run(oracle, shape, setup, T)
    say(f"[split edges] {shape}")
    got, dft = run_variant(gpu, shape, setup, T, gpu.TILED, None)
    assert_fields_equal(got, want, "split edges")
    for n, c in enumerate(setup["dft"]["comps"]):
        assert bits_equal(dft[c][0], wre[n]) and bits_equal(dft[c][1], wim[n]), \
            f"split edges DFT comp {c}"


def test_random_initial_fields_full_material(gpu, oracle):
    """Random fields (boundary included) + random full coefficient arrays."""
    shape = (41, 38, 27)
    nx, ny, nz = shape
    rng = np.random.default_rng(5)
    coeffs = []
    for a in range(12):
        if a % 2 == 0:
            coeffs.append(rng.uniform(0.9, 1.0, (nz, ny, nx)).astype(np.float32))
        else:
            coeffs.append(rng.uniform(0.05, 0.5, (nz, ny, nx)).astype(np.float32))
    fields = random_fields(shape, 9)
    setup = dict(coeffs=coeffs, fields=fields, dx=np.float32(0.7))
    T = 9
    want, _ = oracle_run(oracle, shape, setup, T)
    got, _ = run_variant(gpu, shape, setup, T, gpu.PLAIN)
    assert_fields_equal(got, want, "plain full random")
    for t in TILINGS[:4]:
        say(f"[full-random] tiled {t}")
        got, _ = run_variant(gpu, shape, setup, T, gpu.TILED, t)
        assert_fields_equal(got, want, f"tiled {t} full random")
    # boundary cells never change (PEC walls: only 1..n-2 are updated)
    for c in range(6):
        for sl in (np.s_[0], np.s_[-1], np.s_[:, 0], np.s_[:, -1], np.s_[:, :, 0], np.s_[:, :, -1]):
            assert bits_equal(got[c][sl], fields[c][sl])


def test_runs_continue_across_calls(gpu, oracle):
    """run(10); run(13) == run(23): source and DFT step indices are absolute."""
    shape = (33, 31, 29)
    eps, cb, db = iso_problem(gpu, shape, seed=3)
    T = 23
    src = gpu.gaussian_pulse_table(T, 6.0, 3.0, 0.05)
    setup = dict(cb=cb, db=db, sources=[(gpu.EZ, *centre(shape), src)],
                 dft=dict(comps=[0, 2, 4], axis=2, index=15, f_cps=np.array([0.03, 0.05, 0.07])))
    want, wdft = oracle_run(oracle, shape, setup, T)
    for variant, t in ((gpu.PLAIN, None), (gpu.TILED, dict(blt=4, rows=4, waves=4, zchunk=8)),
                       (gpu.TILED, dict(blt=3, rows=2, waves=8, zchunk=7))):
        got, gdft = run_variant(gpu, shape, setup, T, variant, t, chunks=[10, 13])
        assert_fields_equal(got, want, f"chunks variant {variant}")
        for s, c in enumerate([0, 2, 4]):
            assert bits_equal(gdft[c][0], wdft[0][s]), \
                f"variant {variant} comp {c} re: {describe_diff(gdft[c][0], wdft[0][s])}"
            assert bits_equal(gdft[c][1], wdft[1][s]), \
                f"variant {variant} comp {c} im: {describe_diff(gdft[c][1], wdft[1][s])}"


@pytest.mark.parametrize("axis,index", [(2, 11), (0, 7), (1, 20), (2, 0)])
def test_dft_monitor_fp64(gpu, oracle, axis, index):
    """FP64 running DFT (kernels.cuh:87-113 convention) on every plane orientation."""
    shape = (30, 26, 22)
    eps, cb, db = iso_problem(gpu, shape, seed=21)
    T = 31
    src = gpu.gaussian_pulse_table(T, 10.0, 4.0, 0.05)
    comps = [0, 1, 2, 3, 4, 5]
    setup = dict(cb=cb, db=db, sources=[(gpu.EZ, *centre(shape), src)],
                 dft=dict(comps=comps, axis=axis, index=index,
                          f_cps=np.array([0.02, 0.05, 0.11]), scale=0.5))
    want, (wre, wim) = oracle_run(oracle, shape, setup, T)
    for variant, t in ((gpu.PLAIN, None), (gpu.TILED, dict(blt=3, rows=2, waves=4, zchunk=6)),
                       (gpu.TILED, dict(blt=2, rows=4, waves=8, zchunk=64))):
        say(f"[dft] axis {axis} variant {variant} {t}")
        got, gdft = run_variant(gpu, shape, setup, T, variant, t)
        assert_fields_equal(got, want, "dft fields")
        for s, c in enumerate(comps):
            re, im = gdft[c]
            assert bits_equal(re, wre[s]), f"variant {variant} comp {c} re: {describe_diff(re, wre[s])}"
            assert bits_equal(im, wim[s]), f"variant {variant} comp {c} im: {describe_diff(im, wim[s])}"


def test_dft_closed_form_driven_cell(gpu, oracle):
    """A cell driven so that its Ez after step n is cos(2 pi f0 n) (Caz = 0,
    Cbz = 1 there, every other Cb and Db zero, J(n) = -cos): the device's FP64
    monitor sums equal the long-double geometric-series closed form to the
    float32 rounding of the driven values, for PLAIN and TILED, over a run
    split into three calls, and equal the oracle bit for bit."""
    from test_oracle_invariants import _driven_cell_setup, dft_closed_form
    shape = (21, 19, 17)
    cell = (7, 11, 9)
    f0, T, scale = 0.05, 397, 0.5
    co, src = _driven_cell_setup(shape, cell, f0, T, np.float32)
    f_cps = np.array([0.05, 0.031, 0.07, 0.0, 0.113])
    setup = dict(coeffs=co, sources=[(gpu.EZ, *cell, src)],
                 dft=dict(comps=[2], axis=2, index=cell[2], f_cps=f_cps, scale=scale))
    want, (wre, wim) = oracle_run(oracle, shape, setup, T)
    p = cell[0] + cell[1] * shape[0]
    tol = T * scale * 2.0 ** -23
    for variant, t in ((gpu.PLAIN, None), (gpu.TILED, dict(blt=3, rows=2, waves=4, zchunk=5)),
                       (gpu.TILED, dict(blt=4, rows=4, waves=8, zchunk=8))):
        say(f"[dft closed form] variant {variant} {t}")
        got, gdft = run_variant(gpu, shape, setup, T, variant, t, chunks=[150, 100, 147])
        assert_fields_equal(got, want, "driven cell fields")
        re, im = gdft[2]
        assert bits_equal(re, wre[0]) and bits_equal(im, wim[0])
        for q, fq in enumerate(f_cps):
            closed = dft_closed_form(f0, fq, 0, T, scale)
            assert abs(complex(re[q, p], im[q, p]) - closed) <= tol, (variant, fq)


# ------------------------------------------------------- benchmark workloads, full size


def test_config_vacuum_128_200_steps(gpu, oracle):
    """128^3 vacuum, 200 steps, Courant 0.99, Gaussian Jz at the centre, PEC walls:
    all six fields vs the sequential CPU update (bar: max rel error <= 1e-5;
    observed: bit equality)."""
    shape = (128, 128, 128)
    nx, ny, nz = shape
    cb = np.full((nz, ny, nx), np.float32(np.float64(DT_C1) / 1.0), np.float32)
    db = np.float32(np.float64(DT_C1) / 1.0)
    T = 200
    src = gpu.gaussian_pulse_table(T, 30.0, 10.0, 0.05)
    setup = dict(cb=cb, db=db, sources=[(gpu.EZ, *centre(shape), src)])
    want, _ = oracle_run(oracle, shape, setup, T, threads=16)
    for variant, t in ((gpu.PLAIN, None), (gpu.TILED, None)):
        say(f"[c1] variant {variant}")
        got, _ = run_variant(gpu, shape, setup, T, variant, t)
        for c in range(6):
            assert rel_err(got[c], want[c]) <= 1e-5
        assert_fields_equal(got, want, f"c1 variant {variant}")


@pytest.mark.timeout(900)
def test_config_spheres_512_full_run_vs_oracle(gpu, oracle):
    """The 512^3 sphere workload (200 spheres, r 4-20, eps 1-12, seed 42,
    Gaussian Jz at the centre, FP64 DFT of Ex and Ez on z = 256 at
    0.03/0.05/0.07 cycles/step) for all 1000 steps: TILED (the planner's
    tiling, the benchmark's material form), PLAIN and FUSED each equal the oracle's
    OpenMP update of the whole run on every element of the six fields and of
    the DFT sums, bit for bit -- the reference's whole-run relation
    (unittests/checker_gpu.cu:9-22, gdiamond.hpp:1330-1347)."""
    shape = (512, 512, 512)
    T = 1000
    eps, cb, db = iso_problem(gpu, shape, seed=42, nspheres=200, rmin=4, rmax=20, emin=1,
                              emax=12)
    del eps
    src = gpu.gaussian_pulse_table(T, 30.0, 10.0, 0.05)
    f_cps = np.array([0.03, 0.05, 0.07])
    setup = dict(cb=cb, db=db, sources=[(gpu.EZ, *centre(shape), src)],
                 dft=dict(comps=[0, 2], axis=2, index=256, f_cps=f_cps))
    got = {}
    for variant in (gpu.TILED, gpu.PLAIN, gpu.FUSED):
        say(f"[512 x {T}] variant {variant}")
        got[variant] = run_variant(gpu, shape, setup, T, variant)
    # the oracle in chunks of 100 steps (16 threads), continuing its DFT sums
    want = oracle.zero_fields(shape)
    d = dict(setup["dft"])
    for n0 in range(0, T, 100):
        res = oracle.run(shape, want, cb=cb, db=db, dx=1.0, sources=setup["sources"], n0=n0,
                         nsteps=100, dft=d, threads=16)
        d["re"], d["im"] = res
        say(f"[512 x {T}] oracle at step {n0 + 100}")
    for variant, (fields, dft) in got.items():
        assert_fields_equal(fields, want, f"512 x {T} variant {variant} vs oracle")
        for s, c in enumerate((0, 2)):
            re, im = dft[c]
            assert bits_equal(re, d["re"][s]), (variant, c, describe_diff(re, d["re"][s]))
            assert bits_equal(im, d["im"][s]), (variant, c, describe_diff(im, d["im"][s]))
    assert np.abs(d["re"]).max() > 0


@pytest.mark.parametrize("shape,blt", [((120, 80, 40), 4), ((120, 80, 40), 3), ((76, 53, 30), 4),
                                       ((200, 150, 37), 2), ((66, 34, 21), 1),
                                       ((57, 25, 9), 4)])
def test_remainder_tile_packs(gpu, oracle, shape, blt):
    """The lean form runs the tiles of a narrow remainder column (16 or 32
    lanes per tile, four or two per workgroup) and of a short remainder row
    (half-height tiles, two per workgroup) side by side; a partly filled
    workgroup pads with lanes that own nothing. Remainder tiles that hold a
    source or monitored cells stay whole tiles of the full form. 120 = 2*56+8
    cells at B=4 (16-lane tiles), 76 = 56+20 (32-lane tiles), 80 = 3*24+8 rows
    (half-height tiles); ten steps: launches of B levels and a shorter last
    one, each planned anew."""
    T = 10
    eps, cb, db = iso_problem(gpu, shape, seed=13, nspheres=8, rmin=2, rmax=6)
    nx, ny, nz = shape
    src = gpu.gaussian_pulse_table(T, 3.0, 2.0, 0.06)
    setup = dict(cb=cb, db=db, fields=random_fields(shape, 23, scale=0.05),
                 sources=[(gpu.EZ, *centre(shape), src), (gpu.HY, nx - 3, ny // 3, nz // 2, src),
                          (gpu.EX, nx // 3, ny - 2, 2, src * 0.5)],
                 dft=dict(comps=[2, 3], axis=0, index=nx - 2, f_cps=np.array([0.05, 0.08])))
    want, (wre, wim) = oracle_run(oracle, shape, setup, T)
    t = dict(blt=blt, rows=4, waves=8, zchunk=8)
    say(f"[packs] {shape} {t}")
    got, dft = run_variant(gpu, shape, setup, T, gpu.TILED, t)
    assert_fields_equal(got, want, f"packs {shape} {t}")
    for n, c in enumerate(setup["dft"]["comps"]):
        assert bits_equal(dft[c][0], wre[n]) and bits_equal(dft[c][1], wim[n]), \
            f"packs {shape} {t} DFT comp {c}"
    # without sources and DFT every remainder tile is packed
    setup2 = dict(cb=cb, db=db, fields=setup["fields"])
    want2, _ = oracle_run(oracle, shape, setup2, T)
    got2, _ = run_variant(gpu, shape, setup2, T, gpu.TILED, t)
    assert_fields_equal(got2, want2, f"packs (lean only) {shape} {t}")
