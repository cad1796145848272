"""FUSED variant (one launch per step, E then H in one pass, two field copies)
vs the oracle and the reference's golden vectors (GPU), bit for bit.

The fused kernel's decomposition has its own edges: 62-cell x segments (lanes
1..62 of a wave), 3 rows per wave (each wave recomputing E on the row above
its rows), 8-plane z chunks (E recomputed on the plane after the chunk, H of
the grid's last plane copied after the loop). The shapes and source cells
below sit on and next to each of those edges; every result must equal the
sequential update, as PLAIN's does.
"""
import numpy as np
import pytest

from cases import DT_C1, bits_equal, centre, describe_diff, iso_problem, random_fields
from conftest import golden_cases, load_golden
from test_gpu_parity import assert_fields_equal, oracle_run, run_variant, say
from test_gpu_storage import SCALES, _assert_same, _f16_problem, _gpu, _oracle_f16

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_cases())
def test_fused_reference_default_problem_golden(gpu, name):
    """gDiamond(nx,ny,nz).update_FDTD_seq(T) of the reference (full coefficient
    arrays, Hz point source), bit for bit."""
    g = load_golden(name)
    nx, ny, nz = (int(v) for v in g["shape_xyz"])
    T = int(g["steps"])
    coeffs = [g[n] for n in gpu.COEFF_NAMES]
    i, j, k = nx // 2, ny // 2, nz // 2
    setup = dict(coeffs=coeffs, dx=np.float32(0.1), sources=[(gpu.HZ, i, j, k, g["Mz_table"])])
    say(f"[fused golden] {name}")
    got, _ = run_variant(gpu, (nx, ny, nz), setup, T, gpu.FUSED)
    assert_fields_equal(got, [g[c] for c in gpu.COMPONENTS], f"fused {name}")


# x: 62-cell segments; y: 3 rows per wave; z: 8-plane chunks
EDGE_SHAPES = [(5, 4, 3), (62, 16, 32), (63, 17, 33), (125, 33, 65), (124, 47, 64), (70, 45, 33),
               (3, 3, 3), (62, 8, 32), (63, 9, 31)]


@pytest.mark.parametrize("shape", EDGE_SHAPES)
def test_fused_iso_sphere_medium_edges(gpu, oracle, shape):
    """Sphere medium, random initial fields (boundary included), sources of
    every component on and next to the segment, row-group and chunk edges,
    FP64 DFT of all six components on a plane that crosses them."""
    nx, ny, nz = shape
    eps, cb, db = iso_problem(gpu, shape, seed=11, nspheres=6, rmin=2, rmax=5)
    T = 17
    cells = {(min(i, nx - 1), min(j, ny - 1), min(k, nz - 1))
             for i, j, k in [centre(shape), (61, 15, 31), (62, 16, 32), (1, 1, 1), (61, 7, 31),
                             (62, 8, 32), (nx - 2, ny - 2, nz - 2), (nx - 1, 0, nz // 2)]}
    src = [((q % 6), *c, gpu.gaussian_pulse_table(T, 5.0, 3.0, 0.05 + 0.01 * q))
           for q, c in enumerate(sorted(cells))]
    setup = dict(cb=cb, db=db, sources=src, fields=random_fields(shape, 3, scale=0.05),
                 dft=dict(comps=[0, 1, 2, 3, 4, 5], axis=2, index=min(32, nz - 1),
                          f_cps=np.array([0.03, 0.07]), scale=0.5))
    want, (wre, wim) = oracle_run(oracle, shape, setup, T)
    say(f"[fused edges] {shape}")
    got, dft = run_variant(gpu, shape, setup, T, gpu.FUSED)
    assert_fields_equal(got, want, f"fused {shape}")
    for s, c in enumerate(setup["dft"]["comps"]):
        assert bits_equal(dft[c][0], wre[s]), f"{shape} comp {c} re: {describe_diff(dft[c][0], wre[s])}"
        assert bits_equal(dft[c][1], wim[s]), f"{shape} comp {c} im: {describe_diff(dft[c][1], wim[s])}"


def test_fused_random_fields_full_material(gpu, oracle):
    """Random fields + random full coefficient arrays (the reference's 12
    arrays); boundary cells never change."""
    shape = (67, 38, 37)
    nx, ny, nz = shape
    rng = np.random.default_rng(5)
    coeffs = []
    for a in range(12):
        lo, hi = (0.9, 1.0) if a % 2 == 0 else (0.05, 0.5)
        coeffs.append(rng.uniform(lo, hi, (nz, ny, nx)).astype(np.float32))
    fields = random_fields(shape, 9)
    setup = dict(coeffs=coeffs, fields=fields, dx=np.float32(0.7))
    T = 9
    want, _ = oracle_run(oracle, shape, setup, T)
    got, _ = run_variant(gpu, shape, setup, T, gpu.FUSED)
    assert_fields_equal(got, want, "fused full random")
    for c in range(6):
        for sl in (np.s_[0], np.s_[-1], np.s_[:, 0], np.s_[:, -1], np.s_[:, :, 0], np.s_[:, :, -1]):
            assert bits_equal(got[c][sl], fields[c][sl])


@pytest.mark.parametrize("axis,index", [(2, 11), (0, 61), (1, 16), (2, 0), (2, 21)])
def test_fused_dft_monitor_fp64(gpu, oracle, axis, index):
    """FP64 running DFT on every plane orientation, the grid's last plane
    included (axis 2, index 21 of 22)."""
    shape = (70, 26, 22)
    eps, cb, db = iso_problem(gpu, shape, seed=21)
    T = 31
    src = gpu.gaussian_pulse_table(T, 10.0, 4.0, 0.05)
    comps = [0, 1, 2, 3, 4, 5]
    setup = dict(cb=cb, db=db, sources=[(gpu.EZ, *centre(shape), src)],
                 fields=random_fields(shape, 8, scale=0.05),
                 dft=dict(comps=comps, axis=axis, index=index,
                          f_cps=np.array([0.02, 0.05, 0.11]), scale=0.5))
    want, (wre, wim) = oracle_run(oracle, shape, setup, T)
    say(f"[fused dft] axis {axis} index {index}")
    got, gdft = run_variant(gpu, shape, setup, T, gpu.FUSED)
    assert_fields_equal(got, want, "fused dft fields")
    for s, c in enumerate(comps):
        re, im = gdft[c]
        assert bits_equal(re, wre[s]), f"comp {c} re: {describe_diff(re, wre[s])}"
        assert bits_equal(im, wim[s]), f"comp {c} im: {describe_diff(im, wim[s])}"


def test_fused_mixed_with_other_variants(gpu, oracle):
    """run(5, PLAIN); run(7, FUSED); run(4, TILED); run(6, FUSED) == the
    oracle's 22 steps: the variants hand the current field copy to each other
    and source and DFT step indices stay absolute."""
    shape = (71, 35, 40)
    eps, cb, db = iso_problem(gpu, shape, seed=3)
    T = 22
    src = gpu.gaussian_pulse_table(T, 6.0, 3.0, 0.05)
    setup = dict(cb=cb, db=db, sources=[(gpu.EZ, *centre(shape), src), (gpu.HY, 62, 16, 32, src)],
                 dft=dict(comps=[0, 2, 4], axis=2, index=15, f_cps=np.array([0.03, 0.05])))
    want, (wre, wim) = oracle_run(oracle, shape, setup, T)
    ctx = gpu.Context(*shape)
    try:
        ctx.set_isotropic(cb, db)
        for s in setup["sources"]:
            ctx.add_source(*s)
        d = setup["dft"]
        ctx.dft_config(d["comps"], d["axis"], d["index"], d["f_cps"])
        for n, v in ((5, gpu.PLAIN), (7, gpu.FUSED), (4, gpu.TILED), (6, gpu.FUSED)):
            ctx.run(n, v)
        got = ctx.get_fields()
        gdft = {c: ctx.dft_get(c) for c in d["comps"]}
    finally:
        ctx.close()
    assert_fields_equal(got, want, "mixed variants")
    for s, c in enumerate([0, 2, 4]):
        assert bits_equal(gdft[c][0], wre[s]) and bits_equal(gdft[c][1], wim[s]), c


@pytest.mark.parametrize("mat", ["dense", "indexed"])
@pytest.mark.parametrize("scale_exp", SCALES)
def test_fused_f16_matches_oracle(gd, oracle, scale_exp, mat):
    """binary16 fields, rounded every half step as PLAIN does == the
    oracle's binary16 twin, fields and DFT, over a run split in two calls."""
    shape = (70, 33, 35)
    T = 30
    cb, db, src, dft = _f16_problem(gd, shape, T)
    init = random_fields(shape, 5, scale=0.05, boundary_zero=True)
    want = _oracle_f16(oracle, shape, init, scale_exp, cb, db, src, T, dft, chunks=[13, 17])
    say("fused f16", shape, scale_exp, mat)
    got = _gpu(gd, shape, gd.F16, scale_exp, cb, db, src, T, dft, init=init, variant=gd.FUSED,
               mat=mat, chunks=[13, 17])
    _assert_same(got, want, f"fused f16 {mat}")


@pytest.mark.parametrize("mat", ["dense", "indexed"])
def test_fused_bf16_matches_oracle(gd, oracle, mat):
    shape = (63, 38, 33)
    T = 21
    scale_exp = SCALES[1]
    cb, db, src, dft = _f16_problem(gd, shape, T, seed=13)
    init = random_fields(shape, 4, scale=0.05, boundary_zero=True)
    want = _oracle_f16(oracle, shape, init, scale_exp, cb, db, src, T, dft, fmt="bf16")
    say("fused bf16", shape, mat)
    got = _gpu(gd, shape, gd.BF16, scale_exp, cb, db, src, T, dft, init=init, variant=gd.FUSED,
               mat=mat)
    _assert_same(got, want, f"fused bf16 {mat}")


def test_fused_f64_matches_oracle(gd, oracle):
    shape = (65, 31, 34)
    T = 35
    cb, db, src, dft = _f16_problem(gd, shape, T, seed=21)
    init = random_fields(shape, 2, scale=0.1, boundary_zero=True)
    wf = [a.astype(np.float64) for a in init]
    wre, wim = oracle.run(shape, wf, cb=cb, db=db, sources=src, nsteps=T, dft=dft,
                          dtype=np.float64)
    say("fused f64", shape)
    got = _gpu(gd, shape, gd.F64, None, cb, db, src, T, dft, init=init, variant=gd.FUSED,
               f64=True)
    _assert_same(got, (wf, (wre, wim)), "fused f64")


def test_fused_config_vacuum_128_200_steps(gpu, oracle):
    """128^3 vacuum, 200 steps, Courant 0.99, Gaussian Jz at the centre: all
    six fields equal the sequential CPU update."""
    shape = (128, 128, 128)
    nx, ny, nz = shape
    cb = np.full((nz, ny, nx), np.float32(np.float64(DT_C1) / 1.0), np.float32)
    db = np.float32(np.float64(DT_C1) / 1.0)
    T = 200
    src = gpu.gaussian_pulse_table(T, 30.0, 10.0, 0.05)
    setup = dict(cb=cb, db=db, sources=[(gpu.EZ, *centre(shape), src)])
    want, _ = oracle_run(oracle, shape, setup, T, threads=16)
    say("[fused c1]")
    got, _ = run_variant(gpu, shape, setup, T, gpu.FUSED)
    assert_fields_equal(got, want, "fused c1")
