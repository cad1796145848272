"""Pin the oracle to the reference: golden vectors written by the reference's
own CPU path (tests/golden/make_golden.py), and the point-source / isotropic
shortcuts to the literal reference loop."""
import numpy as np
import pytest

from cases import bits_equal, centre, random_fields
from conftest import golden_cases, load_golden

COEFF_NAMES = ("Cax", "Cbx", "Cay", "Cby", "Caz", "Cbz",
               "Dax", "Dbx", "Day", "Dby", "Daz", "Dbz")
COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


@pytest.mark.parametrize("name", golden_cases() + ["ref_coeff_40x33x27.npz"])
def test_default_problem_coefficients_match_reference(oracle, name):
    """set_FDTD_matrices_3D_structure (utils.h:542-690) via gDiamond(nx,ny,nz)."""
    g = load_golden(name)
    shape = tuple(int(v) for v in g["shape_xyz"])
    mine = oracle.default_problem_coefficients(shape)
    for a, nm in enumerate(COEFF_NAMES):
        assert bits_equal(mine[a], g[nm]), nm


@pytest.mark.parametrize("name", golden_cases())
def test_default_source_matches_reference(oracle, name):
    g = load_golden(name)
    assert bits_equal(oracle.default_mz_table(int(g["steps"])), g["Mz_table"])


@pytest.mark.parametrize("name", golden_cases())
@pytest.mark.parametrize("threads", [1, 4])
def test_seq_update_matches_reference(oracle, name, threads):
    """update_FDTD_seq (gdiamond.hpp:1164-1263), bit for bit; threads=4 is the
    OpenMP twin of update_FDTD_omp (gdiamond_omp.hpp:110-172)."""
    g = load_golden(name)
    shape = tuple(int(v) for v in g["shape_xyz"])
    T = int(g["steps"])
    f = oracle.zero_fields(shape)
    i, j, k = oracle.default_source_ijk(shape)
    oracle.run(shape, f, coeffs=[g[n] for n in COEFF_NAMES], dx=np.float32(0.1),
               sources=[(5, i, j, k, g["Mz_table"])], nsteps=T, threads=threads)
    for a, nm in enumerate(COMPONENTS):
        assert bits_equal(f[a], g[nm]), nm


def test_golden_fields_are_nontrivial():
    """The fixtures exercise the absorber and the source (not all zeros)."""
    for name in golden_cases():
        g = load_golden(name)
        for nm in COMPONENTS:
            assert np.abs(g[nm]).max() > 0.1, (name, nm)
        assert np.unique(g["Cax"]).size > 3


def test_point_source_form_equals_jm_arrays(oracle):
    """Subtracting J*dx only at source cells == the reference's full J/M arrays."""
    shape = (17, 15, 13)
    rng = np.random.default_rng(1)
    coeffs = [(rng.uniform(0.5, 1.0, shape[::-1]) if a % 2 == 0 else rng.uniform(0.1, 0.4, shape[::-1])).astype(np.float32)
              for a in range(12)]
    T = 11
    srcs = [(c, 2 + c, 3 + c, 4 + (c % 3), rng.standard_normal(T).astype(np.float32))
            for c in range(6)]
    f0 = random_fields(shape, 3)
    a = [x.copy() for x in f0]
    b = [x.copy() for x in f0]
    oracle.run(shape, a, coeffs=coeffs, dx=np.float32(0.3), sources=srcs, nsteps=T)
    oracle.run_jm_arrays(shape, b, coeffs, np.float32(0.3), srcs, 0, T)
    for c in range(6):
        assert bits_equal(a[c], b[c])


def test_isotropic_form_equals_full_arrays(oracle):
    """Ca = Da = 1, Cbx = Cby = Cbz = cb, Db scalar: iso path == full path bitwise."""
    shape = (19, 16, 14)
    nx, ny, nz = shape
    rng = np.random.default_rng(2)
    cb = rng.uniform(0.05, 0.6, (nz, ny, nx)).astype(np.float32)
    db = np.float32(0.57)
    ones = np.ones_like(cb)
    dbs = np.full_like(cb, db)
    coeffs = [ones, cb, ones, cb, ones, cb, ones, dbs, ones, dbs, ones, dbs]
    f0 = random_fields(shape, 4)
    T = 9
    src = [(2, *centre(shape), rng.standard_normal(T).astype(np.float32))]
    a = [x.copy() for x in f0]
    b = [x.copy() for x in f0]
    oracle.run(shape, a, cb=cb, db=db, dx=1.0, sources=src, nsteps=T)
    oracle.run(shape, b, coeffs=coeffs, dx=1.0, sources=src, nsteps=T)
    for c in range(6):
        assert bits_equal(a[c], b[c])


def test_oracle_plants_detected(oracle):
    """A planted operand mistake in the oracle inputs (transposed source
    component) is caught by the golden comparison."""
    name = golden_cases()[0]
    g = load_golden(name)
    shape = tuple(int(v) for v in g["shape_xyz"])
    f = oracle.zero_fields(shape)
    i, j, k = oracle.default_source_ijk(shape)
    oracle.run(shape, f, coeffs=[g[n] for n in COEFF_NAMES], dx=np.float32(0.1),
               sources=[(4, i, j, k, g["Mz_table"])], nsteps=int(g["steps"]))
    assert not bits_equal(f[5], g["Hz"])


def test_reference_build_reports_its_seq_rate(oracle):
    """The reference's own update_FDTD_seq, built from its sources into
    oracle/_ref, runs and reports the Mcells/s it prints (its rate is quoted
    as context beside the CPU baseline); skipped where that build is
    absent."""
    import ctypes as C
    import os
    if not os.path.exists(oracle.REF_LIB_PATH):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    v = C.c_double(0.0)
    assert oracle.ref_lib().ref_seq_rate(20, 18, 16, 3, C.byref(v)) == 0
    assert v.value > 0.0
