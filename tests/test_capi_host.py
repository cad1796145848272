"""C ABI of libgdiamond_hip.so without a GPU: the library loads, exports every
symbol include/gdiamond/gd_capi.h declares, and its host-side problem builders
match the reference (golden vectors) and their stated formulas."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from cases import bits_equal
from conftest import ROOT, golden_cases, load_golden

HEADER = os.path.join(ROOT, "include", "gdiamond", "gd_capi.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gd_[a-z0-9_]+)\s*\(", txt)))


def test_header_declarations_are_exported(gd):
    L = C.CDLL(gd.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header(gd):
    assert set(gd.exported_symbols()) == set(declared_symbols())


def test_version(gd):
    assert "gfx950" in gd.version()


@pytest.mark.parametrize("name", golden_cases() + ["ref_coeff_40x33x27.npz"])
def test_library_default_coefficients_match_reference(gd, name):
    """gd_default_problem_coefficients == gDiamond(nx,ny,nz) coefficient arrays."""
    g = load_golden(name)
    nx, ny, nz = (int(v) for v in g["shape_xyz"])
    mine = gd.default_problem_coefficients(nx, ny, nz)
    for a, nm in enumerate(gd.COEFF_NAMES):
        assert bits_equal(mine[a], g[nm]), nm


@pytest.mark.parametrize("name", golden_cases())
def test_library_default_mz_matches_reference(gd, name):
    g = load_golden(name)
    assert bits_equal(gd.default_mz_table(int(g["steps"])), g["Mz_table"])


def test_gaussian_pulse_formula(gd):
    T = 300
    t = gd.gaussian_pulse_table(T, 30.0, 10.0, 0.05)
    want = np.array([np.float32(math.exp(-(((n - 30.0) / 10.0) ** 2)) *
                                math.sin(2 * math.pi * 0.05 * n)) for n in range(T)],
                    np.float32)
    assert np.abs(t - want).max() <= 1e-7
    assert 0.5 < float(np.abs(t).max()) <= 1.0


def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def test_sphere_medium_definition(gd):
    """SplitMix64 draws (cx, cy, cz, r, eps) per sphere; later spheres win."""
    nx, ny, nz = 40, 36, 30
    eps = gd.sphere_medium(nx, ny, nz, 42, 5, 3.0, 7.0, 1.0, 12.0)
    assert eps.shape == (nz, ny, nx)
    s = 42
    spheres = []
    for _ in range(5):
        vals = []
        for _ in range(5):
            s, z = splitmix64(s)
            vals.append((z >> 11) * 2.0 ** -53)
        cx, cy, cz = vals[0] * nx, vals[1] * ny, vals[2] * nz
        r = 3.0 + vals[3] * 4.0
        e = np.float32(1.0 + vals[4] * 11.0)
        spheres.append((cx, cy, cz, r, e))
    want = np.ones((nz, ny, nx), np.float32)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for cx, cy, cz, r, e in spheres:
        inside = (i - cx) ** 2 + (j - cy) ** 2 + (k - cz) ** 2 <= r * r
        want[inside] = e
    assert bits_equal(eps, want)
    assert (eps > 1).sum() > 100


def test_sphere_medium_reproducible(gd):
    a = gd.sphere_medium(64, 48, 40, 7)
    b = gd.sphere_medium(64, 48, 40, 7)
    c = gd.sphere_medium(64, 48, 40, 8)
    assert bits_equal(a, b) and not bits_equal(a, c)
    assert a.min() >= 1.0 and a.max() <= 12.0


def test_create_without_device_or_with_bad_sizes(gd):
    h = C.c_void_p()
    rc = gd.lib().gd_create(C.byref(h), 2, 10, 10, 1.0)
    assert rc == -1 and b"at least 3" in gd.lib().gd_last_error(None)
    # sizes whose cell count (2^60) or byte size would wrap are refused before
    # anything is allocated
    rc = gd.lib().gd_create(C.byref(h), 1 << 20, 1 << 20, 1 << 20, 1.0)
    assert rc == -1 and b"too large" in gd.lib().gd_last_error(None) and h.value is None
    dims = (C.c_size_t * 3)(1 << 14, 1 << 14, 1 << 14)
    rc = gd.lib().gd_batch_create(C.byref(h), 1, dims, 1.0)
    assert rc == -1 and b"2^36 cells" in gd.lib().gd_batch_last_error(None)
    if gd.device_count() == 0:
        rc = gd.lib().gd_create(C.byref(h), 10, 10, 10, 1.0)
        assert rc == -6 and h.value is None
        with pytest.raises(gd.GdError):
            gd.Context(8, 8, 8)


def test_planner_without_device(gd):
    """The planner's tilings: 4 rows x 8 waves; 4 levels (FP32, FP64) or 3
    (16-bit); z chunks sized for ~1024 workgroups within 16..32 planes (4- and
    8-byte fields) or 16..64 (16-bit)."""
    p = gd.plan_tiling_for
    assert p(512, 512, 512) == dict(blt=4, rows=4, waves=8, zchunk=32)
    assert p(512, 512, 512, gd.F16) == dict(blt=3, rows=4, waves=8, zchunk=64)
    assert p(512, 512, 512, gd.BF16)["blt"] == 3
    assert p(1024, 1024, 1024, gd.F16)["zchunk"] == 64
    assert p(1024, 1024, 1024)["zchunk"] == 32
    # 128^3: 3 x 6 tiles -> 128*18/1024 = 2 planes, raised to 16
    assert p(128, 128, 128) == dict(blt=4, rows=4, waves=8, zchunk=16)
    # 256 x 256 x 8: the chunk is the whole (short) z extent
    assert p(256, 256, 8)["zchunk"] == 8
    with pytest.raises(gd.GdError):
        p(2, 10, 10)
    with pytest.raises(gd.GdError):
        p(10, 10, 10, 7)


def test_host_builders_reject_bad_args(gd):
    L = gd.lib()
    assert L.gd_sphere_medium(4, 4, 4, 1, 1, 5.0, 2.0, 1.0, 2.0,
                              np.empty(64, np.float32).ctypes.data_as(gd._FP)) == -1
    assert L.gd_gaussian_pulse_table(4, 0.0, 0.0, 0.1,
                                     np.empty(4, np.float32).ctypes.data_as(gd._FP)) == -1
    null12 = (gd._FP * 12)()
    assert L.gd_default_problem_coefficients(4, 4, 4, null12) == -1


def test_library_has_no_oracle_dependency(gd):
    """The product library never links the oracle."""
    import subprocess
    out = subprocess.run(["ldd", gd.LIB_PATH], capture_output=True, text=True).stdout
    assert "liboracle" not in out and "gdref" not in out
    src = "".join(open(os.path.join(ROOT, "gdiamond_amd", "csrc", f)).read()
                  for f in os.listdir(os.path.join(ROOT, "gdiamond_amd", "csrc")))
    assert "oracle" not in src.lower().replace("the oracle", "")


def test_noalloc_launch_struct_layout(gd, tmp_path):
    """The ctypes mirror of gd_noalloc_launch / gd_point_source has the C
    struct's size and field offsets (compiled from the header with gcc)."""
    import subprocess
    fields = [f[0] for f in gd.NoallocLaunch._fields_]
    src = ["#include <stddef.h>", "#include <stdio.h>", '#include "gdiamond/gd_capi.h"',
           "int main(void) {",
           '  printf("%zu %zu\\n", sizeof(gd_noalloc_launch), sizeof(gd_point_source));']
    for f in fields:
        src.append(f'  printf("%zu\\n", offsetof(gd_noalloc_launch, {f}));')
    src += ["  return 0;", "}"]
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert int(out[0]) == C.sizeof(gd.NoallocLaunch)
    assert int(out[1]) == C.sizeof(gd.PointSource)
    for f, off in zip(fields, out[2:]):
        assert int(off) == getattr(gd.NoallocLaunch, f).offset, f


def test_noalloc_launch_rejects_bad_sizes_on_the_host(gd):
    """Argument checks run before any device work: out-of-range sizes give
    GD_ERR_ARG with a message (no GPU needed, nothing is launched)."""
    def launch(**kw):
        L = gd.NoallocLaunch()
        L.nx, L.ny, L.nz, L.dx, L.nsteps = 16, 16, 16, 1.0, 4
        for a in range(6):
            L.fields_in[a] = 0x1000 + 0x100 * a
            L.fields_out[a] = 0x2000 + 0x100 * a
        L.cb = 0x3000
        for k, v in kw.items():
            setattr(L, k, v)
        return gd.lib().gd_tiled_launch_noalloc(C.byref(L))
    assert launch(nsteps=5) == -1
    assert b"nsteps" in gd.lib().gd_last_error(None)
    assert launch(nsteps=0) == -1
    assert launch(nx=2) == -1
    assert launch(nsrc=9) == -1
    assert launch(cb=None) == -1
    assert launch(dft_mask=3, dft_axis=2, dft_index=4, nf=17) == -1
    assert launch(nx=2048, ny=2048, nz=3, dft_mask=3, dft_axis=2, dft_index=1, nf=1) == -1
    L = gd.NoallocLaunch()
    L.nx = L.ny = L.nz = 16
    L.nsteps = 1
    for a in range(6):
        L.fields_in[a] = 0x1000 + 0x100 * a
        L.fields_out[a] = 0x1000 + 0x100 * a  # in place
    L.cb = 0x3000
    assert gd.lib().gd_tiled_launch_noalloc(C.byref(L)) == -1
    assert b"must not be" in gd.lib().gd_last_error(None)
