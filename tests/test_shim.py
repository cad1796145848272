"""The C++ drop-in include/gdiamond/gdiamond.hpp, driven as the reference's
own test and examples drive gdiamond::gDiamond (tests/cpp/shim_check.cpp)."""
import os
import subprocess

import numpy as np
import pytest

from cases import bits_equal
from conftest import ROOT, golden_cases, load_golden

EXE = os.path.join(ROOT, "tests", "cpp", "shim_check")


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/shim_check is not built (run __graft_entry__.build())")
    return EXE


@pytest.mark.parametrize("name", golden_cases())
def test_shim_seq_engine_matches_reference(exe, tmp_path, name):
    """gDiamond(nx,ny,nz).update_FDTD_seq(T) through the drop-in == the reference."""
    g = load_golden(name)
    nx, ny, nz = (int(v) for v in g["shape_xyz"])
    T = int(g["steps"])
    out = tmp_path / "seq.bin"
    r = subprocess.run([exe, "cpu", str(nx), str(ny), str(nz), str(T), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = np.fromfile(out, np.float32).reshape(6, nz, ny, nx)
    for c, nm in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        assert bits_equal(data[c], g[nm]), nm


@pytest.mark.gpu
def test_shim_gpu_engines(exe):
    """checker_gpu (seq == gpu_2D at 100^3 x 100), gpu_mm_ver2/ver3 (naive ==
    tiled) and gpu_figures (DFT monitor) through the drop-in."""
    print("[shim] gpu", flush=True)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    for tag in ("checker_gpu ok", "gpu_mm_ver2 ok", "gpu_mm_ver3 ok", "gpu_figures ok"):
        assert tag in r.stdout


@pytest.mark.gpu
def test_shim_gpu_figures_files(exe, tmp_path):
    """update_FDTD_gpu_figures(2100) writes the reference's figure set: after
    steps 100, 205, ... (t % (T/20) == 99) the Ex mid-plane and the real and
    imaginary DFT planes at frequency 11."""
    from png_read import read_png
    print("[shim] gpufig", flush=True)
    r = subprocess.run([exe, "gpufig"], capture_output=True, text=True, timeout=240,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    steps = [t + 1 for t in range(99, 2100, 105)]
    names = sorted(os.listdir(tmp_path / "gpu_figures"))
    want = sorted([f"{p}_{s:04d}.png" for s in steps
                   for p in ("Ex_naive_gpu", "fft_Ex_real", "fft_Ex_imag")])
    assert names == want
    img = read_png(tmp_path / "gpu_figures" / f"fft_Ex_real_{steps[-1]:04d}.png")
    assert img.shape == (50, 60, 3) and (img != 255).any()
