"""Figure output: the reference's field colormap and PNG figures.

* gd_field_rgb == the pixels the reference's own save_field_png writes
  (golden fixture from its stb-based writer, tests/golden/make_golden.py png).
* gd_save_field_png / gd_save_mask_png files decode (any PNG reader; here
  tests/png_read.py with zlib) to those pixels; multi-block images too.
* The drop-in's update_FDTD_seq_figures writes Hz_seq_%04d.png every T/10
  steps, and the last one shows the final Hz mid-plane.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from png_read import read_png

GOLD = os.path.join(ROOT, "tests", "golden", "png_field_23x17.npz")
EXE = os.path.join(ROOT, "tests", "cpp", "shim_check")


def test_colormap_matches_reference(gd):
    g = np.load(GOLD)
    assert np.array_equal(gd.field_rgb(g["u"], float(g["vmax"])), g["rgb"])


def test_png_files_decode_to_reference_pixels(gd, tmp_path):
    g = np.load(GOLD)
    f = tmp_path / "u.png"
    gd.save_field_png(g["u"], f, float(g["vmax"]))
    assert np.array_equal(read_png(f), g["rgb"])
    m = tmp_path / "m.png"
    gd.save_mask_png(g["mask"], m)
    assert np.array_equal(read_png(m), g["mask_rgb"])


def test_png_large_image(gd, tmp_path):
    # > 65535 bytes of rows: several stored deflate blocks
    rng = np.random.default_rng(3)
    u = rng.standard_normal((150, 201)).astype(np.float32)
    f = tmp_path / "big.png"
    gd.save_field_png(u, f, 2.0)
    assert np.array_equal(read_png(f), gd.field_rgb(u, 2.0))


def test_shim_seq_figures(gd, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/shim_check is not built (run __graft_entry__.build())")
    nx, ny, nz, T = 24, 18, 24, 21
    out = tmp_path / "hz.bin"
    r = subprocess.run([EXE, "seqfig", str(nx), str(ny), str(nz), str(T), str(out)],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    figs = sorted(os.listdir(tmp_path / "seq_figures"))
    assert figs == [f"Hz_seq_{t:04d}.png" for t in range(0, T, T // 10)]
    hz = np.fromfile(out, np.float32).reshape(nz, ny, nx)
    last = read_png(tmp_path / "seq_figures" / figs[-1])  # written after step T-1
    assert np.array_equal(last, gd.field_rgb(hz[nz // 2], 1.0))
