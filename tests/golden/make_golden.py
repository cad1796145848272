"""Write the golden vectors of the reference CPU path into tests/golden/.

Runs in the development container only (it needs /root/reference): it builds
the reference's own CPU path from its headers (oracle/Makefile target `ref`,
output in oracle/_ref/) and records, as data:

* ``ref_seq_<nx>x<ny>x<nz>_t<T>.npz`` -- the six fields after
  ``gDiamond(nx,ny,nz).update_FDTD_seq(T)`` (gdiamond.hpp:1164-1263), i.e. the
  default problem (graded absorber from utils.h:542-690, Mz sine source at the
  centre, gdiamond.hpp:1189-1191), plus the 12 coefficient arrays the
  constructor built and the per-step Mz values.
* ``ref_coeff_<nx>x<ny>x<nz>.npz`` -- coefficient arrays only, for a grid with a
  vacuum interior.

* ``png_field_23x17.npz`` -- a field and a mask with the pixels the reference's
  own figure writers (save_field_png / save_mask_png, utils.h:168-269) put in
  their PNG files.

Usage: python tests/golden/make_golden.py [png]   (png: the figure fixture only)
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

SEQ_CASES = [((20, 18, 22), 37), ((24, 24, 24), 30), ((13, 9, 11), 23)]
PNG_CASE = (23, 17)
COEFF_CASES = [(40, 33, 27)]


def png_fixture(R):
    """A field and a mask through the reference's own PNG writers; the fixture
    holds the inputs and the decoded pixels (png_field_23x17.npz)."""
    import tempfile
    sys.path.insert(0, os.path.dirname(HERE))
    from png_read import read_png
    nx, ny = PNG_CASE
    rng = np.random.default_rng(17)
    vmax = np.float32(0.37)
    u = (rng.standard_normal((ny, nx)) * vmax).astype(np.float32)
    u[0, :8] = vmax * np.array([0, -0.0, 1, -1, 1.5, -1.5, 0.5, -0.5], np.float32)
    mask = (rng.random((ny, nx)) < 0.3).astype(np.uint8)
    with tempfile.TemporaryDirectory() as d:
        fu, fm = os.path.join(d, "u.png"), os.path.join(d, "m.png")
        R.ref_save_field_png(u.ctypes.data_as(C.POINTER(C.c_float)), fu.encode(), nx, ny,
                             float(vmax))
        R.ref_save_mask_png(mask.ctypes.data_as(C.POINTER(C.c_uint8)), fm.encode(), nx, ny)
        rgb_u, rgb_m = read_png(fu), read_png(fm)
    path = os.path.join(HERE, f"png_field_{nx}x{ny}.npz")
    np.savez_compressed(path, u=u, vmax=vmax, rgb=rgb_u, mask=mask, mask_rgb=rgb_m)
    print("wrote", path, os.path.getsize(path))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    R = O.ref_lib()
    png_fixture(R)
    if "png" in sys.argv[1:]:
        return
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    for (nx, ny, nz), T in SEQ_CASES:
        N = nx * ny * nz
        fields = np.empty(6 * N, np.float32)
        R.ref_seq(nx, ny, nz, T, fp(fields))
        coeff = np.empty(12 * N, np.float32)
        R.ref_default_coefficients(nx, ny, nz, fp(coeff))
        mz = np.array([R.ref_default_mz(t) for t in range(T)], np.float32)
        out = {f"{O.COMPONENTS[a]}": fields[a * N:(a + 1) * N].reshape(nz, ny, nx)
               for a in range(6)}
        out.update({O.COEFF_NAMES[a]: coeff[a * N:(a + 1) * N].reshape(nz, ny, nx)
                    for a in range(12)})
        out["Mz_table"] = mz
        out["shape_xyz"] = np.array([nx, ny, nz], np.int64)
        out["steps"] = np.array(T, np.int64)
        path = os.path.join(HERE, f"ref_seq_{nx}x{ny}x{nz}_t{T}.npz")
        np.savez_compressed(path, **out)
        print("wrote", path, os.path.getsize(path))
    for nx, ny, nz in COEFF_CASES:
        N = nx * ny * nz
        coeff = np.empty(12 * N, np.float32)
        R.ref_default_coefficients(nx, ny, nz, fp(coeff))
        out = {O.COEFF_NAMES[a]: coeff[a * N:(a + 1) * N].reshape(nz, ny, nx)
               for a in range(12)}
        out["shape_xyz"] = np.array([nx, ny, nz], np.int64)
        path = os.path.join(HERE, f"ref_coeff_{nx}x{ny}x{nz}.npz")
        np.savez_compressed(path, **out)
        print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    main()
