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

Usage: python tests/golden/make_golden.py
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
COEFF_CASES = [(40, 33, 27)]


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    R = O.ref_lib()
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
