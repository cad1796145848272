"""Shared problem builders for the tests (inputs only; identical arrays go to
the HIP library and to the oracle)."""
import numpy as np

DT_C1 = np.float32(0.99 / np.sqrt(3.0))  # Courant 0.99 at dx = 1


def iso_problem(gd, shape, seed=7, nspheres=12, rmin=2.0, rmax=6.0, emin=1.0, emax=12.0,
                dt=DT_C1, dx=1.0):
    """Sphere medium -> (cb, db) in the library's isotropic form."""
    nx, ny, nz = shape
    eps = gd.sphere_medium(nx, ny, nz, seed, nspheres, rmin, rmax, emin, emax)
    cb, db = gd.iso_cb_from_eps(eps, dt, dx)
    return eps, cb, db


def centre(shape):
    nx, ny, nz = shape
    return nx // 2, ny // 2, nz // 2


def random_fields(shape, seed, scale=1.0, boundary_zero=False):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    out = [(rng.standard_normal((nz, ny, nx)) * scale).astype(np.float32) for _ in range(6)]
    if boundary_zero:
        for f in out:
            f[0], f[-1] = 0, 0
            f[:, 0], f[:, -1] = 0, 0
            f[:, :, 0], f[:, :, -1] = 0, 0
    return out


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def first_diff(a, b):
    d = np.argwhere(np.ascontiguousarray(a).view(np.uint32) != np.ascontiguousarray(b).view(np.uint32))
    if d.size == 0:
        return None
    k, j, i = d[0]
    return (int(i), int(j), int(k)), float(a[k, j, i]), float(b[k, j, i]), int(d.shape[0])


def rel_err(a, b):
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a.astype(np.float64) - b))) / den if den > 0 else float(
        np.max(np.abs(a)))


def describe_diff(got, want):
    """Short description of how two float arrays differ (for failure messages)."""
    g = np.ascontiguousarray(got).ravel()
    w = np.ascontiguousarray(want).ravel()
    ib = g.view(np.uint64 if g.dtype == np.float64 else np.uint32)
    wb = w.view(np.uint64 if w.dtype == np.float64 else np.uint32)
    bad = np.flatnonzero(ib != wb)
    if bad.size == 0:
        return "equal"
    vals_equal = np.flatnonzero(g[bad] != w[bad])
    q = bad[vals_equal[0]] if vals_equal.size else bad[0]
    return (f"{bad.size} bit-differences, {vals_equal.size} value-differences, "
            f"max |d| {np.abs(g - w).max():.3e}, first at flat {q}: got {g[q]!r} want {w[q]!r}")


def spectral_error(re, im, ref_re, ref_im):
    """Relative DFT-spectrum error ||X - X_ref||_2 / ||X_ref||_2 over the
    monitor plane, per (component, frequency); arrays shaped (ncomp, nf, plane)."""
    X = np.asarray(re, np.float64) + 1j * np.asarray(im, np.float64)
    R = np.asarray(ref_re, np.float64) + 1j * np.asarray(ref_im, np.float64)
    return np.linalg.norm(X - R, axis=-1) / np.linalg.norm(R, axis=-1)
