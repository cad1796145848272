"""Mathematical pins of the oracle beyond the golden vectors.

* Discrete energy of the lossless Yee scheme is conserved exactly in exact
  arithmetic (the E and H curls are adjoint when boundary cells stay zero):
  W^n = sum eps E^n.E^n + sum mu H^(n-1/2).H^(n+1/2). A wrong sign or index in
  any curl term breaks it.
* Discrete divergence: forward-difference div H and backward-difference
  div(eps E) are constant in time away from sources and walls.
* DFT convention (kernels.cuh:87-113): re += v cos(2 pi f n) s, im -= v sin(..) s
  with v the value after step n, checked against a direct numpy sum over the
  recorded time series, and against the closed form: a cell driven so that its
  value after step n is cos(2 pi f0 n) accumulates s times a sum of two
  geometric series, evaluated in long double (a wrong sign, a step index off
  by one or a dropped scale fails it).
* FP32 and FP64 oracles agree to single-precision rounding.
* Discrete cavity eigenmodes: separable standing waves with the frozen
  boundary cells are exact eigenvectors of the update; their six amplitudes
  follow a closed-form 2-step recursion, met to 1e-12 in FP64 after 60 steps.
"""
import numpy as np
import pytest

from cases import centre, random_fields


def _interior_random(shape, seed):
    f = random_fields(shape, seed, boundary_zero=True)
    return [a.astype(np.float64) for a in f]


def energy(f, f_prev_h, eps, mu):
    E = sum(np.sum(eps * a * a) for a in f[:3])
    H = sum(np.sum(mu * a * b) for a, b in zip(f[3:], f_prev_h))
    return E + H


@pytest.mark.parametrize("shape", [(12, 10, 9), (16, 7, 11)])
def test_energy_conserved_lossless(oracle, shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(3)
    eps = rng.uniform(1.0, 4.0, (nz, ny, nx))
    dt, dx = 0.4, 1.0
    cb = (dt / (eps * dx)).astype(np.float32)
    db = np.float32(dt / dx)
    f = _interior_random(shape, 1)
    # cb is float32; use exactly those values as eps for the energy weights
    eps_eff = dt / (cb.astype(np.float64) * dx)
    mu_eff = dt / (np.float64(db) * dx)
    Ws = []
    for n in range(40):
        h_prev = [a.copy() for a in f[3:]]
        oracle.run(shape, f, cb=cb, db=db, dx=dx, nsteps=1, n0=n, dtype=np.float64)
        # E^(n+1) with H^(n+1/2) . H^(n+3/2): use E after the step and H before/after
        Ws.append(sum(np.sum(eps_eff * a * a) for a in f[:3]) +
                  sum(np.sum(mu_eff * a * b) for a, b in zip(h_prev, f[3:])))
    Ws = np.array(Ws)
    assert np.abs(Ws - Ws[0]).max() / abs(Ws[0]) < 1e-12


def test_energy_breaks_with_wrong_sign(oracle):
    """Plant a sign error in the H coefficient (db -> -db): energy is no longer
    conserved, so the check above can fail."""
    shape = (12, 10, 9)
    nx, ny, nz = shape
    cb = np.full((nz, ny, nx), 0.4, np.float32)
    f = _interior_random(shape, 1)
    Ws = []
    for n in range(20):
        h_prev = [a.copy() for a in f[3:]]
        oracle.run(shape, f, cb=cb, db=-0.4, dx=1.0, nsteps=1, n0=n, dtype=np.float64)
        Ws.append(sum(np.sum(a * a) for a in f[:3]) / 0.4 +
                  sum(np.sum(a * b) for a, b in zip(h_prev, f[3:])) / -0.4)
    Ws = np.array(Ws)
    assert np.abs(Ws - Ws[0]).max() / abs(Ws[0]) > 1e-3


def div_h(f):
    """Forward-difference divergence: annihilates the forward-difference curl
    that updates H (gdiamond.hpp:1209-1221)."""
    Hx, Hy, Hz = f[3], f[4], f[5]
    return ((Hx[:-1, :-1, 1:] - Hx[:-1, :-1, :-1]) + (Hy[:-1, 1:, :-1] - Hy[:-1, :-1, :-1]) +
            (Hz[1:, :-1, :-1] - Hz[:-1, :-1, :-1]))


def div_eps_e(f, eps):
    """Backward-difference divergence of eps*E (E is updated with backward
    differences of H, gdiamond.hpp:1194-1206); entry [k-1,j-1,i-1] is cell (i,j,k)."""
    Ex, Ey, Ez = (eps * f[0], eps * f[1], eps * f[2])
    return ((Ex[1:, 1:, 1:] - Ex[1:, 1:, :-1]) + (Ey[1:, 1:, 1:] - Ey[1:, :-1, 1:]) +
            (Ez[1:, 1:, 1:] - Ez[:-1, 1:, 1:]))


def test_divergence_constant_away_from_source(oracle):
    shape = (15, 13, 11)
    nx, ny, nz = shape
    dt = 0.5
    rng = np.random.default_rng(8)
    eps = rng.uniform(1.0, 3.0, (nz, ny, nx))
    cb = (dt / eps).astype(np.float32)
    eps_eff = dt / cb.astype(np.float64)
    f = _interior_random(shape, 2)
    dh0 = div_h(f)
    de0 = div_eps_e(f, eps_eff)
    T = 25
    i, j, k = centre(shape)
    src = np.sin(np.arange(T) * 0.3).astype(np.float32)
    oracle.run(shape, f, cb=cb, db=dt, dx=1.0, sources=[(2, i, j, k, src)], nsteps=T,
               dtype=np.float64)
    # only cells whose whole stencil is updated (frozen boundary cells are not
    # part of the curl identity): drop one layer on each side
    dh = (div_h(f) - dh0)[1:-1, 1:-1, 1:-1]
    assert np.abs(dh).max() < 1e-10
    de = (div_eps_e(f, eps_eff) - de0)[1:-1, 1:-1, 1:-1]  # entry [k-2,j-2,i-2] = cell (i,j,k)
    # the z-directed current charges exactly the two cells it connects
    assert np.abs(de[k - 2, j - 2, i - 2]) > 1e-3 and np.abs(de[k - 1, j - 2, i - 2]) > 1e-3
    de[k - 2:k, j - 2, i - 2] = 0.0
    assert np.abs(de).max() < 1e-10


def test_dft_convention_direct_sum(oracle):
    """Accumulated monitor == direct sum over the recorded plane values."""
    shape = (14, 12, 10)
    nx, ny, nz = shape
    cb = np.full((nz, ny, nx), 0.5, np.float32)
    T = 30
    src = (np.exp(-((np.arange(T) - 8) / 3.0) ** 2) * np.sin(0.6 * np.arange(T))).astype(
        np.float32)
    i, j, k = centre(shape)
    f_cps = np.array([0.05, 0.13])
    scale = 0.25
    f = oracle.zero_fields(shape)
    re, im = oracle.run(shape, f, cb=cb, db=0.5, dx=1.0, sources=[(2, i, j, k, src)], nsteps=T,
                        dft=dict(comps=[2, 3], axis=2, index=k, f_cps=f_cps, scale=scale))
    g = oracle.zero_fields(shape)
    series = []
    for n in range(T):
        oracle.run(shape, g, cb=cb, db=0.5, dx=1.0, sources=[(2, i, j, k, src)], n0=n, nsteps=1)
        series.append([g[2][k].astype(np.float64).ravel(), g[3][k].astype(np.float64).ravel()])
    series = np.array(series)  # (T, 2, plane)
    n = np.arange(T)
    for q, fq in enumerate(f_cps):
        w_c = np.cos(2 * np.pi * fq * n) * scale
        w_s = np.sin(2 * np.pi * fq * n) * scale
        for s in range(2):
            want_re = np.einsum("t,tp->p", w_c, series[:, s])
            want_im = -np.einsum("t,tp->p", w_s, series[:, s])
            tol = 1e-12 * (np.abs(series[:, s]).sum(0) + 1e-30)
            assert np.all(np.abs(re[s, q] - want_re) <= tol)
            assert np.all(np.abs(im[s, q] - want_im) <= tol)
    assert np.abs(re).max() > 0


def test_fp32_vs_fp64_oracle(oracle):
    shape = (24, 22, 20)
    nx, ny, nz = shape
    dt = np.float32(0.99 / np.sqrt(3.0))
    cb = np.full((nz, ny, nx), dt, np.float32)
    T = 60
    n = np.arange(T)
    src = (np.exp(-((n - 30) / 10.0) ** 2) * np.sin(2 * np.pi * 0.05 * n)).astype(np.float32)
    c = centre(shape)
    a = oracle.zero_fields(shape)
    b = oracle.zero_fields(shape, np.float64)
    oracle.run(shape, a, cb=cb, db=dt, dx=1.0, sources=[(2, *c, src)], nsteps=T)
    oracle.run(shape, b, cb=cb, db=dt, dx=1.0, sources=[(2, *c, src)], nsteps=T,
               dtype=np.float64)
    # Hz of a z-directed dipole is zero by symmetry: normalise by the E and H
    # scales rather than per component
    e_scale = max(np.abs(y).max() for y in b[:3])
    h_scale = max(np.abs(y).max() for y in b[3:])
    for c, (x, y) in enumerate(zip(a, b)):
        scale = e_scale if c < 3 else h_scale
        assert np.abs(x - y).max() / scale < 1e-5


def test_dft_weights_are_separate_libm_cos_sin(oracle):
    """Weights are libm cos(2 pi f n)*s and sin(2 pi f n)*s as separate calls
    (glibc's fused sincos() differs from them in the last bit for ~0.1% of
    arguments; the library evaluates them the same way)."""
    import math
    bad = 0
    for f in (0.02, 0.03, 0.05, 0.07, 0.11, 0.0446):
        for n in range(0, 3000, 7):
            ang = 2.0 * 3.14159265358979323846 * f * float(n)
            wc, ws = oracle.dft_weights(f, n, 1.0)
            bad += (wc != math.cos(ang)) + (ws != math.sin(ang))
    assert bad == 0


# ------------------------------------------------------------ DFT closed form


def dft_closed_form(f0, f, n0, n1, scale):
    """s * sum_{n=n0}^{n1-1} cos(2 pi f0 n) exp(-2 pi i f n) in long double:
    cos = (e^{i a n} + e^{-i a n}) / 2, each half a geometric series
    sum_{n=n0}^{n1-1} q^n = q^n0 (1 - q^(n1-n0)) / (1 - q)."""
    two_pi = np.longdouble(2) * np.longdouble(np.pi)
    a = two_pi * np.longdouble(f0)
    w = two_pi * np.longdouble(f)
    tot = np.clongdouble(0)
    for phi in (a - w, -a - w):
        q = np.exp(np.clongdouble(1j) * phi)
        if abs(1 - q) < 1e-15:
            g = np.clongdouble(n1 - n0)
        else:
            g = q ** n0 * (1 - q ** (n1 - n0)) / (1 - q)
        tot += g / 2
    return complex(tot * np.longdouble(scale))


def _driven_cell_setup(shape, cell, f0, T, dtype):
    """Full (12-array) material that makes the value of Ez at `cell` after
    step n equal -J(n) * dx with J(n) = -cos(2 pi f0 n): Caz = 0, Cbz = 1 at
    the cell, Cb = 0 elsewhere and Db = 0 everywhere (H stays 0, so the curl
    is 0 and E = 0 * E + 1 * (0 - J dx)); every other E keeps 0."""
    nx, ny, nz = shape
    i, j, k = cell
    co = [np.zeros((nz, ny, nx), np.float32) for _ in range(12)]
    for a in (0, 2, 4, 6, 8, 10):
        co[a][:] = 1.0  # Ca, Da
    co[4][k, j, i] = 0.0
    co[5][k, j, i] = 1.0
    n = np.arange(T)
    src = (-np.cos(2 * np.pi * f0 * n)).astype(np.float32)
    return co, src


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dft_closed_form(oracle, dtype):
    shape = (9, 8, 7)
    cell = (4, 3, 3)
    f0 = 0.05
    T = 397  # not a whole number of periods: every frequency has a sizeable sum
    scale = 0.5
    co, src = _driven_cell_setup(shape, cell, f0, T, dtype)
    f_cps = np.array([0.05, 0.031, 0.07, 0.0, 0.113])
    f = oracle.zero_fields(shape, dtype)
    # two calls: the weights follow the absolute step n0 + r
    dft = dict(comps=[2], axis=2, index=cell[2], f_cps=f_cps, scale=scale)
    re, im = oracle.run(shape, f, coeffs=co, dx=1.0, sources=[(2, *cell, src)], nsteps=150,
                        dft=dft, dtype=dtype)
    dft.update(re=re, im=im)
    oracle.run(shape, f, coeffs=co, dx=1.0, sources=[(2, *cell, src)], n0=150, nsteps=T - 150,
               dft=dft, dtype=dtype)
    p = cell[0] + cell[1] * shape[0]
    # the driven value is cos rounded to float32 (source tables are float32 in
    # both oracles: <= 2^-24 relative); the FP64 sums add T such terms
    tol = T * scale * 2.0 ** -23
    for q, fq in enumerate(f_cps):
        want = dft_closed_form(f0, fq, 0, T, scale)
        got = complex(re[0, q, p], im[0, q, p])
        assert abs(got - want) <= tol, (fq, got, want)
        # the checks are discriminating: one step late, the opposite sign of
        # the imaginary part or no scale each miss by far more than tol
        late = dft_closed_form(f0, fq, 1, T + 1, scale) * np.exp(2j * np.pi * fq)
        assert abs(got - late) > 100 * tol or abs(want) < 1.0
        assert abs(got - want.conjugate()) > 100 * tol or abs(want.imag) < 1.0
        assert abs(got - want / scale) > 100 * tol or abs(want) < 1.0
    # every other cell of the plane stays 0
    mask = np.ones(re.shape[-1], bool)
    mask[p] = False
    assert not re[0][:, mask].any() and not im[0][:, mask].any()


def test_dft_closed_form_resonance_magnitude():
    """Sanity of the closed form itself: at f = f0 the sum grows as T/2, away
    from resonance it stays bounded; brute force in long double agrees."""
    for f0, f, T in ((0.05, 0.05, 400), (0.05, 0.031, 400), (0.07, 0.02, 123)):
        n = np.arange(T, dtype=np.longdouble)
        two_pi = np.longdouble(2) * np.longdouble(np.pi)
        e = np.cos(two_pi * np.longdouble(f0) * n)
        brute = complex(np.sum(e * np.cos(two_pi * np.longdouble(f) * n)),
                        -np.sum(e * np.sin(two_pi * np.longdouble(f) * n)))
        assert abs(dft_closed_form(f0, f, 0, T, 1.0) - brute) < 1e-9
    assert abs(dft_closed_form(0.05, 0.05, 0, 400, 1.0) - 200.0) < 1.0


# --------------------------------------------------- discrete cavity eigenmodes
#
# With the boundary cells frozen, the reference update (gdiamond.hpp:1193-1221)
# has separable standing-wave eigenvectors. Along an axis of size N (L = N-1)
# a component sits at integer positions p (Ey, Ez, Hx along x; Ex, Ez, Hy
# along y; Ex, Ey, Hz along z) or half positions p+1/2 (the others); take
# f(p) = sin(k(L-p)) for the first kind and g(p) = cos(k(L-p-1/2)) for the
# second, with k = (m+1/2) pi / (L-1/2). Then f(L) = 0 and g(0) = 0, so every
# boundary cell an interior update reads is zero, the others are never read,
# and the differences map the profiles onto each other exactly:
# g(p) - g(p-1) = s f(p), f(p+1) - f(p) = -s g(p), s = 2 sin(k/2). With
# E = (ax gx fy fz, ay fx gy fz, az fx fy gz) and
# H = (bx fx gy gz, by gx fy gz, bz gx gy fz) one step is
#   a' = a + Cb (s x b),   b' = b + Db (s x a'),   s = (sx, sy, sz),
# a closed form for the six amplitudes. A wrong sign or offset in any curl
# term, a transposed axis or swapped Cb/Db breaks it.
def _profiles(N, m):
    L = N - 1
    k = (m + 0.5) * np.pi / (L - 0.5)
    p = np.arange(N, dtype=np.float64)
    f = np.sin(k * (L - p))
    g = np.cos(k * (L - p - 0.5))
    f[L] = 0.0   # exact zeros where the mode vanishes on a read boundary cell
    g[0] = 0.0
    return 2.0 * np.sin(k / 2.0), f, g


def _mode_fields(shape, prof, a, b):
    """The six fields (z, y, x) of amplitudes a, b on the per-axis profiles."""
    (sx, fx, gx), (sy, fy, gy), (sz, fz, gz) = prof
    def outer(px, py, pz):
        return pz[:, None, None] * py[None, :, None] * px[None, None, :]
    return [a[0] * outer(gx, fy, fz), a[1] * outer(fx, gy, fz), a[2] * outer(fx, fy, gz),
            b[0] * outer(fx, gy, gz), b[1] * outer(gx, fy, gz), b[2] * outer(gx, gy, fz)]


def _mode_amplitudes(a, b, cb, db, s, T):
    a = np.array(a, np.float64)
    b = np.array(b, np.float64)
    for _ in range(T):
        a = a + cb * np.cross(s, b)
        b = b + db * np.cross(s, a)
    return a, b


@pytest.mark.parametrize("modes", [(1, 2, 0), (2, 0, 3), (0, 1, 1)])
def test_cavity_eigenmode_closed_form(oracle, modes):
    shape = (12, 10, 14)                       # nx, ny, nz
    cb_v, db_v, T = 0.45, 0.3, 60              # stable: cb*db*|s|^2 < 4
    prof = [_profiles(shape[q], modes[q]) for q in range(3)]
    s = np.array([p[0] for p in prof])
    a0 = np.array([0.7, -0.4, 1.0])
    b0 = np.array([0.2, 0.5, -0.3])
    f = _mode_fields(shape, prof, a0, b0)
    init = [x.copy() for x in f]
    cb = np.full(shape[::-1], cb_v, np.float32)
    oracle.run(shape, f, cb=cb, db=np.float32(db_v), dx=1.0, nsteps=T, dtype=np.float64)
    aT, bT = _mode_amplitudes(a0, b0, np.float64(np.float32(cb_v)),
                              np.float64(np.float32(db_v)), s, T)
    want = _mode_fields(shape, prof, aT, bT)
    inner = (slice(1, -1),) * 3
    for c in range(6):
        assert np.abs(f[c][inner] - want[c][inner]).max() < 1e-12, c
        # boundary cells keep their initial values
        ring = np.ones(f[c].shape, bool)
        ring[inner] = False
        assert np.array_equal(f[c][ring], init[c][ring]), c
    # discrimination: the amplitudes move far from their start during the
    # run, and the recursion with the cross products transposed (a sign error
    # in every curl term) or Cb and Db swapped ends far from the oracle
    excursion = max(np.abs(_mode_amplitudes(a0, b0, cb_v, db_v, s, t)[0] - a0).max()
                    for t in range(1, T + 1))
    assert excursion > 0.1
    for cbx, dbx, sgn in ((cb_v, db_v, -1.0), (db_v, cb_v, 1.0)):
        aW, _ = _mode_amplitudes(a0, b0, cbx, sgn * dbx, sgn * s, T)
        assert np.abs(aW - aT).max() > 1e-3
    # a mistake planted in the oracle's input (Cb and Db swapped) fails the
    # closed form
    g = _mode_fields(shape, prof, a0, b0)
    oracle.run(shape, g, cb=np.full(shape[::-1], db_v, np.float32), db=np.float32(cb_v),
               dx=1.0, nsteps=T, dtype=np.float64)
    assert max(np.abs(g[c][inner] - want[c][inner]).max() for c in range(6)) > 1e-3
