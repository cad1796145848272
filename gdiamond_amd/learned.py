"""Learned coarse-grid corrector: training data from the solver itself,
training with the library's network kernels, and the fine/coarse comparison.

The coarse grid has twice the cell size (dx = 2 in fine units) and, at the same
Courant number, twice the time step; with eps identical per physical point the
isotropic coefficients are numerically the same (cb = dt/(eps*dx)), so the
coarse run is the fine update on an 8x smaller grid, a quarter of the steps.

Correspondence (fine cell size 1, fine step dt):
  * coarse E after coarse step m  <->  fine E after fine step 2m, restricted by
    the mean of the two fine edges along the component's axis;
  * coarse H after coarse step m (time (2m+1) dt)  <->  the mean of the fine H
    after fine steps 2m and 2m+1, each restricted by the mean of the four fine
    faces (gd_restrict_fields does the spatial part on the device);
  * a fine point current J at one cell  <->  a coarse one of J/8 (same dipole
    moment), sampled as the mean of the two fine steps of the coarse step;
  * the same spheres: centres and radii in fine cells halve on the coarse grid
    (gd_sphere_medium draws centres as u*n and radii in cells).

Network input (12 channels per coarse cell, see gd_learned.hip): E, H and
E - E_prev scaled by se, sh, sd, and three material channels; output (12): the
scaled corrections of E, H at the current step (applied) and at the previous
step (trained, not applied).
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import (EX, EY, EZ, NN_LAYER_PARAMS, NN_PARAMS, Context, Net, gaussian_pulse_table,
               sphere_medium)

DT_FINE = np.float32(0.99 / math.sqrt(3.0))
LEAK = 0.125


def coarse_source(jf):
    """Coarse current of a fine one (same dipole moment, mean of the fine pair)."""
    jf = np.asarray(jf, np.float32)
    m = len(jf) // 2
    return ((jf[0:2 * m:2] + jf[1:2 * m:2]) * np.float32(0.5 / 8.0)).astype(np.float32)


def features(F, e_prev, cb, cfg):
    """Host twin of nn_features_kernel: (nz, ny, nx, 16) float32."""
    se, sh, sd, _, _, dx_over_dt = [np.float32(v) for v in cfg]
    nz, ny, nx = F[0].shape
    x = np.zeros((nz, ny, nx, 16), np.float32)
    for q in range(3):
        x[..., q] = F[q] * se
        x[..., 3 + q] = F[3 + q] * sh
        x[..., 6 + q] = (F[q] - e_prev[q]) * sd
    inv_eps = cb.astype(np.float32) * dx_over_dt
    x[..., 9] = (np.float32(1.0) / inv_eps - np.float32(1.0)) * np.float32(0.2)
    x[..., 10] = inv_eps - np.float32(1.0)
    edge = np.zeros(cb.shape, bool)
    for ax in range(3):
        d = np.diff(cb, axis=ax) != 0
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        edge[tuple(lo)] |= d
        edge[tuple(hi)] |= d
    x[..., 11] = edge.astype(np.float32)
    return x


def init_params(seed, last_scale=0.1):
    rng = np.random.default_rng(seed)
    p = np.zeros(NN_PARAMS, np.float32)
    for l, (cin, cout) in enumerate(((12, 16), (16, 16), (16, 12))):
        w = np.zeros((27, 16, 16), np.float32)
        std = math.sqrt(2.0 / (27 * cin)) * (last_scale if l == 2 else 1.0)
        w[:, :cout, :cin] = rng.standard_normal((27, cout, cin)) * std
        p[l * NN_LAYER_PARAMS:l * NN_LAYER_PARAMS + 27 * 256] = w.ravel()
    return p


class Problem:
    """One seeded sphere problem on a fine grid (n^3) and its coarse twin."""

    def __init__(self, n, seed, nspheres, rmin, rmax, f_cps, nsteps_fine, emax=6.0):
        assert n % 2 == 0
        self.n, self.nc = n, n // 2
        self.seed = seed
        self.eps_f = sphere_medium(n, n, n, seed, nspheres, rmin, rmax, 1.0, emax)
        self.eps_c = sphere_medium(self.nc, self.nc, self.nc, seed, nspheres, rmin / 2, rmax / 2,
                                   1.0, emax)
        self.jf = gaussian_pulse_table(nsteps_fine + 2, 30.0, 10.0, f_cps)
        self.jc = coarse_source(self.jf)
        self.dt_c = np.float32(2.0) * DT_FINE

    def fine_ctx(self):
        n = self.n
        ctx = Context(n, n, n, dx=1.0)
        ctx.set_isotropic_eps(self.eps_f, DT_FINE)
        ctx.add_source(EZ, n // 2, n // 2, n // 2, self.jf)
        return ctx

    def coarse_ctx(self, step0=0):
        nc = self.nc
        ctx = Context(nc, nc, nc, dx=2.0)
        ctx.set_isotropic_eps(self.eps_c, self.dt_c)
        ctx.add_source(EZ, nc // 2, nc // 2, nc // 2, self.jc[step0:])
        return ctx

    def cb_coarse(self):
        e = np.asarray(self.eps_c, np.float64)
        return (np.float64(self.dt_c) / (e * 2.0)).astype(np.float32)


def restricted_state(fine, m):
    """Coarse-grid image of the fine run at coarse step m; the fine context
    must be at fine step 2m (it is left at 2m + 1)."""
    assert fine.step == 2 * m, (fine.step, m)
    a = fine.restrict_fields()
    fine.run(1)
    b = fine.restrict_fields()
    return [a[0], a[1], a[2]] + [np.float32(0.5) * (a[q] + b[q]) for q in (3, 4, 5)]


def make_samples(npairs=64, n=96, nsnap=8, first=24, every=8, interval=4, seed0=5000,
                 log=None):
    """Training pairs generated by the solver: for each seeded problem, fine
    states restricted to the coarse grid at coarse steps t-interval, t-1 and t
    (t = first, first+every, ...), and the coarse solver run for `interval`
    steps from the restricted state at t-interval."""
    raw = []
    for k in range(npairs):
        seed = seed0 + k
        f = float(np.random.default_rng(seed).uniform(0.02, 0.08))
        last = first + every * (nsnap - 1)
        prob = Problem(n, seed, 10, 4.0, 16.0, f, 2 * last + 4)
        fine = prob.fine_ctx()
        want = {}
        for t in range(first, last + 1, every):
            for m in (t - interval, t - 1, t):
                want[m] = None
        for m in sorted(want):
            fine.run(2 * m - fine.step)
            want[m] = restricted_state(fine, m)
        fine.close()
        cb = prob.cb_coarse()
        for t in range(first, last + 1, every):
            ctx = prob.coarse_ctx(step0=t - interval)
            ctx.set_fields(want[t - interval])
            ctx.run(interval - 1)
            prev = ctx.get_fields()
            ctx.run(1)
            cur = ctx.get_fields()
            ctx.close()
            raw.append({"cur": cur, "prev": prev, "ref": want[t], "ref_prev": want[t - 1],
                        "cb": cb})
        if log:
            log(f"pair {k}: seed {seed} f {f:.3f} samples {len(raw)}")
    return raw


def apply_correction(net, cur, prev, cb, cfg):
    """The corrector's step on host arrays, as gd_run_corrected does on the
    device: features of (cur, prev E), network forward, F += y / r on interior
    cells. Returns the corrected fields."""
    net.set_input(features(cur, prev[:3], cb, cfg))
    y = net.forward()
    ie = np.float32(1.0) / np.float32(cfg[3])
    ih = np.float32(1.0) / np.float32(cfg[4])
    out = [f.copy() for f in cur]
    for q in range(6):
        s = ie if q < 3 else ih
        out[q][1:-1, 1:-1, 1:-1] = out[q][1:-1, 1:-1, 1:-1] + y[1:-1, 1:-1, 1:-1, q] * s
    return out


def make_rollout_samples(params, cfg, npairs=64, n=96, start=24, nint=8, interval=4,
                         seed0=5000, log=None):
    """Training pairs along the corrected solver's own trajectory: from the
    restricted fine state at coarse step `start`, the coarse solver runs nint
    intervals with the corrector `params` applied after each; every interval
    gives a sample (the coarse state before the correction and the restricted
    fine state it should have been), so the network sees the errors its own
    corrections leave behind."""
    raw = []
    for k in range(npairs):
        seed = seed0 + k
        f = float(np.random.default_rng(seed).uniform(0.02, 0.08))
        last = start + nint * interval
        prob = Problem(n, seed, 10, 4.0, 16.0, f, 2 * last + 4)
        fine = prob.fine_ctx()
        want = {}
        for m in [start] + [start + q * interval + d for q in range(1, nint + 1) for d in (-1, 0)]:
            want[m] = None
        for m in sorted(want):
            fine.run(2 * m - fine.step)
            want[m] = restricted_state(fine, m)
        fine.close()
        cb = prob.cb_coarse()
        nc = prob.nc
        ctx = prob.coarse_ctx(step0=start)
        net = Net(nc, nc, nc)
        net.set_params(params)
        ctx.set_fields(want[start])
        for q in range(1, nint + 1):
            t = start + q * interval
            ctx.run(interval - 1)
            prev = ctx.get_fields()
            ctx.run(1)
            cur = ctx.get_fields()
            raw.append({"cur": cur, "prev": prev, "ref": want[t], "ref_prev": want[t - 1],
                        "cb": cb})
            ctx.set_fields(apply_correction(net, cur, prev, cb, cfg))
        ctx.close()
        net.close()
        if log:
            log(f"rollout pair {k}: seed {seed} f {f:.3f} samples {len(raw)}")
    return raw


def fit_scales(raw, dx_over_dt):
    def rms(vals):
        s = sum(float(np.mean(np.square(v, dtype=np.float64))) for v in vals)
        return math.sqrt(s / len(vals)) or 1.0

    se = 1.0 / rms([r["cur"][q] for r in raw for q in range(3)])
    sh = 1.0 / rms([r["cur"][q] for r in raw for q in range(3, 6)])
    sd = 1.0 / rms([r["cur"][q] - r["prev"][q] for r in raw for q in range(3)])
    re = 1.0 / rms([r["ref"][q] - r["cur"][q] for r in raw for q in range(3)])
    rh = 1.0 / rms([r["ref"][q] - r["cur"][q] for r in raw for q in range(3, 6)])
    return (se, sh, sd, re, rh, float(dx_over_dt))


def sample_tensors(r, cfg):
    x = features(r["cur"], r["prev"][:3], r["cb"], cfg)
    nz, ny, nx = r["cb"].shape
    t = np.zeros((nz, ny, nx, 16), np.float32)
    re, rh = np.float32(cfg[3]), np.float32(cfg[4])
    for q in range(6):
        s = re if q < 3 else rh
        t[..., q] = (r["ref"][q] - r["cur"][q]) * s
        t[..., 6 + q] = (r["ref_prev"][q] - r["prev"][q]) * s
    mask = np.zeros((nz, ny, nx), np.uint8)
    mask[1:-1, 1:-1, 1:-1] = 1
    return x, t, mask


def train(raw, cfg, epochs=8, lr=1e-3, lr_min=1e-4, seed=0, log=None, params0=None):
    """Adam on all samples, cosine learning-rate decay, from params0 (or a
    seeded initialisation); returns the parameters and the mean loss of each
    epoch."""
    nz, ny, nx = raw[0]["cb"].shape
    rng = np.random.default_rng(seed)
    net = Net(nx, ny, nz)
    net.set_params(init_params(seed) if params0 is None else params0)
    tensors = [sample_tensors(r, cfg) for r in raw]
    steps = epochs * len(tensors)
    hist = []
    k = 0
    for ep in range(epochs):
        losses = []
        for i in rng.permutation(len(tensors)):
            x, t, mask = tensors[i]
            a = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * k / steps))
            net.set_input(x)
            net.set_target(t, mask)
            losses.append(net.train_step(a))
            k += 1
        hist.append(float(np.mean(losses)))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4e}")
    params = net.get_params()
    net.close()
    return params, hist


def baseline_loss(raw, cfg):
    """Loss of the zero correction (the coarse solver alone) on the samples."""
    tot = 0.0
    for r in raw:
        _, t, mask = sample_tensors(r, cfg)
        m = mask.astype(bool)
        tot += float(np.mean(np.square(t[m][:, :12], dtype=np.float64)))
    return tot / len(raw)


def restrict_spectrum(re, im, axis_comp):
    """Fine DFT plane (ny, nx) of Ex (axis_comp 0) or Ey (1) -> coarse plane,
    the mean of the two fine edges along the component's axis."""
    X = re + 1j * im
    if axis_comp == 0:
        return 0.5 * (X[..., 0::2, 0::2] + X[..., 0::2, 1::2])
    return 0.5 * (X[..., 0::2, 0::2] + X[..., 1::2, 0::2])


EPS_BINS = ((1.0, 1.0), (1.0, 3.0), (3.0, 6.0))  # vacuum, eps in (1,3], (3,6]


def _errors(spec, ref, f, nc, eps_plane):
    """Relative spectrum error per component and frequency over the plane, and
    per component, frequency and eps bin (cells of the coarse plane grouped
    by their eps)."""
    errs, bins = [], []
    masks = [(eps_plane == lo) if lo == hi else ((eps_plane > lo) & (eps_plane <= hi))
             for lo, hi in EPS_BINS]
    for X, R in zip(spec, ref):
        R = R.reshape(len(f), nc, nc)
        errs.append([float(np.linalg.norm(X[q] - R[q]) / np.linalg.norm(R[q]))
                     for q in range(len(f))])
        bins.append([[float(np.linalg.norm((X[q] - R[q])[m]) / np.linalg.norm(R[q][m]))
                      if m.any() else None for m in masks] for q in range(len(f))])
    return errs, bins


def _coarse_spectra(ctx, f, nc):
    # coarse sample m is fine time 2m+2, weighted as fine step 2m: shift
    shift = np.exp(-2j * np.pi * np.asarray(f))[:, None, None]
    spec = []
    for c in (EX, EY):
        re, im = ctx.dft_get(c)
        spec.append((re + 1j * im).reshape(len(f), nc, nc) * shift)
    return spec


def ideal_correction(n_fine, nsteps_fine, seed=6000, nspheres=200, rmin=4.0, rmax=20.0,
                     f_cps=(0.03, 0.05, 0.07), interval=4, log=None):
    """The best any corrector applied every `interval` coarse steps can do:
    the coarse run whose state is replaced, at every correction, by the
    restricted fine state itself (an exact correction), against the fine
    run's restricted spectrum. Errors as compare()."""
    prob = Problem(n_fine, seed, nspheres, rmin, rmax, 0.05, nsteps_fine)
    n, nc = prob.n, prob.nc
    f = np.asarray(f_cps, np.float64)
    fine = prob.fine_ctx()
    fine.dft_config([EX, EY], axis=2, index=n // 2, f_cps=f)
    ctx = prob.coarse_ctx()
    ctx.dft_config([EX, EY], axis=2, index=nc // 2, f_cps=2 * f, scale=2.0)
    t0 = time.perf_counter()
    m, M = 0, nsteps_fine // 2
    while m + interval <= M:
        ctx.run(interval)
        m += interval
        if m < M:  # a correction after the last step changes no DFT sample
            fine.run(2 * m - fine.step)
            ctx.set_fields(restricted_state(fine, m))
    if m < M:
        ctx.run(M - m)
    fine.run(nsteps_fine - fine.step)
    secs = time.perf_counter() - t0
    ref = []
    for q, c in enumerate((EX, EY)):
        re, im = fine.dft_get(c)
        ref.append(restrict_spectrum(re.reshape(len(f), n, n), im.reshape(len(f), n, n), q))
    fine.close()
    spec = _coarse_spectra(ctx, f, nc)
    ctx.close()
    errs, bins = _errors(spec, ref, f, nc, np.asarray(prob.eps_c)[nc // 2])
    out = {"ideal_err": errs, "ideal_err_eps_bins": bins, "seconds": secs}
    if log:
        log(str(out))
    return out


def compare(n_fine, nsteps_fine, params, cfg, seed=6000, nspheres=200, rmin=4.0, rmax=20.0,
            f_cps=(0.03, 0.05, 0.07), interval=4, log=None):
    """Fine run (the reference) against the coarse run without and with the
    corrector: relative DFT-spectrum error of Ex and Ey on the mid plane per
    component and frequency (and per eps bin of the coarse plane's cells),
    and wall time of each run."""
    prob = Problem(n_fine, seed, nspheres, rmin, rmax, 0.05, nsteps_fine)
    n, nc = prob.n, prob.nc
    f = np.asarray(f_cps, np.float64)
    out = {}

    fine = prob.fine_ctx()
    fine.dft_config([EX, EY], axis=2, index=n // 2, f_cps=f)
    t0 = time.perf_counter()
    fine.run(nsteps_fine)
    out["fine_ms"] = (time.perf_counter() - t0) * 1e3
    ref = []
    for q, c in enumerate((EX, EY)):
        re, im = fine.dft_get(c)
        ref.append(restrict_spectrum(re.reshape(len(f), n, n), im.reshape(len(f), n, n), q))
    fine.close()

    def coarse(with_net):
        ctx = prob.coarse_ctx()
        ctx.dft_config([EX, EY], axis=2, index=nc // 2, f_cps=2 * f, scale=2.0)
        net = None
        if with_net:
            net = Net(nc, nc, nc)
            net.set_params(params)
        t0 = time.perf_counter()
        if net is None:
            ctx.run(nsteps_fine // 2)
        else:
            ctx.run_corrected(net, nsteps_fine // 2, interval, cfg)
        ms = (time.perf_counter() - t0) * 1e3
        spec = _coarse_spectra(ctx, f, nc)
        ctx.close()
        if net is not None:
            net.close()
        errs, bins = _errors(spec, ref, f, nc, np.asarray(prob.eps_c)[nc // 2])
        return ms, errs, bins

    out["coarse_ms"], out["coarse_err"], out["coarse_err_eps_bins"] = coarse(False)
    if params is not None:
        out["learned_ms"], out["learned_err"], out["learned_err_eps_bins"] = coarse(True)
    if log:
        log(str(out))
    return out
