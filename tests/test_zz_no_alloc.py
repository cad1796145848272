"""gd_tiled_launch_noalloc (one TILED launch on caller-owned device buffers,
scratch in a fixed device array, no allocation) against the context entry
point gd_run on the same problems, bit for bit: fields and FP64 DFT sums.

Two calls in a row, the second on a smaller grid (the scratch array keeps the
first call's contents), each with its own material form; then a sequence of
launches with the field buffers swapped and n0 advancing, against one gd_run
of the same steps."""
import ctypes as C

import numpy as np
import pytest

from cases import DT_C1, bits_equal, describe_diff, random_fields
from hipmem import DeviceArray, hip
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

F_CPS = np.array([0.03, 0.05, 0.07])


def _problem(gd, shape, seed):
    nx, ny, nz = shape
    eps = gd.sphere_medium(nx, ny, nz, seed, 8, 2.0, 7.0, 1.0, 12.0)
    cb, db = gd.iso_cb_from_eps(eps, DT_C1, 1.0)
    T = 40
    srcs = [(gd.EZ, nx // 2, ny // 2, nz // 2, gd.gaussian_pulse_table(T, 6.0, 3.0, 0.05)),
            (gd.HY, nx // 3, ny // 2 + 1, nz // 3, gd.gaussian_pulse_table(T, 5.0, 2.0, 0.07))]
    f0 = random_fields(shape, seed, 0.05)
    return cb, db, srcs, f0


def _context_run(gd, shape, prob, nsteps, dft):
    cb, db, srcs, f0 = prob
    with gd.Context(*shape, dx=1.0) as ctx:
        ctx.set_isotropic(cb, db)
        for s in srcs:
            ctx.add_source(*s)
        ctx.set_fields(f0)
        ctx.dft_config(dft["comps"], dft["axis"], dft["index"], F_CPS, dft["scale"])
        ctx.run(nsteps, gd.TILED)
        return ctx.get_fields(), {c: ctx.dft_get(c) for c in dft["comps"]}


def _noalloc_run(gd, shape, prob, chunks, dft, indexed):
    """Launches of `chunks` steps each through gd_tiled_launch_noalloc,
    swapping the two device field sets between launches."""
    cb, db, srcs, f0 = prob
    nx, ny, nz = shape
    bufs = [[DeviceArray(f) for f in f0], [DeviceArray(np.zeros_like(f)) for f in f0]]
    keep = []
    L = gd.NoallocLaunch()
    L.nx, L.ny, L.nz, L.dx, L.db = nx, ny, nz, 1.0, float(db)
    if indexed:
        vals, ids = np.unique(cb, return_inverse=True)
        table = np.zeros(256, np.float32)
        table[:vals.size] = vals
        keep += [DeviceArray(ids.astype(np.uint8).reshape(cb.shape)), DeviceArray(table)]
        L.material_id, L.material_table = keep[-2].ptr, keep[-1].ptr
    else:
        keep.append(DeviceArray(cb))
        L.cb = keep[-1].ptr
    ps = (gd.PointSource * len(srcs))()
    for q, (comp, i, j, k, table) in enumerate(srcs):
        keep.append(DeviceArray(table))
        ps[q].comp, ps[q].i, ps[q].j, ps[q].k, ps[q].table = comp, i, j, k, keep[-1].ptr
    L.nsrc, L.sources = len(srcs), ps
    mask = 0
    for c in dft["comps"]:
        mask |= 1 << c
    plane = {0: ny * nz, 1: nx * nz, 2: nx * ny}[dft["axis"]]
    re = DeviceArray(np.zeros((len(dft["comps"]), F_CPS.size, plane)))
    im = DeviceArray(np.zeros((len(dft["comps"]), F_CPS.size, plane)))
    f_cps = np.ascontiguousarray(F_CPS)
    L.dft_mask, L.dft_axis, L.dft_index, L.nf = mask, dft["axis"], dft["index"], F_CPS.size
    L.f_cycles_per_step = f_cps.ctypes.data_as(C.POINTER(C.c_double))
    L.dft_scale, L.dft_re, L.dft_im = dft["scale"], re.ptr, im.ptr
    L.stream = None
    cur, n0 = 0, 0
    for n in chunks:
        for a in range(6):
            L.fields_in[a] = bufs[cur][a].ptr
            L.fields_out[a] = bufs[1 - cur][a].ptr
        L.n0, L.nsteps = n0, n
        rc = gd.lib().gd_tiled_launch_noalloc(C.byref(L))
        assert rc == 0, gd.lib().gd_last_error(None)
        cur, n0 = 1 - cur, n0 + n
    assert hip().hipDeviceSynchronize() == 0
    fields = [b.download() for b in bufs[cur]]
    r, i_ = re.download(), im.download()
    return fields, {c: (r[s], i_[s]) for s, c in enumerate(dft["comps"])}


def _compare(got, want, label):
    gf, gd_ = got
    wf, wd = want
    for a in range(6):
        assert bits_equal(gf[a], wf[a]), f"{label} component {a}: {describe_diff(gf[a], wf[a])}"
    for c in wd:
        for q in range(2):
            assert bits_equal(gd_[c][q], wd[c][q]), f"{label} DFT {c}: " \
                f"{describe_diff(gd_[c][q], wd[c][q])}"


def test_noalloc_launch_matches_gd_run(gd):
    # two calls in a row, the second on a smaller grid
    cases = [((70, 50, 40), 4, False, dict(comps=[0, 2], axis=2, index=20, scale=0.5)),
             ((33, 21, 18), 3, True, dict(comps=[4], axis=0, index=16, scale=1.0))]
    for shape, n, indexed, dft in cases:
        say(f"[noalloc] {shape} nsteps {n} indexed {indexed}")
        prob = _problem(gd, shape, 31)
        got = _noalloc_run(gd, shape, prob, [n], dft, indexed)
        want = _context_run(gd, shape, prob, n, dft)
        _compare(got, want, f"noalloc {shape}")
        assert np.abs(got[1][dft["comps"][0]][0]).max() > 0


def test_noalloc_launch_sequence(gd):
    """Launches of 4, 4, 1 and 3 steps with n0 advancing == gd_run(12)."""
    shape = (61, 43, 37)
    dft = dict(comps=[1, 5], axis=1, index=21, scale=0.25)
    prob = _problem(gd, shape, 32)
    say("[noalloc] sequence")
    got = _noalloc_run(gd, shape, prob, [4, 4, 1, 3], dft, False)
    want = _context_run(gd, shape, prob, 12, dft)
    _compare(got, want, "noalloc sequence")
