"""Batches of independent problems on the GPU (gd_batch_*).

Every instance of a batch must equal the oracle's sequential update of the
same problem bit for bit, whatever the run mode (persistent launches, the
same launches replayed from a hipGraph, one launch sequence per instance),
the tilings of the size classes (several launch groups run as parallel
branches), the split of a run into calls (graph replays continue the step
count and the field-copy parity from device memory) and remainder launches
(nsteps not a multiple of blt).
"""
import numpy as np
import pytest

from cases import DT_C1, bits_equal, describe_diff
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

DIMS = [(40, 36, 30), (33, 33, 33), (40, 36, 30), (64, 20, 24), (20, 70, 18), (50, 41, 37),
        (33, 33, 33)]


def _problem(gd, dims, T, seed):
    nx, ny, nz = dims
    eps = gd.sphere_medium(nx, ny, nz, seed, 6, 2.0, 5.0, 1.0, 8.0)
    cb, db = gd.iso_cb_from_eps(eps, DT_C1, 1.0)
    f = 0.02 + 0.06 * ((seed * 37) % 11) / 10.0
    src = [(2, nx // 2, ny // 2, nz // 2, gd.gaussian_pulse_table(T, 8.0, 4.0, f) * 2.0)]
    if seed % 2:
        src.append((4, nx // 3, ny // 3 + 1, nz // 2 - 1, gd.gaussian_pulse_table(T, 6.0, 3.0, f)))
    return cb, db, src


def _make_batch(gd, dims, T, init=False):
    b = gd.Batch(dims, dx=1.0)
    probs = []
    for i, d in enumerate(dims):
        cb, db, src = _problem(gd, d, T, 1000 + i)
        b.set_isotropic(i, cb, db)
        for s in src:
            b.add_source(i, *s)
        fields = None
        if init:
            rng = np.random.default_rng(i)
            fields = [(rng.standard_normal((d[2], d[1], d[0])) * 0.05).astype(np.float32)
                      for _ in range(6)]
            for f in fields:
                f[0], f[-1], f[:, 0], f[:, -1], f[:, :, 0], f[:, :, -1] = 0, 0, 0, 0, 0, 0
            b.set_fields(i, fields)
        probs.append((cb, db, src, fields))
    return b, probs


def _oracle(oracle, dims, probs, T):
    out = []
    for d, (cb, db, src, init) in zip(dims, probs):
        f = [x.copy() for x in init] if init is not None else oracle.zero_fields(d)
        oracle.run(d, f, cb=cb, db=db, sources=src, nsteps=T)
        out.append(f)
    return out


def _check(b, want, label):
    for i, w in enumerate(want):
        got = b.get_fields(i)
        for c in range(6):
            assert bits_equal(got[c], w[c]), f"{label}: instance {i} {b.dims[i]} component {c}: " \
                                             f"{describe_diff(got[c], w[c])}"


@pytest.mark.parametrize("mode", ["persistent", "graph", "per_instance"])
def test_batch_matches_oracle(gd, oracle, mode):
    T = 13
    b, probs = _make_batch(gd, DIMS, T, init=True)
    try:
        m = {"persistent": gd.BATCH_PERSISTENT, "graph": gd.BATCH_GRAPH,
             "per_instance": gd.BATCH_PER_INSTANCE}[mode]
        say("batch", mode, b.classes())
        b.run(T, m)
        assert b.step == T
        _check(b, _oracle(oracle, DIMS, probs, T), mode)
    finally:
        b.close()


def test_batch_groups_replays_and_remainders(gd, oracle):
    """Classes on different (blt, rows, waves) form parallel launch groups;
    graph replays of the same nsteps continue from device state; runs of
    other lengths add remainder launches."""
    T = 7 + 7 + 7 + 5
    b, probs = _make_batch(gd, DIMS, T)
    try:
        tilings = [dict(blt=4, rows=4, waves=8, zchunk=16), dict(blt=3, rows=2, waves=8, zchunk=11),
                   dict(blt=2, rows=4, waves=4, zchunk=64), dict(blt=3, rows=2, waves=4, zchunk=9),
                   dict(blt=1, rows=2, waves=4, zchunk=7)]
        for c in range(len(b.classes())):
            b.set_class_tiling(c, tilings[c % len(tilings)])
        say("batch groups", b.classes())
        for n, mode in ((7, gd.BATCH_GRAPH), (7, gd.BATCH_GRAPH), (7, gd.BATCH_PERSISTENT),
                        (5, gd.BATCH_GRAPH)):
            b.run(n, mode)
        assert b.step == T
        _check(b, _oracle(oracle, DIMS, probs, T), "groups")
    finally:
        b.close()


def test_batch_autotune_keeps_state(gd, oracle):
    T = 12
    b, probs = _make_batch(gd, DIMS[:4], T, init=True)
    try:
        b.run(3, gd.BATCH_PERSISTENT)
        before = [b.get_fields(i) for i in range(len(b))]
        tuned = b.autotune(reps=1)
        say("autotuned", tuned)
        assert b.step == 3
        for i in range(len(b)):
            after = b.get_fields(i)
            for c in range(6):
                assert bits_equal(after[c], before[i][c])
        b.run(T - 3, gd.BATCH_GRAPH)
        _check(b, _oracle(oracle, DIMS[:4], probs, T), "after autotune")
    finally:
        b.close()


@pytest.mark.parametrize("continuous", [False, True])
def test_batch_material_forms(gd, oracle, continuous):
    """Launches run the 1-byte indexed material when every instance's Cb has
    at most 256 values (sphere media), the dense FP32 Cb as soon as one
    instance has more (a continuous medium); both == the oracle."""
    dims = DIMS[:3]
    T = 11
    b, probs = _make_batch(gd, dims, T, init=True)
    try:
        if continuous:
            d = dims[1]
            rng = np.random.default_rng(7)
            cb = (DT_C1 / rng.uniform(1.0, 6.0, (d[2], d[1], d[0]))).astype(np.float32)
            assert np.unique(cb).size > 256
            b.set_isotropic(1, cb, probs[1][1])
            probs[1] = (cb,) + probs[1][1:]
        say(f"[batch material] continuous={continuous}")
        b.run(T, gd.BATCH_GRAPH)
        _check(b, _oracle(oracle, dims, probs, T), f"material continuous={continuous}")
    finally:
        b.close()


def test_batch_errors(gd):
    b = gd.Batch([(20, 20, 20), (24, 20, 16)])
    try:
        with pytest.raises(gd.GdError):
            b.run(4, gd.BATCH_GRAPH)  # no material yet
        cb = np.full((20, 20, 20), 0.3, np.float32)
        b.set_isotropic(0, cb, 0.5)
        b.set_isotropic(1, np.full((16, 20, 24), 0.3, np.float32), 0.5)
        b.add_source(0, 2, 10, 10, 10, np.ones(5, np.float32))
        with pytest.raises(gd.GdError):
            b.run(6, gd.BATCH_GRAPH)  # table covers 5 steps
        with pytest.raises(gd.GdError):
            b.add_source(1, 2, 30, 1, 1, np.ones(5, np.float32))  # outside the grid
        with pytest.raises(gd.GdError):
            b.set_class_tiling(0, dict(blt=5, rows=4, waves=8, zchunk=8))
        b.run(5, gd.BATCH_GRAPH)
        assert b.step == 5
    finally:
        b.close()
    with pytest.raises(gd.GdError):
        gd.Batch([(2, 20, 20)])
