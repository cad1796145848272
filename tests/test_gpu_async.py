"""gd_run_async / gd_sync: runs enqueued without waiting give the same results
as gd_run, bit for bit, whether they are completed by gd_sync or by the next
call that reads the context (fields, DFT sums, run statistics)."""
import numpy as np
import pytest

from cases import bits_equal, centre, iso_problem
from test_gpu_parity import assert_fields_equal, oracle_run, say

pytestmark = pytest.mark.gpu


def _setup(gpu):
    shape = (70, 45, 39)
    eps, cb, db = iso_problem(gpu, shape, seed=17)
    T = 24
    src = gpu.gaussian_pulse_table(T, 6.0, 3.0, 0.05)
    setup = dict(cb=cb, db=db, sources=[(gpu.EZ, *centre(shape), src)],
                 dft=dict(comps=[0, 2, 4], axis=2, index=19, f_cps=np.array([0.03, 0.07])))
    return shape, T, setup


def _ctx(gpu, shape, setup):
    ctx = gpu.Context(*shape)
    ctx.set_isotropic(setup["cb"], setup["db"])
    for s in setup["sources"]:
        ctx.add_source(*s)
    d = setup["dft"]
    ctx.dft_config(d["comps"], d["axis"], d["index"], d["f_cps"])
    ctx.set_launch_timing(True)
    return ctx


def test_async_chain_equals_oracle(gpu, oracle):
    """Four runs enqueued back to back (TILED, FUSED, PLAIN, TILED), then the
    fields and DFT sums read without an explicit sync == the oracle's 24
    steps."""
    shape, T, setup = _setup(gpu)
    want, (wre, wim) = oracle_run(oracle, shape, setup, T)
    ctx = _ctx(gpu, shape, setup)
    try:
        for n, v in ((7, gpu.TILED), (5, gpu.FUSED), (4, gpu.PLAIN), (8, gpu.TILED)):
            ctx.run(n, v, wait=False)
        say("[async] chain enqueued")
        got = ctx.get_fields()
        gdft = {c: ctx.dft_get(c) for c in setup["dft"]["comps"]}
        assert ctx.step == T
    finally:
        ctx.close()
    assert_fields_equal(got, want, "async chain")
    for s, c in enumerate(setup["dft"]["comps"]):
        assert bits_equal(gdft[c][0], wre[s]) and bits_equal(gdft[c][1], wim[s]), c


@pytest.mark.parametrize("variant", ["tiled", "fused", "plain"])
def test_async_then_sync_equals_sync_run(gpu, variant):
    """run(wait=False) + sync() == run(): fields, DFT sums; the statistics of
    the run are available after sync()."""
    v = {"tiled": gpu.TILED, "fused": gpu.FUSED, "plain": gpu.PLAIN}[variant]
    shape, T, setup = _setup(gpu)
    out = []
    for wait in (True, False):
        ctx = _ctx(gpu, shape, setup)
        try:
            ctx.run(T, v, wait=wait)
            if not wait:
                ctx.sync()
                ctx.sync()  # a second sync has nothing to wait for
            st = ctx.last_run_stats()
            assert st["total_ms"] > 0.0 and st["kernel_ms"] > 0.0 and st["nlaunch"] > 0
            out.append((ctx.get_fields(), ctx.dft_get(gpu.EZ)))
        finally:
            ctx.close()
    (f0, d0), (f1, d1) = out
    assert_fields_equal(f1, f0, f"async {variant}")
    assert bits_equal(d1[0], d0[0]) and bits_equal(d1[1], d0[1])


def test_async_stats_complete_the_run(gpu):
    """last_run_stats() right after run(wait=False) waits for the run; a
    setter after an async run sees the run's fields (reset zeroes them)."""
    shape, T, setup = _setup(gpu)
    ctx = _ctx(gpu, shape, setup)
    try:
        ctx.run(T // 2, gpu.TILED, wait=False)
        st = ctx.last_run_stats()
        assert st["total_ms"] > 0.0
        assert max(float(np.abs(a).max()) for a in ctx.get_fields()) > 0.0
        ctx.run(T // 2, gpu.FUSED, wait=False)
        ctx.reset()
        assert max(float(np.abs(a).max()) for a in ctx.get_fields()) == 0.0
    finally:
        ctx.close()


def test_async_argument_errors_return_at_once(gpu):
    shape, T, setup = _setup(gpu)
    ctx = _ctx(gpu, shape, setup)
    try:
        with pytest.raises(gpu.GdError):
            ctx.run(T, 7, wait=False)  # no such variant
        with pytest.raises(gpu.GdError):
            ctx.run(T + 100, gpu.TILED, wait=False)  # past the source table
        ctx.sync()  # nothing pending
    finally:
        ctx.close()
