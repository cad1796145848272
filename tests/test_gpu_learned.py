"""Learned coarse-grid corrector in the solver loop (GPU).

* gd_restrict_fields == the Yee-consistent edge / face means in numpy.
* gd_run_corrected == its host emulation: run interval-1 steps, keep E, one
  more step, features on the host (learned.features, the twin of the device
  kernel), the same network forward (gd_nn_forward), F += y / r on interior
  cells. Both paths round the same FP32 features to BF16 and run the same
  kernels, so the corrected fields agree bit for bit.
* A network whose output is zero leaves the run equal to the uncorrected one.
* The training pipeline runs end to end on a small case and lowers the loss
  below the zero-correction baseline.
"""
import numpy as np
import pytest

from cases import bits_equal, describe_diff
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _restrict_np(F):
    ex, ey, ez, hx, hy, hz = [np.asarray(f, np.float32) for f in F]
    h = np.float32(0.5)
    q = np.float32(0.25)
    out = [h * (ex[0::2, 0::2, 0::2] + ex[0::2, 0::2, 1::2]),
           h * (ey[0::2, 0::2, 0::2] + ey[0::2, 1::2, 0::2]),
           h * (ez[0::2, 0::2, 0::2] + ez[1::2, 0::2, 0::2]),
           q * ((hx[0::2, 0::2, 0::2] + hx[0::2, 1::2, 0::2]) +
                (hx[1::2, 0::2, 0::2] + hx[1::2, 1::2, 0::2])),
           q * ((hy[0::2, 0::2, 0::2] + hy[0::2, 0::2, 1::2]) +
                (hy[1::2, 0::2, 0::2] + hy[1::2, 0::2, 1::2])),
           q * ((hz[0::2, 0::2, 0::2] + hz[0::2, 0::2, 1::2]) +
                (hz[0::2, 1::2, 0::2] + hz[0::2, 1::2, 1::2]))]
    return out


def test_restrict_fields(gd):
    shape = (20, 14, 12)
    nx, ny, nz = shape
    rng = np.random.default_rng(0)
    F = [rng.standard_normal((nz, ny, nx)).astype(np.float32) for _ in range(6)]
    with gd.Context(*shape) as ctx:
        ctx.set_isotropic(np.full((nz, ny, nx), 0.3, np.float32), 0.5)
        ctx.set_fields(F)
        got = ctx.restrict_fields()
    want = _restrict_np(F)
    for c in range(6):
        assert bits_equal(got[c], want[c]), (c, describe_diff(got[c], want[c]))


def _coarse_problem(gd, learned, n=32, seed=11):
    prob = learned.Problem(n * 2, seed, 6, 4.0, 10.0, 0.05, 200)
    return prob


@pytest.mark.parametrize("interval", [2, 4])
def test_run_corrected_matches_host_emulation(gd, interval):
    from gdiamond_amd import learned
    prob = _coarse_problem(gd, learned)
    nc = prob.nc
    cfg = (3.0, 2.5, 8.0, 40.0, 50.0, float(np.float32(2.0) / prob.dt_c))
    params = learned.init_params(5, last_scale=1.0)
    T = 3 * interval + 1
    # device path
    with gd.Context(nc, nc, nc, dx=2.0) as ctx, gd.Net(nc, nc, nc) as net:
        ctx.set_isotropic_eps(prob.eps_c, prob.dt_c)
        ctx.add_source(gd.EZ, nc // 2, nc // 2, nc // 2, prob.jc)
        net.set_params(params)
        ctx.run(20)  # something to correct
        ctx.run_corrected(net, T, interval, cfg)
        got = ctx.get_fields()
    # host emulation
    cb = prob.cb_coarse()
    ie = np.float32(1.0) / np.float32(cfg[3])
    ih = np.float32(1.0) / np.float32(cfg[4])
    with gd.Context(nc, nc, nc, dx=2.0) as ctx, gd.Net(nc, nc, nc) as net:
        ctx.set_isotropic_eps(prob.eps_c, prob.dt_c)
        ctx.add_source(gd.EZ, nc // 2, nc // 2, nc // 2, prob.jc)
        net.set_params(params)
        ctx.run(20)
        done = 0
        while done + interval <= T:
            ctx.run(interval - 1)
            prev = ctx.get_fields()
            ctx.run(1)
            cur = ctx.get_fields()
            net.set_input(learned.features(cur, prev[:3], cb, cfg))
            y = net.forward()
            for q in range(6):
                s = ie if q < 3 else ih
                cur[q][1:-1, 1:-1, 1:-1] = cur[q][1:-1, 1:-1, 1:-1] + y[1:-1, 1:-1, 1:-1, q] * s
            ctx.set_fields(cur)
            done += interval
        if done < T:
            ctx.run(T - done)
        want = ctx.get_fields()
    assert np.max(np.abs(y[..., :6])) > 0  # the network does correct
    for c in range(6):
        assert bits_equal(got[c], want[c]), (c, describe_diff(got[c], want[c]))


def test_zero_network_leaves_run_unchanged(gd):
    from gdiamond_amd import learned
    prob = _coarse_problem(gd, learned, n=24, seed=12)
    nc = prob.nc
    cfg = (1.0, 1.0, 1.0, 1.0, 1.0, float(np.float32(2.0) / prob.dt_c))
    out = []
    for corrected in (False, True):
        with gd.Context(nc, nc, nc, dx=2.0) as ctx, gd.Net(nc, nc, nc) as net:
            ctx.set_isotropic_eps(prob.eps_c, prob.dt_c)
            ctx.add_source(gd.EZ, nc // 2, nc // 2, nc // 2, prob.jc)
            net.set_params(np.zeros(gd.NN_PARAMS, np.float32))
            if corrected:
                ctx.run_corrected(net, 17, 4, cfg)
            else:
                ctx.run(17)
            out.append(ctx.get_fields())
    for c in range(6):
        assert np.array_equal(out[0][c], out[1][c]), c


def test_training_pipeline_small(gd):
    from gdiamond_amd import learned
    raw = learned.make_samples(npairs=3, n=32, nsnap=3, first=12, every=4, log=say)
    cfg = learned.fit_scales(raw, float(np.float32(2.0) / (np.float32(2.0) * learned.DT_FINE)))
    base = learned.baseline_loss(raw, cfg)
    params, hist = learned.train(raw, cfg, epochs=40, lr=2e-3, lr_min=2e-4, log=None)
    say("baseline", base, "loss", hist[0], hist[-1])
    assert np.isfinite(hist[-1]) and hist[-1] < hist[0] and hist[-1] < base
    res = learned.compare(32, 120, params, cfg, seed=6000, nspheres=6, rmin=4.0, rmax=10.0,
                          log=say)
    assert np.isfinite(res["learned_err"]).all() and np.isfinite(res["coarse_err"]).all()
