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


def _corrected_vs_emulation(gd, learned, nc, eps_c, dt_c, jc, params, cfg, T, interval, warm):
    """gd_run_corrected over T steps after `warm` plain steps, and its host
    emulation on a second context: interval-1 steps, keep E, one more step,
    features on the host, the same network forward, F += y / r on interior
    cells. Returns both field sets and the last (features, network output)."""
    with gd.Context(nc, nc, nc, dx=2.0) as ctx, gd.Net(nc, nc, nc) as net:
        ctx.set_isotropic_eps(eps_c, dt_c)
        ctx.add_source(gd.EZ, nc // 2, nc // 2, nc // 2, jc)
        net.set_params(params)
        ctx.run(warm)  # something to correct
        ctx.run_corrected(net, T, interval, cfg)
        got = ctx.get_fields()
    cb = (np.float64(dt_c) / (np.asarray(eps_c, np.float64) * 2.0)).astype(np.float32)
    ie = np.float32(1.0) / np.float32(cfg[3])
    ih = np.float32(1.0) / np.float32(cfg[4])
    x = y = None
    with gd.Context(nc, nc, nc, dx=2.0) as ctx, gd.Net(nc, nc, nc) as net:
        ctx.set_isotropic_eps(eps_c, dt_c)
        ctx.add_source(gd.EZ, nc // 2, nc // 2, nc // 2, jc)
        net.set_params(params)
        ctx.run(warm)
        done = 0
        while done + interval <= T:
            ctx.run(interval - 1)
            prev = ctx.get_fields()
            ctx.run(1)
            cur = ctx.get_fields()
            x = learned.features(cur, prev[:3], cb, cfg)
            del prev
            net.set_input(x)
            y = net.forward()
            for q in range(6):
                s = ie if q < 3 else ih
                cur[q][1:-1, 1:-1, 1:-1] = cur[q][1:-1, 1:-1, 1:-1] + y[1:-1, 1:-1, 1:-1, q] * s
            ctx.set_fields(cur)
            del cur
            done += interval
        if done < T:
            ctx.run(T - done)
        want = ctx.get_fields()
    return got, want, x, y


@pytest.mark.parametrize("interval", [2, 4])
def test_run_corrected_matches_host_emulation(gd, interval):
    from gdiamond_amd import learned
    prob = _coarse_problem(gd, learned)
    nc = prob.nc
    cfg = (3.0, 2.5, 8.0, 40.0, 50.0, float(np.float32(2.0) / prob.dt_c))
    params = learned.init_params(5, last_scale=1.0)
    T = 3 * interval + 1
    got, want, _, y = _corrected_vs_emulation(gd, learned, nc, prob.eps_c, prob.dt_c, prob.jc,
                                              params, cfg, T, interval, 20)
    assert np.max(np.abs(y[..., :6])) > 0  # the network does correct
    for c in range(6):
        assert bits_equal(got[c], want[c]), (c, describe_diff(got[c], want[c]))


@pytest.mark.timeout(600)
def test_learned_form_at_its_size(gd):
    """The learned form's own configuration: the coarse 256^3 grid (dx = 2,
    dt = 2 dt_fine) of the 512^3 sphere medium with eps 1-6 (seed 6000, 200
    spheres, radius 4-20 fine cells), the fine Gaussian Jz mapped to the
    coarse grid. Two corrector intervals of 4 coarse steps plus one plain step
    through gd_run_corrected == the host emulation bit for bit; and a 32^3
    block of the network forward at 256^3 (the emulation's last application)
    == the FP32 torch reference with the device's BF16 roundings, to the
    tolerance of the network tests (2e-2 relative: BF16 activations)."""
    import nn_ref
    from gdiamond_amd import learned
    nc = 256
    eps_c = gd.sphere_medium(nc, nc, nc, 6000, 200, 2.0, 10.0, 1.0, 6.0)
    dt_c = np.float32(2.0) * learned.DT_FINE
    jc = learned.coarse_source(gd.gaussian_pulse_table(2 * 200 + 2, 30.0, 10.0, 0.05))
    cfg = (3.0, 2.5, 8.0, 40.0, 50.0, float(np.float32(2.0) / dt_c))
    params = learned.init_params(5, last_scale=1.0)
    say("[learned 256^3] corrected run and host emulation")
    got, want, x, y = _corrected_vs_emulation(gd, learned, nc, eps_c, dt_c, jc, params, cfg,
                                              T=9, interval=4, warm=40)
    assert np.max(np.abs(y[..., :6])) > 0
    for c in range(6):
        assert bits_equal(got[c], want[c]), (c, describe_diff(got[c], want[c]))
    del got, want
    # 32^3 output block around the source, with the 3-cell receptive-field halo
    # of three 3x3x3 layers taken from the same input
    lo, m, h = 112, 32, 3
    xs = np.ascontiguousarray(x[lo - h:lo + m + h, lo - h:lo + m + h, lo - h:lo + m + h])
    ref, _ = nn_ref.forward(params, xs)
    ref = ref.detach().numpy()[h:h + m, h:h + m, h:h + m, :12]
    blk = y[lo:lo + m, lo:lo + m, lo:lo + m, :12]
    err = np.max(np.abs(blk - ref)) / np.max(np.abs(ref))
    say(f"[learned 256^3] forward block rel err {err:.3e}")
    assert np.max(np.abs(ref)) > 0 and err < 2e-2, err


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


def test_rollout_samples_and_exact_correction(gd):
    """Rollout samples come from the corrected coarse trajectory (one per
    interval, coarse state before the correction against the restricted fine
    state); the exact correction (the restricted fine state every 4 coarse
    steps) runs and reports finite errors per component, frequency and eps
    bin. (It is not a bound: a restricted state is not a state of the coarse
    scheme, and in vacuum the resets can raise the error.)"""
    from gdiamond_amd import learned
    raw = learned.make_samples(npairs=1, n=32, nsnap=2, first=12, every=4, log=say)
    cfg = learned.fit_scales(raw, float(np.float32(2.0) / (np.float32(2.0) * learned.DT_FINE)))
    params = learned.init_params(3)
    roll = learned.make_rollout_samples(params, cfg, npairs=2, n=32, start=12, nint=3, log=say)
    assert len(roll) == 6
    for r in roll:
        assert r["cur"][0].shape == (16, 16, 16) and np.isfinite(r["cur"][2]).all()
        assert not np.array_equal(r["cur"][2], r["ref"][2])
    ideal = learned.ideal_correction(32, 160, nspheres=6, rmin=4.0, rmax=10.0, log=say)
    plain = learned.compare(32, 160, None, cfg, nspheres=6, rmin=4.0, rmax=10.0, log=say)
    assert np.isfinite(ideal["ideal_err"]).all() and np.isfinite(plain["coarse_err"]).all()
    assert np.asarray(ideal["ideal_err"]).shape == (2, 3)
    assert len(ideal["ideal_err_eps_bins"]) == 2 and len(ideal["ideal_err_eps_bins"][0][0]) == 3
