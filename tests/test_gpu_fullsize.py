"""Checks at the full sizes of the benchmark workloads (GPU).

* 1024^3 spheres with binary16 fields (the precision workload: seed 43, 200
  spheres, eps 1-12, scale 2^-8, material index): one TILED launch of the
  planner's 3 steps == the oracle's blocked binary16 twin (fields and the FP64
  DFT on z = 512) bit for bit, from initial fields that are non-zero in every
  cell; and, over 30 steps, the result does not depend on the tile shape
  (rows, waves, z chunk) at a fixed number of levels per launch.
* The 256-instance batch of the batch workload (sizes from
  {32,48,64,96,128,160}^3, default_rng(2024), 2.71e8 cells) with autotuned
  plans: the graph-launched run == one launch per instance, every instance bit
  for bit, and three instances of different sizes == the oracle.
"""
import numpy as np
import pytest

from cases import DT_C1, bits_equal, describe_diff
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

N3 = 1024
SCALE = [-8] * 6
DFT3 = dict(comps=[0, 2], axis=2, index=N3 // 2, f_cps=np.array([0.03, 0.05, 0.07]), scale=1.0)


@pytest.fixture(scope="module")
def c3_problem(gpu):
    say("[c3] sphere medium 1024^3")
    eps = gpu.sphere_medium(N3, N3, N3, 43, 200, 4.0, 20.0, 1.0, 12.0)
    cb, db = gpu.iso_cb_from_eps(eps, DT_C1, 1.0)
    del eps
    say("[c3] initial fields")
    init = [_pattern((N3, N3, N3), c) for c in range(6)]
    return cb, db, init


def _pattern(shape, c):
    """Deterministic initial field with values in [-0.5, 0.5] in every cell."""
    nx, ny, nz = shape
    i = np.arange(nx, dtype=np.int32)[None, None, :]
    j = np.arange(ny, dtype=np.int32)[None, :, None]
    k = np.arange(nz, dtype=np.int32)[:, None, None]
    v = (i * 7 + j * 13 + k * 29 + 31 * c) % 97
    return (v - 48).astype(np.float32) * np.float32(1.0 / 97.0)


def _c3_context(gpu, cb, db, init, T):
    ctx = gpu.Context(N3, N3, N3)
    ctx.set_storage(gpu.F16, SCALE)
    ctx.set_isotropic(cb, db)
    assert ctx.get_material_format()[0] == gpu.MAT_INDEXED
    c = N3 // 2
    ctx.add_source(gpu.EZ, c, c, c, gpu.gaussian_pulse_table(T, 30.0, 10.0, 0.05))
    ctx.set_fields(init)
    ctx.dft_config(DFT3["comps"], DFT3["axis"], DFT3["index"], DFT3["f_cps"], DFT3["scale"])
    return ctx


def test_config_c3_fp16_1024_one_launch_vs_oracle(gpu, oracle, c3_problem):
    cb, db, init = c3_problem
    shape = (N3, N3, N3)
    T = 3
    say("[c3] gpu: one TILED launch")
    ctx = _c3_context(gpu, cb, db, init, T)
    try:
        assert ctx.plan_tiling()["blt"] == T
        ctx.run(T, gpu.TILED)
        bits = [oracle.encode_f16(init[c], SCALE[c]) for c in range(6)]
        say("[c3] oracle: blocked binary16 twin, 16 threads")
        c = N3 // 2
        re, im = oracle.run_f16(shape, bits, SCALE, cb, db,
                                sources=[(gpu.EZ, c, c, c,
                                          gpu.gaussian_pulse_table(T, 30.0, 10.0, 0.05))],
                                nsteps=T, dft=dict(DFT3), block=T, threads=16)
        for q in range(6):
            got = ctx.get_field(q)
            want = oracle.decode_f16(bits[q], SCALE[q])
            assert bits_equal(got, want), f"component {q}: {describe_diff(got, want)}"
            del got, want
        for n, comp in enumerate(DFT3["comps"]):
            gre, gim = ctx.dft_get(comp)
            assert bits_equal(gre, re[n]), f"DFT re {comp}: {describe_diff(gre, re[n])}"
            assert bits_equal(gim, im[n]), f"DFT im {comp}: {describe_diff(gim, im[n])}"
    finally:
        ctx.close()


def test_config_c3_fp16_1024_tile_shape_independent(gpu, c3_problem):
    cb, db, init = c3_problem
    T = 30
    out = []
    for tiling in (None, dict(blt=3, rows=2, waves=8, zchunk=16)):
        say(f"[c3] gpu: 30 steps, tiling {tiling}")
        ctx = _c3_context(gpu, cb, db, init, T)
        try:
            ctx.run(T, gpu.TILED, tiling)
            if not out:
                out.append(([ctx.get_field(q) for q in range(6)],
                            [ctx.dft_get(comp) for comp in DFT3["comps"]]))
            else:
                fa, da = out[0]
                for q in range(6):
                    got = ctx.get_field(q)
                    assert bits_equal(got, fa[q]), f"component {q}: {describe_diff(got, fa[q])}"
                for n, comp in enumerate(DFT3["comps"]):
                    re, im = ctx.dft_get(comp)
                    assert bits_equal(re, da[n][0]) and bits_equal(im, da[n][1]), comp
                assert np.isfinite(fa[2]).all()
        finally:
            ctx.close()


def test_config_c4_batch_256_graph_equals_per_instance_and_oracle(gpu, oracle):
    T = 24
    rng = np.random.default_rng(2024)
    sizes = [int(s) for s in rng.choice([32, 48, 64, 96, 128, 160], size=256)]
    freqs = rng.uniform(0.02, 0.08, size=256)
    say("[c4] batch setup, 256 instances")
    b = gpu.Batch([(s, s, s) for s in sizes], dx=1.0)
    probs = {}
    check = {sizes.index(s) for s in (32, 96, 160)}
    try:
        for i, (s, f) in enumerate(zip(sizes, freqs)):
            eps = gpu.sphere_medium(s, s, s, 1000 + i, 200, 4.0, 20.0, 1.0, 12.0)
            b.set_isotropic_eps(i, eps, DT_C1)
            table = gpu.gaussian_pulse_table(T, 30.0, 10.0, f)
            b.add_source(i, gpu.EZ, s // 2, s // 2, s // 2, table)
            if i in check:
                cb, db = gpu.iso_cb_from_eps(eps, DT_C1, 1.0)
                probs[i] = (cb, db, [(gpu.EZ, s // 2, s // 2, s // 2, table)])
        say("[c4] autotune")
        b.autotune(reps=1)
        say("[c4] graph run")
        b.run(T, gpu.BATCH_GRAPH)
        first = [b.get_fields(i) for i in range(256)]
        b.reset()
        say("[c4] per-instance run")
        b.run(T, gpu.BATCH_PER_INSTANCE)
        for i in range(256):
            got = b.get_fields(i)
            for c in range(6):
                assert bits_equal(got[c], first[i][c]), \
                    f"instance {i} ({sizes[i]}^3) component {c}: {describe_diff(got[c], first[i][c])}"
        for i, (cb, db, src) in probs.items():
            say(f"[c4] oracle instance {i} ({sizes[i]}^3)")
            s = sizes[i]
            want = oracle.zero_fields((s, s, s))
            oracle.run((s, s, s), want, cb=cb, db=db, sources=src, nsteps=T, threads=16)
            for c in range(6):
                assert bits_equal(first[i][c], want[c]), \
                    f"instance {i} component {c}: {describe_diff(first[i][c], want[c])}"
    finally:
        b.close()
