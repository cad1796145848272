"""Seeded random cases of the TILED path against the oracle (GPU), bit for bit.

Each case draws, from its own seed: a grid (3..130 x 3..90 x 3..50 cells), a
tiling (levels per launch, rows x waves, z chunk; or the planner's choice),
field storage (FP32, or binary16 / bfloat16 with per-component scales) and material
(dense or indexed isotropic Cb, or the reference's 12 coefficient arrays for
FP32), up to three point sources, zero or random initial fields, a DFT plane
(any axis, a subset of components, 1-3 frequencies) and a run split into up to
three calls. The fields and the FP64 DFT sums must equal the oracle's:
`oracle.run` for FP32, the blocked 16-bit twin `oracle.run_f16` (rounding
once per launch of `blt` steps) for the 16-bit formats.
"""
import numpy as np
import pytest

from cases import bits_equal, describe_diff
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

NCASES = 200
RW = [(2, 4), (2, 8), (4, 4), (4, 8)]


def _draw(gd, seed):
    rng = np.random.default_rng(5000 + seed)
    nx, ny, nz = int(rng.integers(3, 131)), int(rng.integers(3, 91)), int(rng.integers(3, 51))
    shape = (nx, ny, nz)
    r = rng.random()
    fmt = "bf16" if r < 0.12 else ("f16" if r < 0.3 else None)  # 16-bit storage
    f16 = fmt is not None
    mat = rng.choice(["dense", "indexed"] if f16 else ["dense", "indexed", "full"])
    if rng.random() < 0.2:
        tiling = None
    else:
        blt = int(rng.integers(1, 5))
        rows, waves = RW[int(rng.integers(0, 4))]
        while rows * waves - 2 * blt < 1:
            rows, waves = RW[int(rng.integers(0, 4))]
        tiling = dict(blt=blt, rows=rows, waves=waves, zchunk=int(rng.integers(1, nz + 4)))
    eps_vals = rng.uniform(1.0, 9.0, size=int(rng.integers(1, 20))).astype(np.float32)
    eps = eps_vals[rng.integers(0, eps_vals.size, size=(nz, ny, nx))]
    dt = np.float32(0.55)
    cb = (np.float64(dt) / eps.astype(np.float64)).astype(np.float32)
    db = np.float32(dt)
    coeffs = None
    if mat == "full":
        coeffs = []
        for a in range(12):
            base = 0.97 if a % 2 == 0 else 0.4
            coeffs.append((base + 0.03 * rng.random((nz, ny, nx))).astype(np.float32))
    sources, used = [], set()
    for _ in range(int(rng.integers(0, 4))):
        comp = int(rng.integers(0, 6))
        # an interior cell (every axis has at least 3 cells)
        cell = (int(rng.integers(1, nx - 1)), int(rng.integers(1, ny - 1)),
                int(rng.integers(1, nz - 1)))
        if (comp, cell) in used:
            continue
        used.add((comp, cell))
        sources.append((comp, *cell, (rng.standard_normal(64) * 2).astype(np.float32)))
    init = None
    if rng.random() < 0.5:
        init = [(rng.standard_normal((nz, ny, nx)) * 0.5).astype(np.float32) for _ in range(6)]
    axis = int(rng.integers(0, 3))
    index = int(rng.integers(0, shape[axis]))
    comps = sorted(set(int(c) for c in rng.integers(0, 6, size=int(rng.integers(1, 4)))))
    dft = dict(comps=comps, axis=axis, index=index,
               f_cps=rng.uniform(0.01, 0.2, size=int(rng.integers(1, 4))),
               scale=float(rng.uniform(0.1, 2.0)))
    total = int(rng.integers(1, 15))
    cuts = sorted(set(int(c) for c in rng.integers(1, total + 1, size=int(rng.integers(0, 3)))))
    chunks = [b - a for a, b in zip([0] + cuts, cuts + [total]) if b > a]
    scale = [int(v) for v in rng.integers(-6, 4, size=6)] if f16 else None
    return dict(shape=shape, f16=f16, fmt=fmt, mat=str(mat), tiling=tiling, cb=cb, db=db, coeffs=coeffs,
                sources=sources, init=init, dft=dft, chunks=chunks, scale=scale)


def _gpu_run(gd, case):
    ctx = gd.Context(*case["shape"])
    try:
        if case["f16"]:
            ctx.set_storage(gd.BF16 if case["fmt"] == "bf16" else gd.F16, case["scale"])
        if case["coeffs"] is not None:
            ctx.set_coefficients(case["coeffs"])
        else:
            ctx.set_material_format(gd.MAT_DENSE if case["mat"] == "dense" else gd.MAT_INDEXED)
            ctx.set_isotropic(case["cb"], case["db"])
        for s in case["sources"]:
            ctx.add_source(*s)
        if case["init"] is not None:
            ctx.set_fields(case["init"])
        d = case["dft"]
        ctx.dft_config(d["comps"], d["axis"], d["index"], d["f_cps"], d["scale"])
        blt = (case["tiling"] or ctx.plan_tiling())["blt"]
        for n in case["chunks"]:
            ctx.run(n, gd.TILED, case["tiling"])
        fields = ctx.get_fields()
        spec = {c: ctx.dft_get(c) for c in d["comps"]}
    finally:
        ctx.close()
    return fields, spec, blt


def _oracle_run(oracle, case, blt):
    shape = case["shape"]
    nx, ny, nz = shape
    d = dict(case["dft"])
    if not case["f16"]:
        f = [a.copy() for a in case["init"]] if case["init"] is not None else oracle.zero_fields(shape)
        kw = dict(coeffs=case["coeffs"]) if case["coeffs"] is not None else dict(cb=case["cb"],
                                                                                 db=case["db"])
        n0 = 0
        for n in case["chunks"]:
            d["re"], d["im"] = oracle.run(shape, f, sources=case["sources"], n0=n0, nsteps=n,
                                          dft=d, threads=8, **kw)
            n0 += n
        return f, (d["re"], d["im"])
    se, fmt = case["scale"], case["fmt"]
    enc, dec = ((oracle.encode_bf16, oracle.decode_bf16) if fmt == "bf16" else
                (oracle.encode_f16, oracle.decode_f16))
    init = case["init"] or [np.zeros((nz, ny, nx), np.float32)] * 6
    bits = [enc(init[c], se[c]) for c in range(6)]
    cb = case["cb"]  # exact FP32 Cb in both material forms
    n0 = 0
    for n in case["chunks"]:
        d["re"], d["im"] = oracle.run_f16(shape, bits, se, cb, case["db"], sources=case["sources"],
                                          n0=n0, nsteps=n, dft=d, block=blt, fmt=fmt)
        n0 += n
    return [dec(bits[c], se[c]) for c in range(6)], (d["re"], d["im"])


@pytest.mark.parametrize("seed", range(NCASES))
def test_random_case_matches_oracle(gd, oracle, seed):
    case = _draw(gd, seed)
    say(f"[random {seed}] shape {case['shape']} 16-bit {case['fmt']} mat {case['mat']} "
        f"tiling {case['tiling']} sources {len(case['sources'])} chunks {case['chunks']} "
        f"dft axis {case['dft']['axis']} index {case['dft']['index']}")
    fields, spec, blt = _gpu_run(gd, case)
    want, (re, im) = _oracle_run(oracle, case, blt)
    for c in range(6):
        assert bits_equal(fields[c], want[c]), \
            f"seed {seed} comp {c}: {describe_diff(fields[c], want[c])}"
    for n, comp in enumerate(case["dft"]["comps"]):
        gre, gim = spec[comp]
        assert bits_equal(gre, re[n]) and bits_equal(gim, im[n]), \
            f"seed {seed} DFT comp {comp}: {describe_diff(gre, re[n])}"
