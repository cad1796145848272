"""Pins of the FP16-storage oracle.

* binary16 conversions == numpy's float16 casts (round to nearest even,
  subnormals, overflow) over a large sample and the edge cases.
* or_run_f16 == an independent numpy restatement of the storage semantics
  (decode to float32, float32 arithmetic in the reference operand order,
  astype(float16) once per step), bit for bit.
* the bfloat16 variant (fmt="bf16") == the same restatement with torch's
  bfloat16 rounding, and its conversions == torch's casts.
"""
import numpy as np
import pytest

from cases import bits_equal, centre


def test_f32_to_f16_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    parts = [
        np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65519.99, 65520.0, 70000.0, -1e9,
                  2.0 ** -24, 2.0 ** -25, 2.0 ** -25 * 1.0000001, 1.5 * 2.0 ** -24,
                  2.5 * 2.0 ** -24, 2.0 ** -14, 2.0 ** -14 * (1 - 2 ** -12), 6.1e-5,
                  np.inf, -np.inf], np.float32),
        (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 12, 20000))).astype(np.float32),
        # exact halfway points between binary16 neighbours
        ((np.arange(1, 4000, dtype=np.float32) + 0.5) * np.float32(2.0 ** -24)),
        ((np.arange(1024, 3000, dtype=np.float32) + 0.5) * np.float32(2.0 ** -10)),
    ]
    x = np.concatenate(parts)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    got = oracle.f32_to_f16_bits(x)
    assert np.array_equal(got, want)
    assert oracle.f32_to_f16_bits(np.array([np.nan], np.float32))[0] & 0x7c00 == 0x7c00


def test_f16_to_f32_matches_numpy(oracle):
    bits = np.arange(0, 65536, 7, dtype=np.uint16)
    bits = bits[(bits & 0x7c00) != 0x7c00]  # finite only
    want = bits.view(np.float16).astype(np.float32)
    got = oracle.decode_f16(bits)
    assert bits_equal(got, want)


# 16-bit codecs of the restatements: binary16 through numpy's float16 casts,
# bfloat16 through torch's (both round to nearest even, independently of the C)
F16_NP = (lambda v: v.astype(np.float16).view(np.uint16),
          lambda b: b.view(np.float16).astype(np.float32))


def _bf16_torch():
    torch = pytest.importorskip("torch")
    to_bits = lambda v: torch.from_numpy(np.ascontiguousarray(v, np.float32)).to(
        torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    from_bits = lambda b: torch.from_numpy(np.ascontiguousarray(b).view(np.int16)).view(
        torch.bfloat16).to(torch.float32).numpy()
    return to_bits, from_bits


def _numpy_f16_step(f, cb, db, sc, src, n, dx, codec=F16_NP):
    """One step of the 16-bit-storage semantics in numpy (independent of the C)."""
    to_bits, from_bits = codec
    dec = [from_bits(f[c]) * np.float32(sc[c]) for c in range(6)]
    Ex, Ey, Ez, Hx, Hy, Hz = dec
    b = np.asarray(cb, np.float32)
    I = np.s_[1:-1, 1:-1, 1:-1]
    km, jm, im = np.s_[:-2, 1:-1, 1:-1], np.s_[1:-1, :-2, 1:-1], np.s_[1:-1, 1:-1, :-2]
    J = [np.zeros(Ex[I].shape, np.float32) for _ in range(6)]
    for comp, i, j, k, table in src:
        J[comp][k - 1, j - 1, i - 1] = table[n]
    dxf = np.float32(dx)
    cx = ((Hz[I] - Hz[jm]) - (Hy[I] - Hy[km])) - J[0] * dxf
    cy = ((Hx[I] - Hx[km]) - (Hz[I] - Hz[im])) - J[1] * dxf
    cz = ((Hy[I] - Hy[im]) - (Hx[I] - Hx[jm])) - J[2] * dxf
    for c, cur in ((0, cx), (1, cy), (2, cz)):
        v = (dec[c][I] + b[I] * cur) * np.float32(1.0 / sc[c])
        f[c][I] = to_bits(v)
    Ex, Ey, Ez = [from_bits(f[c]) * np.float32(sc[c]) for c in range(3)]
    kp, jp, ip = np.s_[2:, 1:-1, 1:-1], np.s_[1:-1, 2:, 1:-1], np.s_[1:-1, 1:-1, 2:]
    dbf = np.float32(db)
    cx = ((Ey[kp] - Ey[I]) - (Ez[jp] - Ez[I])) - J[3] * dxf
    cy = ((Ez[ip] - Ez[I]) - (Ex[kp] - Ex[I])) - J[4] * dxf
    cz = ((Ex[jp] - Ex[I]) - (Ey[ip] - Ey[I])) - J[5] * dxf
    for c, cur in ((3, cx), (4, cy), (5, cz)):
        v = (dec[c][I] + dbf * cur) * np.float32(1.0 / sc[c])
        f[c][I] = to_bits(v)


@pytest.mark.parametrize("cb_f16", [True, False], ids=["cb-binary16", "cb-exact"])
@pytest.mark.parametrize("scale_exp", [[0] * 6, [-6, -5, -4, 3, 2, 1]])
def test_run_f16_matches_numpy_restatement(oracle, scale_exp, cb_f16):
    shape = (14, 12, 11)
    nx, ny, nz = shape
    rng = np.random.default_rng(4)
    eps = rng.uniform(1.0, 8.0, (nz, ny, nx))
    cbf = (0.5 / eps).astype(np.float32)
    cbv = cbf.astype(np.float16).astype(np.float32) if cb_f16 else cbf
    db = np.float32(0.5)
    T = 15
    src = [(2, *centre(shape), (np.sin(np.arange(T) * 0.4) * 3).astype(np.float32)),
           (4, 3, 4, 5, (np.cos(np.arange(T) * 0.3)).astype(np.float32))]
    sc = [2.0 ** e for e in scale_exp]
    init = [(rng.standard_normal((nz, ny, nx)) * 0.3).astype(np.float32) for _ in range(6)]
    a = [oracle.encode_f16(init[c], scale_exp[c]) for c in range(6)]
    b = [x.copy() for x in a]
    oracle.run_f16(shape, a, scale_exp, cbv, db, dx=1.0, sources=src, nsteps=T)
    for n in range(T):
        _numpy_f16_step(b, cbv, db, sc, src, n, 1.0)
    for c in range(6):
        assert np.array_equal(a[c], b[c]), c
    assert not np.array_equal(a[2], oracle.encode_f16(init[2], scale_exp[2]))


def _numpy_block_run(f16, cb, db, sc, src, n0, nsteps, block, dx=1.0, codec=F16_NP):
    """Block-rounding semantics in numpy: float steps from the decoded state,
    16-bit rounding of the whole state after every `block` steps."""
    to_bits, from_bits = codec
    v = [from_bits(f16[c]) * np.float32(sc[c]) for c in range(6)]
    I = np.s_[1:-1, 1:-1, 1:-1]
    km, jm, im = np.s_[:-2, 1:-1, 1:-1], np.s_[1:-1, :-2, 1:-1], np.s_[1:-1, 1:-1, :-2]
    kp, jp, ip = np.s_[2:, 1:-1, 1:-1], np.s_[1:-1, 2:, 1:-1], np.s_[1:-1, 1:-1, 2:]
    b = np.asarray(cb, np.float32)[I]
    dbf, dxf = np.float32(db), np.float32(dx)
    done = 0
    for n in range(n0, n0 + nsteps):
        J = [np.zeros(v[0][I].shape, np.float32) for _ in range(6)]
        for comp, i, j, k, table in src:
            J[comp][k - 1, j - 1, i - 1] = table[n]
        Ex, Ey, Ez, Hx, Hy, Hz = v
        cx = ((Hz[I] - Hz[jm]) - (Hy[I] - Hy[km])) - J[0] * dxf
        cy = ((Hx[I] - Hx[km]) - (Hz[I] - Hz[im])) - J[1] * dxf
        cz = ((Hy[I] - Hy[im]) - (Hx[I] - Hx[jm])) - J[2] * dxf
        Ex[I], Ey[I], Ez[I] = Ex[I] + b * cx, Ey[I] + b * cy, Ez[I] + b * cz
        cx = ((Ey[kp] - Ey[I]) - (Ez[jp] - Ez[I])) - J[3] * dxf
        cy = ((Ez[ip] - Ez[I]) - (Ex[kp] - Ex[I])) - J[4] * dxf
        cz = ((Ex[jp] - Ex[I]) - (Ey[ip] - Ey[I])) - J[5] * dxf
        Hx[I], Hy[I], Hz[I] = Hx[I] + dbf * cx, Hy[I] + dbf * cy, Hz[I] + dbf * cz
        done += 1
        if done % block == 0 or n == n0 + nsteps - 1:
            for c in range(6):
                f16[c][...] = to_bits(v[c] * np.float32(1.0 / sc[c]))
                v[c] = from_bits(f16[c]) * np.float32(sc[c])


def _small_problem(oracle, T, scale_exp):
    shape = (13, 11, 10)
    nx, ny, nz = shape
    rng = np.random.default_rng(9)
    cb = (0.5 / rng.uniform(1.0, 8.0, (nz, ny, nx))).astype(np.float32)
    src = [(2, *centre(shape), (np.sin(np.arange(T) * 0.4) * 3).astype(np.float32)),
           (3, 3, 4, 5, (np.cos(np.arange(T) * 0.3)).astype(np.float32))]
    init = [(rng.standard_normal((nz, ny, nx)) * 0.3).astype(np.float32) for _ in range(6)]
    f16 = [oracle.encode_f16(init[c], scale_exp[c]) for c in range(6)]
    return shape, cb, src, f16


@pytest.mark.parametrize("block", [1, 3, 4])
def test_run_f16_block_matches_numpy(oracle, block):
    se = [-3, -2, -2, 1, 0, 2]
    T = 14
    shape, cb, src, a = _small_problem(oracle, T, se)
    b = [x.copy() for x in a]
    oracle.run_f16(shape, a, se, cb, 0.5, sources=src, nsteps=T, block=block)
    _numpy_block_run(b, cb, 0.5, [2.0 ** e for e in se], src, 0, T, block)
    for c in range(6):
        assert np.array_equal(a[c], b[c]), c


def test_run_f16_one_block_is_f32_run_then_encode(oracle):
    """A single block over the whole run == the FP32 oracle (pinned to the
    reference) from the decoded state, encoded once; the DFT sees the FP32
    values."""
    se = [-3, -2, -2, 1, 0, 2]
    T = 12
    shape, cb, src, a = _small_problem(oracle, T, se)
    dft = dict(comps=[0, 2, 4], axis=1, index=5, f_cps=np.array([0.04, 0.09]), scale=0.5)
    w = [oracle.decode_f16(a[c], se[c]) for c in range(6)]
    wre, wim = oracle.run(shape, w, cb=cb, db=0.5, sources=src, nsteps=T, dft=dft)
    re, im = oracle.run_f16(shape, a, se, cb, 0.5, sources=src, nsteps=T, dft=dft, block=T + 5)
    for c in range(6):
        assert np.array_equal(a[c], oracle.encode_f16(w[c], se[c])), c
    assert bits_equal(re, wre) and bits_equal(im, wim)


def test_run_f16_block_differs_from_half_step(oracle):
    se = [0] * 6
    T = 10
    shape, cb, src, a = _small_problem(oracle, T, se)
    b = [x.copy() for x in a]
    oracle.run_f16(shape, a, se, cb, 0.5, sources=src, nsteps=T, block=0)
    oracle.run_f16(shape, b, se, cb, 0.5, sources=src, nsteps=T, block=1)
    assert any(not np.array_equal(a[c], b[c]) for c in range(6))


def test_run_f16_block_threads_do_not_change_results(oracle):
    """The threaded block mode (used for the large FP16 checks) == one thread,
    fields and DFT bit for bit; the threaded conversions of long arrays ==
    numpy's float16 casts."""
    se = [-3, -2, -2, 1, 0, 2]
    T = 11
    shape, cb, src, a = _small_problem(oracle, T, se)
    b = [x.copy() for x in a]
    dft = dict(comps=[0, 5], axis=2, index=4, f_cps=np.array([0.05]), scale=1.0)
    r1 = oracle.run_f16(shape, a, se, cb, 0.5, sources=src, nsteps=T, dft=dft, block=3)
    r4 = oracle.run_f16(shape, b, se, cb, 0.5, sources=src, nsteps=T, dft=dft, block=3,
                        threads=4)
    for c in range(6):
        assert np.array_equal(a[c], b[c]), c
    assert bits_equal(r1[0], r4[0]) and bits_equal(r1[1], r4[1])
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(5_000_000) * 100).astype(np.float32)  # above the parallel cutoff
    bits = oracle.f32_to_f16_bits(x)
    assert np.array_equal(bits, x.astype(np.float16).view(np.uint16))
    assert np.array_equal(oracle.decode_f16(bits), bits.view(np.float16).astype(np.float32))


def test_f32_to_bf16_matches_torch(oracle):
    """bfloat16 encoding == torch's float -> bfloat16 cast (round to nearest even)
    over normal, subnormal, halfway, overflowing and infinite values; decoding
    is the exact widening."""
    to_bits, from_bits = _bf16_torch()
    rng = np.random.default_rng(1)
    parts = [
        np.array([0.0, -0.0, 1.0, -1.0, 3.3895314e38, 3.4e38, -3.4028235e38, np.inf, -np.inf,
                  1e-45, -1e-45, 1.1754942e-38, 1.1754944e-38, 9.2e-41], np.float32),
        (rng.standard_normal(50000) * np.exp(rng.uniform(-87, 87, 50000))).astype(np.float32),
        rng.integers(0, 2**32, 50000, dtype=np.uint64).astype(np.uint32).view(np.float32),
        # exact halfway points between bfloat16 neighbours, both parities
        ((np.arange(0, 4000, dtype=np.uint32) << 16) | 0x8000).view(np.float32),
    ]
    x = np.concatenate(parts)
    x = x[~np.isnan(x)]
    got = oracle.f32_to_bf16_bits(x)
    assert np.array_equal(got, to_bits(x))
    bits = np.arange(0, 65536, 3, dtype=np.uint16)
    bits = bits[(bits & 0x7f80) != 0x7f80]  # finite only
    assert bits_equal(oracle.decode_bf16(bits), from_bits(bits))
    assert np.all(oracle.f32_to_bf16_bits(np.array([np.nan], np.float32)) & 0x7fc0 == 0x7fc0)


@pytest.mark.parametrize("block", [0, 1, 3])
def test_run_bf16_matches_restatement(oracle, block):
    """bfloat16 storage, per half step (block 0) and per block of steps ==
    the numpy restatement with torch's bfloat16 rounding."""
    codec = _bf16_torch()
    se = [-3, 0, -2, 1, 0, 2]
    T = 13
    shape = (13, 11, 10)
    nx, ny, nz = shape
    rng = np.random.default_rng(19)
    cb = (0.5 / rng.uniform(1.0, 8.0, (nz, ny, nx))).astype(np.float32)
    src = [(2, *centre(shape), (np.sin(np.arange(T) * 0.4) * 3).astype(np.float32)),
           (3, 3, 4, 5, (np.cos(np.arange(T) * 0.3)).astype(np.float32))]
    init = [(rng.standard_normal((nz, ny, nx)) * 0.3).astype(np.float32) for _ in range(6)]
    a = [oracle.encode_bf16(init[c], se[c]) for c in range(6)]
    for c in range(6):
        assert np.array_equal(a[c], codec[0](init[c] * np.float32(2.0 ** -se[c])))
    b = [x.copy() for x in a]
    oracle.run_f16(shape, a, se, cb, 0.5, sources=src, nsteps=T, block=block, fmt="bf16")
    sc = [2.0 ** e for e in se]
    if block == 0:
        for n in range(T):
            _numpy_f16_step(b, cb, 0.5, sc, src, n, 1.0, codec=codec)
    else:
        _numpy_block_run(b, cb, 0.5, sc, src, 0, T, block, codec=codec)
    for c in range(6):
        assert np.array_equal(a[c], b[c]), c
    # bfloat16 and binary16 storage give different fields
    h = [oracle.encode_f16(init[c], se[c]) for c in range(6)]
    oracle.run_f16(shape, h, se, cb, 0.5, sources=src, nsteps=T, block=block)
    assert not np.allclose(oracle.decode_bf16(a[2], se[2]), oracle.decode_f16(h[2], se[2]),
                           rtol=0, atol=0)
