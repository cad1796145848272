"""Large field copies to the host go through two pinned staging slots in
64 MiB chunks (fields of 8 MiB and more); smaller ones are copied directly.
Both must hand back exactly what the device holds: FP32 fields equal what
was uploaded bit for bit, 16-bit and FP64 fields equal the FP64 read-back
(a direct copy, exact for every storage) rounded to float."""
import numpy as np
import pytest

from cases import bits_equal

pytestmark = pytest.mark.gpu

# 260 x 256 x 300 = 19.97 M cells: FP32 80 MB (two chunks, the second
# partial), 16-bit 40 MB (one partial chunk), FP64 160 MB (three chunks)
BIG = (260, 256, 300)
SMALL = (40, 36, 30)  # below the staging threshold for every storage


def _fields(shape, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    return [(rng.standard_normal((nz, ny, nx)) * 4.0).astype(np.float32) for _ in range(6)]


@pytest.mark.parametrize("shape", [BIG, SMALL], ids=["staged", "direct"])
def test_fp32_fields_round_trip(gd, shape):
    f = _fields(shape, 3)
    with gd.Context(*shape) as ctx:
        ctx.set_fields(f)
        got = ctx.get_fields()
        for a in range(6):
            assert bits_equal(got[a], f[a]), a
        assert bits_equal(ctx.get_field(4), f[4])


@pytest.mark.parametrize("shape", [BIG, SMALL], ids=["staged", "direct"])
@pytest.mark.parametrize("storage", ["f16", "bf16", "f64"])
def test_other_storage_reads_match_fp64_read(gd, shape, storage):
    f = _fields(shape, 5)
    with gd.Context(*shape) as ctx:
        if storage == "f16":
            ctx.set_storage(gd.F16, [-2, -1, 0, 1, 2, 3])
        elif storage == "bf16":
            ctx.set_storage(gd.BF16, [0, 0, 1, 1, -1, -1])
        else:
            ctx.set_storage(gd.F64)
        ctx.set_fields(f)
        want = [w.astype(np.float32) for w in ctx.get_fields_f64()]
        got = ctx.get_fields()
        for a in range(6):
            assert bits_equal(got[a], want[a]), a
        if storage == "f64":
            # FP64 holds every float32 input exactly: the upload and both
            # read-backs must return the input itself
            for a in range(6):
                assert bits_equal(got[a], f[a]), a
        else:
            # the stored values are a rounding of the input, not the input
            assert not bits_equal(got[0], f[0])


def test_staged_read_after_run(gd):
    """After stepping (fields in the second copy of the TILED ping-pong) the
    staged read returns the same values as the FP64 read."""
    shape = BIG
    with gd.Context(*shape) as ctx:
        ctx.set_isotropic_eps(np.full((shape[2], shape[1], shape[0]), 2.0, np.float32),
                              np.float32(0.5))
        ctx.set_fields(_fields(shape, 9))
        ctx.run(3, gd.TILED)
        want = [w.astype(np.float32) for w in ctx.get_fields_f64()]
        got = ctx.get_fields()
        for a in range(6):
            assert bits_equal(got[a], want[a]), a


def test_material_from_host_eps_large_grid(gd):
    """Material set from host eps on a grid that the host indexes on several
    threads (6.3 M cells): the indexed form holds one entry per distinct Cb
    bit pattern and steps to the same fields as the dense form."""
    shape = (256, 256, 96)
    nx, ny, nz = shape
    eps = gd.sphere_medium(nx, ny, nz, 21, 40, 3.0, 12.0, 1.0, 12.0)
    dt = np.float32(0.99 / np.sqrt(3.0))
    cb, _ = gd.iso_cb_from_eps(eps, dt, 1.0)
    distinct = np.unique(cb.view(np.uint32)).size
    assert 2 < distinct <= 256
    src = gd.gaussian_pulse_table(8, 3.0, 2.0, 0.05)
    out = {}
    for fmt in (gd.MAT_INDEXED, gd.MAT_DENSE):
        with gd.Context(*shape) as ctx:
            ctx.set_material_format(fmt)
            ctx.set_isotropic_eps(eps, dt)
            f, ntable = ctx.get_material_format()
            assert f == fmt
            if fmt == gd.MAT_INDEXED:
                assert ntable == distinct
            ctx.add_source(gd.EZ, nx // 2, ny // 2, nz // 2, src)
            ctx.set_fields(_fields(shape, 13))
            ctx.run(8, gd.TILED)
            out[fmt] = ctx.get_fields()
    for a in range(6):
        assert bits_equal(out[gd.MAT_INDEXED][a], out[gd.MAT_DENSE][a]), a
