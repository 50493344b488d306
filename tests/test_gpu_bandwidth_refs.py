"""The bandwidth reference kernels behind the roofline figures: the streaming
copy (vlb_device_copy, which also calibrates rocprofv3's HBM byte counters)
moves the bytes exactly, and the read-only pass (vlb_device_read, the
attainable read bandwidth for the skybox projection) runs over any grid and
refuses sizes that are not whole 16-byte words."""
import numpy as np
import pytest

import vlb

pytestmark = pytest.mark.gpu


def test_copy_moves_bytes_exactly(ctx):
    n = (1 << 20) + 48
    src = np.random.default_rng(7).integers(0, 256, n, dtype=np.uint8)
    d_src, d_dst = ctx.alloc(n), ctx.alloc(n)
    try:
        ctx.upload(d_src, src)
        ms = ctx.device_copy(d_dst, d_src, n)
        got = np.zeros(n, np.uint8)
        ctx.download(d_dst, got)
        assert np.array_equal(got, src)
        assert ms > 0.0
    finally:
        ctx.free(d_dst)
        ctx.free(d_src)


@pytest.mark.parametrize("grid", [0, 1, 3, 7, 4096])
def test_read_pass_runs_and_leaves_memory_unchanged(ctx, grid):
    n = (3 << 20) + 16 * 5
    src = np.random.default_rng(grid).standard_normal(n // 4).astype(np.float32)
    d = ctx.alloc(n)
    try:
        ctx.upload(d, src)
        ms = ctx.device_read(d, n, grid)
        assert ms > 0.0
        got = np.zeros_like(src)
        ctx.download(d, got)
        assert np.array_equal(got.view(np.uint32), src.view(np.uint32))
    finally:
        ctx.free(d)


def test_read_pass_rejects_partial_words(ctx):
    d = ctx.alloc(64)
    try:
        with pytest.raises(vlb.VlbError) as e:
            ctx.device_read(d, 40)
        assert e.value.status == vlb.ERR_INVALID
        with pytest.raises(vlb.VlbError):
            ctx.device_read(d, 64, 1 << 21)
    finally:
        ctx.free(d)
