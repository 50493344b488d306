"""The allocation-free entry point vlb_skybox_sh_noalloc against the entry
point it varies, vlb_skybox_sh_d (the library's form of skybox_sh.comp,
shaders/skybox_sh.comp:23-42), through the C ABI: two calls in a row, the
second on a smaller input (the fixed scratch keeps the first call's
contents), each within 1e-5 relative L2 of the allocating entry point and
of the FP64 oracle; a request beyond the scratch fails cleanly."""
import ctypes as C

import numpy as np
import pytest

import orc
import vlb
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _cubes(N, B, seed):
    rng = np.random.default_rng(seed)
    return np.stack([vlb.procedural_cubemap(N, vlb.sun_direction(rng)) for _ in range(B)]).astype(np.float32)


def test_noalloc_equals_allocating_entry_point_on_two_calls(ctx):
    for N, B, order, seed in ((128, 6, 2, 1), (40, 2, 3, 2)):   # the second call is smaller
        K = 9 if order == 2 else 16
        cubes = _cubes(N, B, seed)
        d_in = ctx.alloc(cubes.nbytes)
        d_a = ctx.alloc(B * K * 3 * 4)
        d_b = ctx.alloc(B * K * 3 * 4)
        try:
            ctx.upload(d_in, cubes)
            ctx.skybox_sh_device(d_in, N, B, order, d_a)
            st = vlb.lib.vlb_skybox_sh_noalloc(ctx.h, C.c_void_p(d_in), N, B, order, C.c_void_p(d_b))
            assert st == vlb.OK, ctx.last_error()
            ctx.sync()
            a = np.zeros((B, K, 3), np.float32)
            b = np.zeros((B, K, 3), np.float32)
            ctx.download(d_a, a)
            ctx.download(d_b, b)
            assert np.isfinite(b).all()
            assert rel_err(b, a).max() <= 1e-5
            for i in range(B):
                assert rel_err(b[i], orc.skybox_sh(cubes[i], K)).max() <= 1e-5
        finally:
            for d in (d_in, d_a, d_b):
                ctx.free(d)


def test_noalloc_rejects_requests_beyond_its_scratch(ctx):
    d = ctx.alloc(1 << 12)
    try:
        # a 2048^2 face needs a 16 MiB solid-angle table alone
        st = vlb.lib.vlb_skybox_sh_noalloc(ctx.h, C.c_void_p(d), 2048, 1, 2, C.c_void_p(d))
        assert st == vlb.ERR_INVALID
        assert "scratch" in ctx.last_error()
        st = vlb.lib.vlb_skybox_sh_noalloc(ctx.h, C.c_void_p(d), 16, 1, 4, C.c_void_p(d))
        assert st == vlb.ERR_INVALID
    finally:
        ctx.free(d)
