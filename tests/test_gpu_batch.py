"""GPU parity of batches: many bakes over several resident scenes as one unit.

Bars: every job of a batch gives the bit-identical coefficients of the same
bake run alone through vlb_bake (same rays per slice), whatever the other
jobs, the round size, the lane map, the work-fetch chunk or graph replay;
FP32 jobs are within 1e-5 relative L2 of the FP64 oracle, BF16 (split) jobs
within 1e-4.
"""
import numpy as np
import pytest

import orc
import vlb
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _params(probes, rays=512, rps=256, **kw):
    rng = np.random.default_rng(7)
    sun = vlb.sun_direction(rng)
    base = dict(probes=probes, rays=rays, sh_order=2, shade_mode=vlb.SHADE_LAMBERT,
                lights=[(vlb.LIGHT_DIRECTIONAL, (-sun).tolist(), (3.0, 2.5, 2.0))],
                sun_dir=sun.tolist(), sky_zenith=(2.0, 1.8, 1.5), sky_horizon=(0.5, 0.6, 0.7),
                sun_rgb=(50.0, 45.0, 40.0), shadow_bias=1e-3, rays_per_slice=rps)
    base.update(kw)
    return vlb.Params(**base)


def _probes(n, seed, lo=(0.3, 0.3, 0.3), hi=(9.7, 3.7, 9.7)):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)


@pytest.fixture(scope="module")
def scenes(ctx):
    room = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    room2 = vlb.generate_scene(vlb.SCENE_ROOM, 11, 16, 3)
    sph = vlb.generate_scene(vlb.SCENE_SPHERES, 4, 12, 2)
    ids = [ctx.upload_scene_dict(s) for s in (room, room2, sph)]
    yield list(zip((room, room2, sph), ids))
    for i in ids:
        ctx.release_scene(i)


def _jobs(scenes, precision=vlb.PREC_FP32):
    """Uneven jobs: ragged probe and ray counts, 0..3 lights, both shading modes."""
    (room, r0), (room2, r1), (sph, r2) = scenes
    lo_s, hi_s = sph["bounds_min"], sph["bounds_max"]
    pt = (vlb.LIGHT_POINT, (5.0, 3.5, 5.0), (40.0, 30.0, 20.0))
    return [
        (r0, _params(_probes(37, 1), rays=512, precision=precision)),
        (r2, _params(_probes(16, 2, lo_s, hi_s), rays=300, rps=128, precision=precision)),
        (r1, _params(_probes(5, 3), rays=1000, rps=512, precision=precision,
                     lights=[(vlb.LIGHT_DIRECTIONAL, (0.3, -0.9, 0.2), (2.0, 2.0, 2.0)), pt,
                             (vlb.LIGHT_POINT, (2.0, 1.0, 8.0), (10.0, 10.0, 10.0))])),
        (r0, _params(_probes(1, 4), rays=64, rps=64, precision=precision, lights=[])),
        (r1, _params(_probes(23, 5), rays=256, rps=128, precision=precision,
                     shade_mode=vlb.SHADE_REFERENCE)),
        (r0, _params(_probes(64, 6), rays=2048, rps=1024, precision=precision, lights=[pt])),
    ]


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_batch_jobs_equal_single_bakes_bitwise(ctx, scenes):
    jobs = _jobs(scenes)
    single = [ctx.bake(sid, p) for sid, p in jobs]
    # small rounds: jobs split across rounds, rounds hold several jobs
    # wavefront and fused kernels
    # (25..36: fused with far-first shadow queries)
    for mrr, var, chunk in [(0, 20, 256), (1 << 16, 13, 1024), (1 << 14, 18, 64), (0, 12, 256),
                            (1 << 16, 1, 256), (1 << 14, 7, 256), (0, 36, 256), (1 << 16, 25, 256)]:
        b = ctx.batch(jobs, max_round_rays=mrr, trace_variant=var, chunk=chunk)
        info = b.info()
        if mrr:
            assert info["rounds"] > 1
        b.execute()
        for j in range(len(jobs)):
            got = b.download(j)
            assert np.array_equal(_bits(got), _bits(single[j])), (mrr, var, chunk, j)
        b.destroy()


def test_batch_fp32_vs_fp64_oracle(ctx, scenes):
    jobs = _jobs(scenes)
    b = ctx.batch(jobs, keep_f64=True)
    b.execute()
    byid = {sid: s for s, sid in scenes}
    for j, (sid, p) in enumerate(jobs):
        ref = orc.Scene.from_dict(byid[sid]).bake(p, mode64=True)["coeffs"]
        err = rel_err(b.download(j), ref)
        assert err.max() <= 1e-5, (j, err.max())
        err64 = rel_err(b.download_f64(j), ref)
        assert err64.max() <= 1e-5, (j, err64.max())
    b.destroy()


def test_batch_bf16_vs_fp64_oracle(ctx, scenes):
    jobs = _jobs(scenes, precision=vlb.PREC_BF16)
    b = ctx.batch(jobs)
    b.execute()
    byid = {sid: s for s, sid in scenes}
    for j, (sid, p) in enumerate(jobs):
        ref = orc.Scene.from_dict(byid[sid]).bake(p, mode64=True)["coeffs"]
        err = rel_err(b.download(j), ref)
        assert err.max() <= 1e-4, (j, err.max())
        assert np.array_equal(_bits(b.download(j)), _bits(ctx.bake(sid, p)))
    b.destroy()


def test_batch_job_alone_equals_job_in_mix(ctx, scenes):
    jobs = _jobs(scenes)
    alone = ctx.batch([jobs[2]])
    alone.execute()
    a = alone.download(0)
    mix = ctx.batch(jobs[::-1] + jobs, max_round_rays=1 << 15)
    mix.execute()
    k = len(jobs) - 1 - 2
    assert np.array_equal(_bits(mix.download(k)), _bits(a))
    assert np.array_equal(_bits(mix.download(len(jobs) + 2)), _bits(a))
    alone.destroy()
    mix.destroy()


def test_batch_graph_replays_match_eager(ctx, scenes):
    jobs = _jobs(scenes)
    b = ctx.batch(jobs, max_round_rays=1 << 16)
    b.execute()
    eager = [b.download(j) for j in range(len(jobs))]
    b.use_graph(True)
    for rep in range(3):
        b.execute()
        b.sync()
        _ = np.ones(1000).sum()  # host work between replays
        for j in range(len(jobs)):
            assert np.array_equal(_bits(b.download(j)), _bits(eager[j])), (rep, j)
    b.destroy()


def test_batch_autotune_and_counts(ctx, scenes):
    jobs = _jobs(scenes)
    b = ctx.batch(jobs)
    b.execute()
    before = [b.download(j) for j in range(len(jobs))]
    v, c = b.autotune(1)
    # fused kernel (1..12, 25..36) at chunk 256 or wavefront kernels (13..24)
    assert ((1 <= v <= 12 or 25 <= v <= 36) and c == 256) or (13 <= v <= 24 and c in (256, 1024))
    b.execute()
    for j in range(len(jobs)):
        assert np.array_equal(_bits(b.download(j)), _bits(before[j]))
    b.count()
    info = b.info()
    # shadow rays traced = sum over jobs of the single-bake counts
    want = 0
    for sid, p in jobs:
        pl = ctx.plan(sid, p)
        pl.count()
        want += pl.stats()["shadow_rays"]
        pl.destroy()
    assert info["shadow_rays"] == want
    assert info["node_visits"] > 0 and info["tri_tests"] > 0
    assert info["primary_rays"] == sum(p.probes.shape[0] * p.R for _, p in jobs)
    b.destroy()


def test_batch_empty_scene_job(ctx, scenes):
    sid0 = ctx.upload_scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    (room, r0) = scenes[0]
    jobs = [(sid0, _params(_probes(5, 8), rays=300, rps=128)), (r0, _params(_probes(9, 9), rays=256))]
    b = ctx.batch(jobs)
    b.execute()
    ref = orc.Scene(np.zeros((0, 3)), np.zeros((0, 3), np.uint32)).bake(jobs[0][1])["coeffs"]
    assert rel_err(b.download(0), ref).max() <= 1e-5
    assert np.array_equal(_bits(b.download(1)), _bits(ctx.bake(r0, jobs[1][1])))
    b.destroy()
    ctx.release_scene(sid0)


def test_batch_rejects_mixed_precision_and_bad_options(ctx, scenes):
    (room, r0) = scenes[0]
    a = (r0, _params(_probes(4, 1)))
    b = (r0, _params(_probes(4, 1), precision=vlb.PREC_BF16))
    with pytest.raises(vlb.VlbError):
        ctx.batch([a, b])
    with pytest.raises(vlb.VlbError):
        ctx.batch([a], trace_variant=37)
    with pytest.raises(vlb.VlbError):
        ctx.batch([a], chunk=100)


def test_context_outlives_its_batches_plans_and_servers(scenes):
    """vlb_ctx_destroy refuses while a batch, plan or server made from the
    context is alive; Context.close() destroys them first."""
    c2 = vlb.Context(0)
    (room, r0) = scenes[0]
    sid = c2.upload_scene_dict(room)
    p = _params(_probes(3, 8), rays=256)
    b = c2.batch([(sid, p)])
    pl = c2.plan(sid, p)
    srv = vlb.Server(c2, state_budget=1 << 22, step_rays=1 << 15)
    for obj in (b, pl, srv):
        assert vlb.lib.vlb_ctx_destroy(c2.h) == vlb.ERR_INVALID
        assert "destroy them first" in c2.last_error()
        obj.destroy()
    pl2 = c2.plan(sid, p)
    c2.close()                      # destroys pl2, then the context
    assert not pl2.h and not c2.h
