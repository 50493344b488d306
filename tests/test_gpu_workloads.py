"""GPU parity on the benchmark workloads themselves, at their full sizes.

The precision form (1,154,050-triangle terrain, 16384 probes x 16384 rays,
L2 and L3), the batched form (64 jobs over the 11k/99k rooms and the
0.5M/1.15M terrains) and the serving form (requests over the ~2M-triangle
city block and those four scenes, arriving while others run) go through
libvlb.so exactly as bench.py runs them (workloads.py builds the inputs for
both); the FP64 oracle checks seeded probe subsets of every job / request.

Bars (as for the small scenes): hit prim ids, t, u, v and visibility are
bit-exact against the oracle (sh.comp:23-42 and env_map.rchit:51-102 run
over the same rays); SH coefficients within 1e-5 relative L2 of FP64 in the
FP32 form and 1e-4 in the split BF16/FP16 forms; a job or request gives
bit-identical coefficients alone and in the mix.
"""
import threading
import time

import numpy as np
import pytest

import orc
import vlb
import workloads as W
from conftest import rel_err

pytestmark = pytest.mark.gpu

FP32_BAR = 1e-5
SPLIT16_BAR = 1e-4


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def world(ctx):
    """The serving form's five scenes (city first), uploaded once, with
    their oracle scenes built on first use."""
    host = W.scenes(W.C5_SCENES)
    sids = [ctx.upload_scene_dict(s) for s in host]
    oracles = {}

    def oracle(i):
        if i not in oracles:
            oracles[i] = orc.Scene.from_dict(host[i])
        return oracles[i]

    yield host, sids, oracle
    for sid in sids:
        ctx.release_scene(sid)


@pytest.fixture(scope="module")
def terrain(ctx):
    s = W.c3_scene()
    sid = ctx.upload_scene_dict(s)
    yield s, sid, orc.Scene.from_dict(s)
    ctx.release_scene(sid)


def _subset(n, k, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, size=min(k, n), replace=False))


# ------------------------------------------------------------ precision form
def test_c3_scene_size(ctx, terrain):
    s, sid, o = terrain
    info = ctx.scene_info(sid)
    assert info["num_tris"] == 1154050
    assert info["depth"] >= 20


@pytest.mark.parametrize("variant", [12, 16])  # fused (float nodes), wavefront (quantized nodes)
def test_c3_primary_hits_bit_exact(ctx, terrain, variant):
    s, sid, o = terrain
    probes = W.c3_probes(s)
    sub = _subset(len(probes), 6, 31)
    p = W.c3_params(probes[sub].copy(), vlb.PREC_FP32, 2, W.c3_sun())
    p.trace_variant = variant
    _, ghit = ctx.bake_radiance(sid, p, want_radiance=False, want_hits=True)
    chit = o.bake(p, mode64=False, want_hits=True)["hits"]
    assert ghit.shape == chit.shape == (6, 16384)
    mism = int((ghit != chit).sum())
    assert mism == 0, f"{mism} of {ghit.size} primary hits differ"
    assert 0.05 < (chit >= 0).mean() < 0.98   # terrain below, sky above


def test_c3_and_city_trace_queries_bit_exact(ctx, terrain, world):
    """Random rays through the terrain and the city: closest hits (prim, t,
    u, v) and shadow queries, against the oracle's traversal."""
    cases = [(terrain[1], terrain[2], terrain[0])]
    host, sids, oracle = world
    cases.append((sids[0], oracle(0), host[0]))
    for sid, o, s in cases:
        rng = np.random.default_rng(sid)
        lo, hi = np.asarray(s["bounds_min"]), np.asarray(s["bounds_max"])
        n = 40000
        org = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
        d = rng.normal(size=(n, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        gp, gt, gu, gv = ctx.trace_closest(sid, org, d, 0.001, 1e5)
        cp, ct, cu, cv = o.trace_closest(org, d, 0.001, 1e5)
        assert (gp == cp).all(), f"prim mismatch on {(gp != cp).sum()} rays"
        hit = cp >= 0
        assert hit.mean() > 0.2
        for g, c in ((gt, ct), (gu, cu), (gv, cv)):
            assert np.array_equal(_bits(g[hit]), _bits(c[hit]))
        tmax = rng.uniform(0.5, float(np.linalg.norm(hi - lo)) / 4, size=n).astype(np.float32)
        go = ctx.trace_occluded(sid, org, d, tmax, 0.0)
        co = o.trace_occluded(org, d, tmax, 0.0)
        assert (go == co).all(), f"visibility mismatch on {(go != co).sum()} rays"


@pytest.mark.parametrize("order", [2, 3])
def test_c3_full_bake_error_per_band(ctx, terrain, order):
    """The whole 16384 x 16384 bake in FP32, split BF16 and split FP16; a
    seeded subset of 24 probes against the FP64 oracle, band by band."""
    s, sid, o = terrain
    probes = W.c3_probes(s)
    sun = W.c3_sun()
    sub = _subset(len(probes), 24, 3 + order)
    ref = o.bake(W.c3_params(probes[sub].copy(), vlb.PREC_FP32, order, sun), mode64=True,
                 nthreads=0)["coeffs"]
    worst = {}
    for name, prec, bar in (("fp32", vlb.PREC_FP32, FP32_BAR), ("bf16", vlb.PREC_BF16, SPLIT16_BAR),
                            ("fp16", vlb.PREC_FP16, SPLIT16_BAR)):
        plan = ctx.plan(sid, W.c3_params(probes, prec, order, sun))
        try:
            plan.execute()
            got = plan.download()[sub]
        finally:
            plan.destroy()
        assert np.isfinite(got).all()
        for l in range(order + 1):
            ks = slice(l * l, (l + 1) * (l + 1))
            e = float(rel_err(got[:, ks], ref[:, ks]).max())
            worst[(name, l)] = e
        e_all = float(rel_err(got, ref).max())
        assert e_all <= bar, f"{name} L{order}: {e_all:.3e} > {bar:.0e}; per band {worst}"


def test_c3_wavefront_and_fused_bitwise(ctx, terrain):
    s, sid, o = terrain
    probes = W.c3_probes(s)[::8].copy()   # 2048 probes
    out = []
    for v in (12, 16, 20):
        p = W.c3_params(probes, vlb.PREC_BF16, 2, W.c3_sun())
        p.trace_variant = v
        out.append(ctx.bake(sid, p))
    assert np.array_equal(_bits(out[0]), _bits(out[1]))
    assert np.array_equal(_bits(out[0]), _bits(out[2]))


# ------------------------------------------------------------- batched form
def _c4(world):
    host, sids, oracle = world
    spec, sun = W.c4_jobs(host[1:])
    return spec, sun, sids[1:], (lambda i: oracle(i + 1))


@pytest.mark.parametrize("prec,bar", [(vlb.PREC_BF16, SPLIT16_BAR), (vlb.PREC_FP32, FP32_BAR)])
def test_c4_batch_vs_oracle(ctx, world, prec, bar):
    """All 64 jobs as one batch (graph replay); 3 seeded probes of every job
    against the FP64 oracle."""
    spec, sun, sids, oracle = _c4(world)
    jobs = [(sids[si], W.sun_params(pr, R, prec, sun)) for si, pr, R in spec]
    b = ctx.batch(jobs)
    try:
        b.use_graph(True)
        b.execute()
        b.execute()            # a replay of the captured graph
        got = [b.download(j) for j in range(len(jobs))]
        info = b.info()
    finally:
        b.destroy()
    assert info["jobs"] == 64 and info["rounds"] >= 2
    worst = 0.0
    for j, (si, pr, R) in enumerate(spec):
        sub = _subset(pr.shape[0], 3, 100 + j)
        ref = oracle(si).bake(W.sun_params(pr[sub].copy(), R, vlb.PREC_FP32, sun), mode64=True)["coeffs"]
        assert np.isfinite(got[j]).all()
        e = float(rel_err(got[j][sub], ref).max())
        worst = max(worst, e)
        assert e <= bar, f"job {j} (scene {si}, {pr.shape[0]} probes x {R} rays): {e:.3e}"


def test_c4_job_alone_equals_in_batch(ctx, world):
    """Jobs on each of the four scenes: the same coefficients, bit for bit,
    in a batch of their own and inside the 64-job batch."""
    spec, sun, sids, _ = _c4(world)
    jobs = [(sids[si], W.sun_params(pr, R, vlb.PREC_BF16, sun)) for si, pr, R in spec]
    picks = {}
    for j, (si, pr, R) in enumerate(spec):
        if si not in picks and pr.shape[0] <= 8192:
            picks[si] = j
    assert len(picks) == 4
    b = ctx.batch(jobs)
    try:
        b.execute()
        mix = {j: b.download(j) for j in picks.values()}
    finally:
        b.destroy()
    for j in picks.values():
        one = ctx.batch([jobs[j]])
        try:
            one.execute()
            alone = one.download(0)
        finally:
            one.destroy()
        assert np.array_equal(_bits(alone), _bits(mix[j])), f"job {j}"


# ------------------------------------------------------------- serving form
def test_c5_requests_in_the_mix_vs_oracle_and_alone(ctx, world):
    """40 requests of the serving form arriving at 40/s (so most arrive while
    others run): every request against the FP64 oracle on 4 seeded probes,
    and three of them bit-identical when served alone afterwards."""
    host, sids, oracle = world
    arrive, reqs, _ = W.c5_requests(host, 40, 40.0)
    assert {r[0] for r in reqs} == set(range(5))        # the city is in the mix
    srv = vlb.Server(ctx, precision=vlb.PREC_BF16, state_budget=256 << 20, step_rays=1 << 25)
    try:
        params = [W.c5_params(r, vlb.PREC_BF16) for r in reqs]
        ids = [None] * len(reqs)

        def feeder():
            t0 = time.perf_counter()
            for i in range(len(reqs)):
                dt = t0 + arrive[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids[i] = srv.submit(sids[reqs[i][0]], params[i])

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
        assert st["completed"] == 40 and st["active_peak"] > 1
        for i, ((got, info), r) in enumerate(zip(res, reqs)):
            assert info["status"] == 0
            sub = _subset(r[1].shape[0], 4, 500 + i)
            q = W.c5_params((r[0], r[1][sub].copy(), r[2], r[3]), vlb.PREC_FP32)
            ref = oracle(r[0]).bake(q, mode64=True)["coeffs"]
            e = float(rel_err(got[sub], ref).max())
            assert e <= SPLIT16_BAR, f"request {i} (scene {r[0]}, {r[1].shape[0]} probes): {e:.3e}"
        for i in (0, 7, 23):
            alone, _ = srv.wait(srv.submit(sids[reqs[i][0]], params[i]))
            assert np.array_equal(_bits(alone), _bits(res[i][0])), f"request {i}"
    finally:
        srv.destroy()
