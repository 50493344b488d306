"""GPU tests of the rebake server (continuous batching).

Bars: a request's coefficients are bit-identical to the same bake run alone
through vlb_bake (same rays per slice) and to the same request served alone,
whatever else is in flight, the step size or the admission budget; FP32
requests are within 1e-5 relative L2 of the FP64 oracle, BF16 (split) within
1e-4.
"""
import threading

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
    ids = [ctx.upload_scene_dict(s) for s in (room, room2)]
    yield list(zip((room, room2), ids))
    for i in ids:
        ctx.release_scene(i)


def _requests(scenes, n=10, seed=3, precision=vlb.PREC_FP32):
    """Uneven rebake requests: a light moves or a point light is added."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        s, sid = scenes[i % len(scenes)]
        P = int(rng.integers(1, 200))
        lights = [(vlb.LIGHT_DIRECTIONAL, (0.3, -0.9, 0.2), (2.0, 2.0, 2.0))]
        if i % 2:
            lights.append((vlb.LIGHT_POINT, tuple(rng.uniform([1, 1, 1], [9, 3, 9])),
                           (float(rng.uniform(10, 500)),) * 3))
        rays = int(rng.choice([256, 512, 1000]))
        out.append((sid, _params(_probes(P, 100 + i), rays=rays, rps=256, lights=lights,
                                 precision=precision)))
    return out


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


# 0: the server's default (the fused kernel, variant 35: lane map 2, far-first
# shadow queries); 12 and 3: the
# fused kernel at 8 waves/SIMD and at the compiler's occupancy; 20: the
# wavefront multi-job kernels (k_rays_jobs / k_shade_jobs, lane map 3); 36:
# variant 12 with far-first shadow queries
@pytest.mark.parametrize("variant", [0, 12, 3, 20, 36])
def test_server_request_equals_single_bake(ctx, scenes, variant):
    reqs = _requests(scenes, n=3)
    srv = vlb.Server(ctx, step_rays=1 << 16, trace_variant=variant)
    assert srv.stats()["trace_variant"] == (variant or 35)
    for sid, p in reqs:
        got, info = srv.wait(srv.submit(sid, p))
        assert info["status"] == 0 and info["done_ms"] >= info["start_ms"] >= info["admit_ms"]
        assert np.array_equal(_bits(got), _bits(ctx.bake(sid, p)))
    srv.destroy()


# age 1: SRPT with an age bound of 1 ms, so that requests pass from SRPT
# order to oldest-first while the mix runs
@pytest.mark.parametrize("policy,variant,age", [(vlb.SCHED_SRPT, 0, 0), (vlb.SCHED_SHARE, 0, 0),
                                                (vlb.SCHED_FIFO, 0, 0), (vlb.SCHED_SRPT, 12, 0),
                                                (vlb.SCHED_SRPT, 20, 0), (vlb.SCHED_SHARE, 20, 0),
                                                (vlb.SCHED_SRPT, 0, 1)])
def test_server_alone_equals_in_mix(ctx, scenes, policy, variant, age):
    reqs = _requests(scenes, n=12, seed=5)
    target = 4
    srv = vlb.Server(ctx, step_rays=1 << 15, policy=policy, trace_variant=variant, age_ms=age)
    alone, _ = srv.wait(srv.submit(*reqs[target]))
    ids = [srv.submit(sid, p) for sid, p in reqs]      # all in flight together
    res = [srv.wait(i) for i in ids]
    assert np.array_equal(_bits(res[target][0]), _bits(alone))
    assert max(r[1]["steps"] for r in res) > 1          # requests spanned several steps
    st = srv.stats()
    assert st["completed"] == 13 and st["active_peak"] > 1
    byid = {sid: s for s, sid in scenes}
    for (sid, p), (got, _) in zip(reqs[:4], res[:4]):
        ref = orc.Scene.from_dict(byid[sid]).bake(p, mode64=True)["coeffs"]
        assert rel_err(got, ref).max() <= 1e-5
    srv.destroy()


def test_server_age_bound_is_for_srpt_only(ctx):
    for policy in (vlb.SCHED_SHARE, vlb.SCHED_FIFO):
        with pytest.raises(vlb.VlbError) as e:
            vlb.Server(ctx, policy=policy, age_ms=50)
        assert e.value.status == vlb.ERR_INVALID
    vlb.Server(ctx, policy=vlb.SCHED_SRPT, age_ms=50).destroy()


def test_server_admission_by_state_budget(ctx, scenes):
    reqs = _requests(scenes, n=8, seed=9)
    srv = vlb.Server(ctx, step_rays=1 << 15, state_budget=1 << 19, max_active=3)
    ids = [srv.submit(sid, p) for sid, p in reqs]
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    assert st["state_bytes_peak"] <= 1 << 19
    assert st["active_peak"] <= 3
    for (sid, p), (got, info) in zip(reqs, res):
        assert np.array_equal(_bits(got), _bits(ctx.bake(sid, p)))
    srv.destroy()


def test_server_bf16_vs_oracle(ctx, scenes):
    reqs = _requests(scenes, n=4, seed=11, precision=vlb.PREC_BF16)
    srv = vlb.Server(ctx, precision=vlb.PREC_BF16)
    ids = [srv.submit(sid, p) for sid, p in reqs]
    byid = {sid: s for s, sid in scenes}
    for (sid, p), i in zip(reqs, ids):
        got, _ = srv.wait(i)
        ref = orc.Scene.from_dict(byid[sid]).bake(p, mode64=True)["coeffs"]
        assert rel_err(got, ref).max() <= 1e-4
    srv.destroy()


def test_server_rejects_bad_requests_and_drains_on_destroy(ctx, scenes):
    (_, sid) = scenes[0]
    srv = vlb.Server(ctx, state_budget=1 << 16, max_lights=1)
    with pytest.raises(vlb.VlbError):
        srv.submit(sid, _params(_probes(4, 1), precision=vlb.PREC_BF16))
    two = [(vlb.LIGHT_DIRECTIONAL, (0, -1, 0), (1, 1, 1))] * 2
    with pytest.raises(vlb.VlbError):
        srv.submit(sid, _params(_probes(4, 1), lights=two))
    with pytest.raises(vlb.VlbError) as e:
        srv.submit(sid, _params(_probes(5000, 1)))
    assert e.value.status == vlb.ERR_BUDGET
    rid = srv.submit(sid, _params(_probes(20, 2)))
    for _ in range(100000):
        if srv.poll(rid):
            break
    got, info = srv.wait(rid)
    assert info["status"] == 0
    srv.submit(sid, _params(_probes(20, 3)))   # pending at destroy: cancelled
    srv.destroy()


def test_concurrent_failing_submits_report_their_errors(ctx, scenes):
    # requests fail from several threads at once; each reads a whole message
    (_, sid) = scenes[0]
    srv = vlb.Server(ctx, precision=vlb.PREC_FP32, max_lights=1)
    two = [(vlb.LIGHT_DIRECTIONAL, (0, -1, 0), (1, 1, 1))] * 2
    bad = [_params(_probes(8, 1), precision=vlb.PREC_BF16), _params(_probes(8, 2), lights=two)]
    words = ["precision", "max_lights"]
    seen, errors = [], []

    def worker(k):
        try:
            for i in range(200):
                j = (i + k) % 2
                try:
                    srv.submit(sid, bad[j])
                    errors.append("accepted a bad request")
                except vlb.VlbError as e:
                    assert e.status == vlb.ERR_INVALID
                    msg = str(e)
                    seen.append(msg)
                    if not any(w in msg for w in words):
                        errors.append(msg)
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append(repr(e))

    ths = [threading.Thread(target=worker, args=(k,)) for k in range(8)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    srv.destroy()
    assert not errors, errors[:3]
    assert len(seen) == 1600
