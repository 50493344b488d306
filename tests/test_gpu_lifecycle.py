"""Failure paths and object lifetimes through the C ABI, forced from the host.

The reference's only failure handling is an exception caught in main
(src/baker/main.cpp:26-40) and fence waits that print a failure instead of
raising it (src/application.cpp:263-276).  Here: a scene cannot be released
while a plan, a batch job or an unfinished server request uses it; injected
failures (vlb_debug_inject) fail a call before it launches anything and
leave the object usable; server requests can be cancelled, waited for with
a deadline, and a failed step fails only its own requests; destroying a
server with queued requests and uncollected results returns promptly.
Every kernel that runs here ends normally: the conditions are made on the
host.
"""
import threading
import time

import numpy as np
import pytest

import vlb

pytestmark = pytest.mark.gpu


def _params(n=40, rays=512, seed=1, rps=256, precision=vlb.PREC_FP32):
    rng = np.random.default_rng(seed)
    probes = rng.uniform((0.3, 0.3, 0.3), (9.7, 3.7, 9.7), size=(n, 3)).astype(np.float32)
    return vlb.Params(probes=probes, rays=rays, sh_order=2, shade_mode=vlb.SHADE_LAMBERT,
                      lights=[(vlb.LIGHT_DIRECTIONAL, (0.3, -0.9, 0.2), (2.0, 2.0, 2.0))],
                      shadow_bias=1e-3, rays_per_slice=rps, precision=precision)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def room(ctx):
    s = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    sid = ctx.upload_scene_dict(s)
    yield sid
    ctx.release_scene(sid)


def _release_refused(ctx, sid):
    with pytest.raises(vlb.VlbError) as e:
        ctx.release_scene(sid)
    assert e.value.status == vlb.ERR_INVALID
    assert "in use" in str(e.value)


def test_scene_release_refused_while_in_use(ctx):
    s = vlb.generate_scene(vlb.SCENE_ROOM, 5)
    sid = ctx.upload_scene_dict(s)
    plan = ctx.plan(sid, _params())
    _release_refused(ctx, sid)
    plan.destroy()
    b = ctx.batch([(sid, _params()), (sid, _params(seed=2))])
    _release_refused(ctx, sid)
    b.destroy()
    srv = vlb.Server(ctx)
    try:
        srv.hold(True)
        rid = srv.submit(sid, _params())
        _release_refused(ctx, sid)           # queued request holds the scene
        srv.cancel(rid)
        with pytest.raises(vlb.VlbError) as e:
            srv.wait(rid)
        assert e.value.status == vlb.ERR_CANCELLED
        rid = srv.submit(sid, _params())
        srv.hold(False)
        got, info = srv.wait(rid)
        assert info["status"] == 0
    finally:
        srv.destroy()
    ctx.release_scene(sid)                   # nothing uses it any more


def test_injected_plan_and_bake_failures(ctx, room):
    p = _params()
    plan = ctx.plan(room, p)
    try:
        plan.execute()
        ref = plan.download()
        ctx.debug_inject(vlb.INJECT_PLAN_EXECUTE, 1)
        with pytest.raises(vlb.VlbError) as e:
            plan.execute()
        assert e.value.status == vlb.ERR_INJECTED
        plan.execute()                       # one failure only: the plan still works
        assert np.array_equal(_bits(plan.download()), _bits(ref))
        ctx.debug_inject(vlb.INJECT_PLAN_EXECUTE, 2)
        for _ in range(2):
            with pytest.raises(vlb.VlbError) as e:
                ctx.bake(room, p)
            assert e.value.status == vlb.ERR_INJECTED
        assert np.array_equal(_bits(ctx.bake(room, p)), _bits(ref))
        ctx.debug_inject(vlb.INJECT_PLAN_EXECUTE, 5)
        ctx.debug_inject(vlb.INJECT_PLAN_EXECUTE, 0)   # disarmed
        plan.execute()
    finally:
        plan.destroy()
    with pytest.raises(vlb.VlbError):
        ctx.debug_inject(vlb.INJECT_SERVER_STEP + 1, 1)


def test_injected_batch_failure(ctx, room):
    b = ctx.batch([(room, _params()), (room, _params(seed=3, rays=300))])
    try:
        b.execute()
        ref = [b.download(0), b.download(1)]
        ctx.debug_inject(vlb.INJECT_BATCH_EXECUTE, 1)
        with pytest.raises(vlb.VlbError) as e:
            b.execute()
        assert e.value.status == vlb.ERR_INJECTED
        b.use_graph(True)
        b.execute()
        assert np.array_equal(_bits(b.download(0)), _bits(ref[0]))
        assert np.array_equal(_bits(b.download(1)), _bits(ref[1]))
    finally:
        b.destroy()


def test_injected_server_step_fails_only_its_requests(ctx, room):
    srv = vlb.Server(ctx, step_rays=1 << 16)
    try:
        srv.hold(True)
        a = srv.submit(room, _params(seed=4))
        b = srv.submit(room, _params(seed=5))
        ctx.debug_inject(vlb.INJECT_SERVER_STEP, 1)
        srv.hold(False)
        for rid in (a, b):                   # admitted together, failed together
            with pytest.raises(vlb.VlbError) as e:
                srv.wait(rid)
            assert e.value.status == vlb.ERR_INJECTED
        p = _params(seed=6)
        got, info = srv.wait(srv.submit(room, p))   # the server keeps serving
        assert info["status"] == 0
        assert np.array_equal(_bits(got), _bits(ctx.bake(room, p)))
        assert srv.stats()["status"] == 0
    finally:
        srv.destroy()


def test_cancel_queued_and_running_requests(ctx, room):
    srv = vlb.Server(ctx, step_rays=1 << 16)
    try:
        srv.hold(True)
        q = srv.submit(room, _params(seed=7))
        srv.cancel(q)                        # still queued: dropped at once
        with pytest.raises(vlb.VlbError) as e:
            srv.wait(q, timeout_ms=0)
        assert e.value.status == vlb.ERR_CANCELLED
        srv.hold(False)
        # 4096 probes x 4096 rays in steps of 65536 rays: hundreds of steps
        big = srv.submit(room, _params(n=4096, rays=4096, rps=4096, seed=8))
        time.sleep(0.005)
        srv.cancel(big)
        with pytest.raises(vlb.VlbError) as e:
            srv.wait(big)
        assert e.value.status == vlb.ERR_CANCELLED
        with pytest.raises(vlb.VlbError) as e:
            srv.cancel(big)                  # forgotten after its wait
        assert e.value.status == vlb.ERR_NOT_FOUND
        p = _params(seed=9)
        got, info = srv.wait(srv.submit(room, p))
        assert info["status"] == 0
        assert np.array_equal(_bits(got), _bits(ctx.bake(room, p)))
        st = srv.stats()
        assert st["completed"] == 1 and st["status"] == 0
    finally:
        srv.destroy()


def test_wait_with_deadline(ctx, room):
    srv = vlb.Server(ctx)
    try:
        srv.hold(True)
        rid = srv.submit(room, _params(seed=10))
        t0 = time.perf_counter()
        with pytest.raises(vlb.VlbError) as e:
            srv.wait(rid, timeout_ms=30.0)
        el = time.perf_counter() - t0
        assert e.value.status == vlb.ERR_TIMEOUT
        assert 0.025 <= el < 5.0
        with pytest.raises(vlb.VlbError) as e:
            srv.wait(rid, timeout_ms=0.0)
        assert e.value.status == vlb.ERR_TIMEOUT
        srv.hold(False)
        got, info = srv.wait(rid, timeout_ms=60000.0)   # the request stayed
        assert info["status"] == 0 and np.isfinite(got).all()
    finally:
        srv.destroy()


def test_destroy_with_queued_requests_and_uncollected_results(ctx, room):
    """A result arena with room for one result: finished results are not
    collected, so the next requests cannot be admitted; destroy cancels them
    and returns."""
    p = _params(n=200, rays=256, seed=11)              # 200 x 9 x 3 x 4 B = 21.6 KB
    srv = vlb.Server(ctx, result_bytes=32 << 10)
    first = srv.submit(room, p)
    t0 = time.perf_counter()
    while not srv.poll(first) and time.perf_counter() - t0 < 20.0:
        time.sleep(0.001)
    assert srv.poll(first)
    for s in range(4):
        srv.submit(room, _params(n=200, rays=256, seed=12 + s))
    done = threading.Event()

    def destroy():
        srv.destroy()
        done.set()

    th = threading.Thread(target=destroy, daemon=True)
    th.start()
    th.join(60.0)
    assert done.is_set(), "vlb_server_destroy did not return"
