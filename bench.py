#!/usr/bin/env python3
"""Benchmark of the probe-baking hot path on one MI355X.

Default workload (the batched form, the configuration the metric and its
target are stated on): 64 bake jobs over four resident scenes (procedural
rooms of 11k and 99k triangles, fBm terrains of 0.5M and 1.15M), probes per
job log-uniform in [64, 32768] in a random region, rays per probe from
{1024, 4096, 16384} (Fibonacci), one directional light + procedural sky,
one shadow ray per lit hit, L2 SH, BF16 MFMA projection (split operands,
FP32 accumulation, FP64 merge).  All jobs run as one batch: one persistent
launch per kernel and round, the rounds captured as one hipGraph, shapes
picked by the batch autotuner.  A step is one execute of the whole batch
with every input resident in HBM; `value` is probes/s over the K timed
steps (device sync on both sides).

The same run also reports: the plain FP32 GPU baker on the same 64 jobs
(one vlb_bake per job, timed warm over whole passes) and the ratio; the
served form at a fixed Poisson arrival rate (a ~2M-triangle city block
joins the four scenes; each request moves the sun or adds a point light and
rebakes 256-65536 probes at 4096 rays) with its p50/p99 latency against the
plain per-request baker; SH error against the FP64 oracle on seeded probes
of every job; primary-hit parity against the oracle at the workload's
sizes; the trace kernel's roofline; and the CPU baseline (the FP32 C
oracle built -march=native on this host, all usable cores).

    python bench.py [--steps K] [--warmup W]
                    [--config c4|c2|c1|c3|c5|copy|sweep|bounce|reference]
                    [--precision ...] [--variant V] [--no-cpu-baseline]

--no-cpu-baseline skips the CPU leg (the oracle's baseline timing and the
SH error / hit checks against the oracle).
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vulkan-light-bakery_amd"))

METRIC = "probes/sec & Grays/s at ≤1e-3 rel SH err vs FP64; p50/p99 rebake latency"
HBM_PEAK_GBS = 8000.0
TRI_BYTES = 48.0     # bytes per triangle test (3 x float4)


def is_fused(variant):
    """trace_variant 1..12 and 25..36 (far-first shadow queries) run the fused
    trace+shade kernel, 13..24 the wavefront kernels."""
    return not 13 <= variant <= 24


def node_bytes(variant):
    """Bytes fetched per BVH4 node visit: the fused kernels read float nodes
    (7 x 16 B in use), the wavefront kernels quantized ones (4 x 16 B)."""
    return 112.0 if is_fused(variant) else 64.0


PREC = {"fp32": 0, "bf16": 1, "fp16": 2, "bf16_fast": 3, "fp16_fast": 4}
SCHED = {"srpt": 0, "share": 1, "fifo": 2}


def rel_err(a, b):
    import numpy as np

    num = np.sqrt(((np.asarray(a, np.float64) - b) ** 2).sum(axis=1))
    den = np.sqrt((np.asarray(b, np.float64) ** 2).sum(axis=1))
    return num / np.maximum(den, 1e-30)


def load_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_threads():
    """Threads of the CPU baseline: every core this process may run on (its
    affinity mask), capped by a cgroup CPU quota when one is set; and what
    the host reports, for the record."""
    n = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(math.ceil(int(q) / int(per))))
    except (OSError, ValueError):
        pass
    used = min(n, quota) if quota else n
    return used, {"nproc": os.cpu_count(), "affinity": n, "cgroup_quota": quota}


CPU_THREADS, CPU_INFO = cpu_threads()
_ORC = {}


def oracle_leg(args):
    """The CPU leg of a run: the CPU oracle (tests/orc.py over oracle/),
    used for the CPU baseline timing and, as the checker, for the SH error
    and hit parity against FP64 / FP32 -- never for what is measured on the
    GPU.  None with --no-cpu-baseline, which skips both.  The oracle's
    sources are built here for this host first (-O3 -march=native
    -ffp-contract=off, results unchanged); if that fails the portable build
    is used and the line says so.  This is the only place bench.py loads the
    oracle."""
    if args.no_cpu_baseline:
        return None
    if "orc" not in _ORC:
        tmp = tempfile.mkdtemp(prefix="vlb_orc_native_")
        r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "native",
                            f"NATIVE_DIR={tmp}"], capture_output=True, text=True)
        lib = os.path.join(tmp, "liborc.so")
        if r.returncode == 0 and os.path.exists(lib):
            os.environ["VLB_ORC_LIB"] = lib
            _ORC["build"] = "gcc -O3 -march=native -ffp-contract=off, built on this host"
        else:
            _ORC["build"] = "gcc -O3 -march=x86-64-v3 (native build failed: " + r.stderr[-200:] + ")"
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import orc

        _ORC["orc"] = orc
    return _ORC["orc"]


def cpu_line(value, unit, sample):
    return {"value": value, "unit": unit, "cores": CPU_THREADS, "kind": "port", "sample": sample,
            "host": dict(CPU_INFO), "build": _ORC.get("build")}


def trace_roofline(kernel, variant, cnt_nodes, cnt_tris, launch_s, launches, profile):
    """Roofline of a trace stage.  `achieved` is the HBM bytes per launch
    measured by rocprofv3 (FETCH_SIZE x2 + WRITE_SIZE, committed under
    profiles/ for this kernel and workload) over the live event-timed launch
    duration, so frac <= 1 is a real HBM fraction.  `alg_rate_gbs` is the
    algorithmic node + triangle byte rate (112 or 64 B per BVH4 node visit +
    48 B per triangle test, counted by the kernels' counting pass): those
    bytes are mostly served by L1/L2/MALL, so it can exceed the HBM peak.
    `issue` (from the profile's SQ counters) is the bound that does hold:
    the share of VALU issue slots used and the lanes busy per instruction."""
    alg = (node_bytes(variant) * cnt_nodes + TRI_BYTES * cnt_tris) / launches
    prof = None
    if profile:
        prof = profile.get("variants", {}).get(str(variant))
    traffic = prof["hbm_bytes_per_launch"] if prof else None
    achieved = traffic / launch_s / 1e9 if traffic else None
    return {"bound": "hbm", "achieved": None if achieved is None else round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "kernel": kernel, "launch_ms": round(launch_s * 1e3, 4),
            "launches_per_step": launches, "alg_bytes_per_launch": round(alg, 1),
            "alg_rate_gbs": round(alg / launch_s / 1e9, 1),
            "issue": prof.get("issue") if prof else None,
            "profile": (profile or {}).get("file") if prof else None}


def _cpu_seconds_per_probe(vlb, oscene, prm, nprobe, min_seconds, seed=0):
    """Wall seconds per probe of the FP32 CPU oracle (CPU_THREADS threads) on
    a seeded subset of the job's probes (per-probe cost is uniform in the
    ray count), baked repeatedly until min_seconds have passed; and the
    number of probes baked."""
    import numpy as np

    P = prm.probes.shape[0]
    rng = np.random.default_rng(seed)
    sub = np.sort(rng.choice(P, size=min(nprobe, P), replace=False))
    q = vlb.Params(**{**prm.__dict__, "probes": prm.probes[sub].copy(), "_keep": []})
    q.precision = 0
    reps, t0 = 0, time.perf_counter()
    while True:
        oscene.bake(q, mode64=False, nthreads=CPU_THREADS)
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds:
            break
    return el / (reps * sub.size), int(reps * sub.size)


def _err_sample(vlb, oracle_of, items, nprobe=8, seed=0):
    """Max and p99 rel. SH error vs the FP64 oracle over nprobe seeded probes
    of each (scene index, params, coefficients) item."""
    import numpy as np

    rng = np.random.default_rng(seed)
    errs = []
    for si, prm, got in items:
        P = prm.probes.shape[0]
        sub = np.sort(rng.choice(P, size=min(nprobe, P), replace=False))
        q = vlb.Params(**{**prm.__dict__, "probes": prm.probes[sub].copy(), "_keep": []})
        q.precision = 0
        ref = oracle_of(si).bake(q, mode64=True, nthreads=CPU_THREADS)["coeffs"]
        errs.append(rel_err(got[sub], ref).ravel())
    allv = np.concatenate(errs)
    return {"max": float(allv.max()), "p99": float(np.percentile(allv, 99)),
            "probes": int(allv.size // 3), "items": len(items)}


def _hit_parity(ctx, vlb, host, sids, oracle_of, picks, variant):
    """Primary-hit parity at the workload's sizes: for each (scene index,
    params) pick, the product's closest-hit prim id per ray (vlb_bake_radiance,
    trace variant `variant`) against the FP32 oracle's, over every ray of the
    pick's probes.  Grazing rays (|d.n_g| < 1e-4, or a barycentric < 1e-6 at
    the oracle's hit) are counted separately."""
    import numpy as np

    tot = same = graze = graze_same = 0
    for si, prm in picks:
        p = vlb.Params(**{**prm.__dict__, "_keep": []})
        p.precision = 0
        p.trace_variant = variant
        _, ghit = ctx.bake_radiance(sids[si], p, want_radiance=False, want_hits=True)
        osc = oracle_of(si)
        chit = osc.bake(p, mode64=False, nthreads=CPU_THREADS, want_hits=True)["hits"]
        dirs, _ = vlb.directions(vlb.DIRS_FIBONACCI, p.rays)
        P, R = chit.shape
        org = np.repeat(p.probes, R, axis=0)
        d = np.tile(dirs, (P, 1))
        prim, t, u, v = osc.trace_closest(org, d, p.t_min, p.t_max)
        assert np.array_equal(prim, chit.ravel())
        xyz, idx = host[si]["xyz"].astype(np.float64), host[si]["idx"]
        h = prim >= 0
        tri = idx[prim[h]]
        n = np.cross(xyz[tri[:, 1]] - xyz[tri[:, 0]], xyz[tri[:, 2]] - xyz[tri[:, 0]])
        n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-300)
        cosang = np.abs((n * d[h]).sum(axis=1))
        bary = np.minimum(np.minimum(u[h], v[h]), 1.0 - u[h] - v[h])
        gz = np.zeros(prim.shape, bool)
        gz[np.flatnonzero(h)] = (cosang < 1e-4) | (bary < 1e-6)
        eq = ghit.ravel() == chit.ravel()
        tot += eq.size
        same += int(eq.sum())
        graze += int(gz.sum())
        graze_same += int(eq[gz].sum())
    ng = tot - graze
    return {"rays": tot, "equal": same, "frac": round(same / tot, 8),
            "non_grazing_frac": round((same - graze_same) / ng, 8) if ng else None,
            "grazing": graze, "grazing_equal": graze_same, "variant": variant}


# --------------------------------------------------------- the served form
def serve(ctx, vlb, W, host, sids, args, oracle_of=None):
    """The served form on resident scenes: args.requests Poisson arrivals at
    args.rate per second through the server (BF16, continuous batching,
    HBM-budget admission), then the same schedule through the plain FP32
    baker, one vlb_bake per request in arrival order.  Returns the result
    dict and (request params, coefficients) for error checks."""
    import threading

    import numpy as np

    n = args.requests
    arrive, reqs, sun0 = W.c5_requests(host, n, args.rate)
    srv = vlb.Server(ctx, precision=PREC["bf16"], state_budget=args.state_budget << 20,
                     step_rays=args.step_rays, policy=SCHED[args.policy], trace_variant=args.variant,
                     age_ms=args.age_ms)
    # warm-up: one small request per scene (direction tables, graph capture)
    warm = []
    for si in range(len(host)):
        w = (si, reqs[0][1][:256], sun0,
             [(vlb.LIGHT_DIRECTIONAL, (-sun0).tolist(), (3.0, 3.0, 3.0))])
        warm.append((si, W.c5_params(w, PREC["bf16"])))
        srv.wait(srv.submit(sids[si], warm[-1][1]))
    ps16 = [W.c5_params(r, PREC["bf16"]) for r in reqs]
    ids = [None] * n

    def feeder():
        t0 = time.perf_counter()
        for i in range(n):
            dt = t0 + arrive[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids[i] = srv.submit(sids[reqs[i][0]], ps16[i])

    th = threading.Thread(target=feeder)
    th.start()
    lat, outs, infos_r = [], [], []
    for i in range(n):
        while ids[i] is None:
            time.sleep(0.0005)
        c, inf = srv.wait(ids[i])
        outs.append(c)
        infos_r.append(inf)
        lat.append(inf["done_ms"] - inf["submit_ms"])
    th.join()
    st = srv.stats()
    first = min(i["submit_ms"] for i in infos_r)
    last = max(i["done_ms"] for i in infos_r)
    P_all = sum(r[1].shape[0] for r in reqs)
    srv_span = (last - first) / 1e3
    srv.destroy()

    ps32 = [W.c5_params(r, PREC["fp32"]) for r in reqs]
    plat = []
    q0 = time.perf_counter()
    for i in range(n):
        now = time.perf_counter() - q0
        if now < arrive[i]:
            time.sleep(arrive[i] - now)
        ctx.bake(sids[reqs[i][0]], ps32[i])
        plat.append((time.perf_counter() - q0 - arrive[i]) * 1e3)
    plain_span = time.perf_counter() - q0
    lat = np.asarray(lat)
    plat = np.asarray(plat)
    res = {"requests": n, "rate_per_s": args.rate, "probes": P_all, "rays_per_probe": W.C5_RAYS,
           "probes_per_s": round(P_all / srv_span, 1),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                          "p99": round(float(np.percentile(lat, 99)), 2),
                          "max": round(float(lat.max()), 2)},
           "server": {"steps": st["steps"], "busy_ms": round(st["busy_ms"], 1),
                      "active_peak": st["active_peak"], "state_bytes_peak": st["state_bytes_peak"],
                      "span_s": round(srv_span, 3), "policy": args.policy,
                      "state_budget_mib": args.state_budget, "step_rays": args.step_rays,
                      "age_ms": args.age_ms,
                      "trace_variant": st["trace_variant"]},
           "plain_fp32_per_request": {"probes_per_s": round(P_all / plain_span, 1),
                                      "p50_ms": round(float(np.percentile(plat, 50)), 2),
                                      "p99_ms": round(float(np.percentile(plat, 99)), 2),
                                      "span_s": round(plain_span, 3)}}
    return res, reqs, ps16, ps32, outs, warm


def served_roofline(ctx, sids, reqs, ps16, warm, server):
    """Roofline of the served form's trace work.  The server's steps run
    the multi-job trace kernel on mixed requests; a step's device time is
    event-timed by the server (busy_ms, which also holds the step's small
    projection and merge, so the rates below are lower bounds for the trace
    kernel).  Algorithmic bytes: node visits and triangle tests of every
    ray the server traced (the requests and the warm-up requests), counted
    by one batch counting pass of the same variant -- a ray's counts do not
    depend on how rays are grouped into launches.  `traffic` is the HBM
    bytes per step from the committed rocprofv3 PMC summary of this
    workload."""
    v = server["trace_variant"]
    jobs = [(sids[si], p) for si, p in warm] + [(sids[reqs[i][0]], ps16[i]) for i in range(len(reqs))]
    b = ctx.batch(jobs, trace_variant=v)
    b.count()
    cnt = b.info()
    b.destroy()
    steps = server["steps"]
    step_s = server["busy_ms"] / 1e3 / steps
    alg = (node_bytes(v) * cnt["node_visits"] + TRI_BYTES * cnt["tri_tests"]) / steps
    profile = load_traffic("served_c5_traffic.json")
    prof = (profile or {}).get("variants", {}).get(str(v))
    traffic = prof["hbm_bytes_per_launch"] if prof else None
    achieved = traffic / step_s / 1e9 if traffic else None
    return {"bound": "hbm", "achieved": None if achieved is None else round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "kernel": "k_fused_jobs (server steps)", "variant": v,
            "launch_ms": round(step_s * 1e3, 4), "launch_timing": "server step device time (HIP events)",
            "launches": steps, "alg_bytes_per_launch": round(alg, 1),
            "alg_rate_gbs": round(alg / step_s / 1e9, 1),
            "node_visits": cnt["node_visits"], "tri_tests": cnt["tri_tests"],
            "issue": prof.get("issue") if prof else None,
            "profile": (profile or {}).get("file") if prof else None}


# ----------------------------------------------------------- batched form
def bench_c4(args):
    """The default line: the batched form (see the module docstring)."""
    import numpy as np

    import vlb
    import workloads as W

    ctx = vlb.Context(0)
    host = W.scenes(W.C5_SCENES)          # the city block first, then the batched form's four
    sids = [ctx.upload_scene_dict(s) for s in host]
    infos = [ctx.scene_info(s) for s in sids]
    spec, sun = W.c4_jobs(host[1:])       # scene indices into host[1:]
    jobs16 = [(sids[1 + si], W.sun_params(pr, R, PREC["bf16"], sun)) for si, pr, R in spec]
    P_all = sum(pr.shape[0] for _, pr, _ in spec)
    rays_all = sum(pr.shape[0] * R for _, pr, R in spec)

    if args.variant:
        b = ctx.batch(jobs16, trace_variant=args.variant)
        variant, chunk = args.variant, 256
    else:
        b = ctx.batch(jobs16)
        # 8 timed runs per candidate on the library's ~16M-ray sample (about
        # 2 ms each): variants 12 and 36 are 1% apart on these jobs, and 2
        # runs let run-to-run noise pick either
        variant, chunk = b.autotune(8)
    b.count()
    cnt = b.info()
    # timed region 1 (`value`): the batch as one captured hipGraph per step
    b.use_graph(True)
    for _ in range(max(1, args.warmup)):
        b.execute()
    b.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.execute()
    b.sync()
    t1 = time.perf_counter()
    el = (t1 - t0) / args.steps
    graph_ms = b.info()["total_ms"]
    got = [b.download(j) for j in range(len(jobs16))]
    # timed region 2 (kernel durations): the same steps launched eagerly with
    # hipEvents around every kernel on the library stream (events inside a
    # captured graph cannot be timed); same kernels, same work
    b.use_graph(False)
    b.set_timing(True)
    b.execute()
    b.sync()
    b.set_timing(True)                    # the means cover the timed executes only
    e0 = time.perf_counter()
    for _ in range(args.steps):
        b.execute()
    b.sync()
    el_eager = (time.perf_counter() - e0) / args.steps
    info = b.info()
    b.destroy()

    # the plain FP32 GPU baker: one synchronous vlb_bake per job, in order;
    # one untimed pass, then whole passes timed
    jobs32 = [(sids[1 + si], W.sun_params(pr, R, PREC["fp32"], sun)) for si, pr, R in spec]
    passes = args.plain_passes
    plain_s = None
    if passes > 0:
        for sid, prm in jobs32:
            ctx.bake(sid, prm)
        p0 = time.perf_counter()
        for _ in range(passes):
            for sid, prm in jobs32:
                ctx.bake(sid, prm)
        plain_s = (time.perf_counter() - p0) / passes

    # the served form, same context and resident scenes
    served = None
    if args.requests > 0:
        served, *_ = serve(ctx, vlb, W, host, sids, args)

    orc = oracle_leg(args)
    err = hits = cpu = None
    if orc is not None:
        oscenes = {}

        def oracle_of(i):          # index into host
            if i not in oscenes:
                oscenes[i] = orc.Scene.from_dict(host[i])
            return oscenes[i]

        err = _err_sample(vlb, oracle_of, [(1 + spec[j][0], jobs16[j][1], got[j]) for j in range(len(spec))],
                          nprobe=3)
        # hit parity: 2 seeded probes of the first job on each scene
        rng = np.random.default_rng(8)
        picks, seen = [], set()
        for j, (si, pr, R) in enumerate(spec):
            if si in seen:
                continue
            seen.add(si)
            sub = np.sort(rng.choice(pr.shape[0], size=2, replace=False))
            picks.append((1 + si, W.sun_params(pr[sub].copy(), R, PREC["fp32"], sun)))
        hits = _hit_parity(ctx, vlb, host, sids, oracle_of, picks, variant)
        # CPU baseline: every job on the FP32 oracle, 16 seeded probes each,
        # scaled to the job's probe count (per-probe cost is uniform in a job)
        est, t_cpu = 0.0, 0.0
        for j, (si, pr, R) in enumerate(spec):
            spp, n = _cpu_seconds_per_probe(vlb, oracle_of(1 + si), jobs32[j][1], 16, args.cpu_seconds / 64,
                                            seed=j)
            est += spp * pr.shape[0]
            t_cpu += spp * n
        cpu = cpu_line(round(P_all / est, 2), "probes/s",
                       f"FP32 C oracle, the same 64 jobs one after another, 16 seeded probes of each baked "
                       f"repeatedly for >= {args.cpu_seconds / 64:.2f} s and scaled to the job's probe count, "
                       f"{CPU_THREADS} threads, {t_cpu:.1f} s")
    ctx.close()

    rays = cnt["primary_rays"] + cnt["shadow_rays"]
    rounds = info["rounds"]
    roof = trace_roofline("k_fused_jobs" if is_fused(variant) else "k_rays_jobs + k_shade_jobs", variant,
                          cnt["node_visits"], cnt["tri_tests"], info["trace_ms"] / 1e3 / rounds, rounds,
                          load_traffic("batch_c4_traffic.json"))
    return {
        "metric": METRIC, "value": round(P_all / el, 1), "unit": "probes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": max(1, args.warmup), "ms_per_step": round(el * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (4 seeded rooms/terrains, 64 seeded jobs, Fibonacci rays; city block added "
                "for the served form)",
        "config": {"workload": "c4 batched form", "jobs": len(spec), "probes": P_all,
                   "primary_rays": rays_all,
                   "scenes": [{"scene": W.C4_SCENES[i][0], "triangles": infos[1 + i]["num_tris"],
                               "bvh_depth": infos[1 + i]["depth"]} for i in range(4)],
                   "rays_per_probe": list(W.C4_RAYS), "sh_order": 2,
                   "precision": "bf16 (split operands, FP32 accumulate, FP64 merge)",
                   "rounds": rounds, "trace_variant": variant, "chunk": chunk, "graph": True},
        "grays_per_s": round(rays / el / 1e9, 3),
        "probe_rays_per_s": round(rays_all / el, 1),
        "shadow_rays_per_step": cnt["shadow_rays"],
        "per_ray": {"inner_nodes": round(cnt["node_visits"] / rays, 2),
                    "triangles": round(cnt["tri_tests"] / rays, 2)},
        "kernel_ms": {"trace": round(info["trace_ms"], 3), "project": round(info["project_ms"], 3),
                      "merge": round(info["merge_ms"], 3), "timed_runs": info["kernel_timed_runs"],
                      "eager_step_ms": round(el_eager * 1e3, 2), "graph_device_ms": round(graph_ms, 3),
                      "note": "per-kernel means over the timed eager steps (sum over the rounds)"},
        "sh_rel_err_vs_fp64": err,
        "hit_parity": hits,
        "roofline": roof,
        "plain_fp32_baker": None if plain_s is None else
        {"probes_per_s": round(P_all / plain_s, 1), "seconds_per_pass": round(plain_s, 4),
         "passes": passes, "calls_per_pass": len(jobs32), "warm": "one untimed pass first"},
        "speedup_vs_plain_fp32": None if plain_s is None else round(plain_s / el, 3),
        "served": served,
        "cpu_baseline": cpu,
        "device_bytes": info["device_bytes"],
    }


def bench_c5(args):
    """The served form on its own line: the serving leg of the default run,
    with the SH error of the served requests and the CPU baseline on the
    request mix."""
    import numpy as np

    import vlb
    import workloads as W

    ctx = vlb.Context(0)
    host = W.scenes(W.C5_SCENES)
    sids = [ctx.upload_scene_dict(s) for s in host]
    infos = [ctx.scene_info(s) for s in sids]
    res, reqs, ps16, ps32, outs, warm = serve(ctx, vlb, W, host, sids, args)
    roof = served_roofline(ctx, sids, reqs, ps16, warm, res["server"])
    orc = oracle_leg(args)
    err = cpu = hits = None
    if orc is not None:
        oscenes = {}

        def oracle_of(i):
            if i not in oscenes:
                oscenes[i] = orc.Scene.from_dict(host[i])
            return oscenes[i]

        n = len(reqs)
        err = _err_sample(vlb, oracle_of, [(reqs[i][0], ps16[i], outs[i]) for i in range(0, n, max(1, n // 20))])
        # hit parity: 2 seeded probes of the first request on each scene
        rng = np.random.default_rng(9)
        picks, seen = [], set()
        for i, (si, pr, _, _) in enumerate(reqs):
            if si in seen:
                continue
            seen.add(si)
            q = vlb.Params(**{**ps32[i].__dict__, "_keep": []})
            q.probes = pr[np.sort(rng.choice(pr.shape[0], size=2, replace=False))].copy()
            picks.append((si, q))
        # the variant the server ran (its default is 35)
        hits = _hit_parity(ctx, vlb, host, sids, oracle_of, picks, res["server"]["trace_variant"])
        spp_scene, t_cpu = {}, 0.0
        for i in range(n):
            si = reqs[i][0]
            if si in spp_scene:
                continue
            spp, m = _cpu_seconds_per_probe(vlb, oracle_of(si), ps32[i], 16, args.cpu_seconds / len(host), seed=i)
            spp_scene[si] = spp
            t_cpu += spp * m
        est = sum(spp_scene[r[0]] * r[1].shape[0] for r in reqs)
        cpu = cpu_line(round(res["probes"] / est, 2), "probes/s",
                       f"FP32 C oracle, 16 probes of one request per scene ({len(spp_scene)} scenes) baked "
                       f"repeatedly and scaled to all {n} requests, {CPU_THREADS} threads, {t_cpu:.1f} s")
    ctx.close()
    lat = res["latency_ms"]
    return {"metric": METRIC, "value": res["probes_per_s"], "unit": "probes/s", "n_gpus": 1,
            "steps": res["requests"], "warmup": len(host), "ms_per_step": lat["p50"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
            "data": "synthetic (seeded city + 4 rooms/terrains, Poisson arrivals, Fibonacci rays)",
            "config": {"workload": "c5 serving form", "requests": res["requests"], "rate_per_s": args.rate,
                       "probes": res["probes"], "rays_per_probe": W.C5_RAYS,
                       "scenes": [{"scene": W.C5_SCENES[i][0], "triangles": infos[i]["num_tris"],
                                   "bvh_depth": infos[i]["depth"]} for i in range(len(host))]},
            "latency_ms": lat, "server": res["server"],
            "plain_fp32_per_request": res["plain_fp32_per_request"],
            "sh_rel_err_vs_fp64": err, "hit_parity": hits, "roofline": roof, "cpu_baseline": cpu}


# -------------------------------------------------------- plain single bake
def bench_c2(args):
    """The plain single-scene bake: procedural room 10x4x10 m (64 boxes + 8
    icospheres, 11,018 triangles, seed 2), 16x8x16 = 2048 probes, 4096
    Fibonacci rays per probe, one directional light + procedural sky, L2 SH,
    FP32 projection.  A step is one full vlb_plan_execute (trace stage, MFMA
    projection, FP64 merge)."""
    import numpy as np

    import vlb

    scene = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    params = vlb.config_c2_params(scene, seed=2, precision=PREC[args.precision])
    params.trace_variant = args.variant
    P = params.probes.shape[0]
    R = params.rays
    ctx = vlb.Context(0)
    sid = ctx.upload_scene_dict(scene)
    info = ctx.scene_info(sid)
    plan = ctx.plan(sid, params)
    # outside the timed region: the autotuner picks the trace-kernel variant
    # on this workload, then one counting pass gives node visits, triangle
    # tests and traced shadow rays of exactly this bake
    variant = plan.autotune(5) if args.variant == 0 else args.variant
    plan.count()
    cnt = plan.stats()
    for _ in range(args.warmup):
        plan.execute()
    plan.sync()
    plan.set_timing(True)
    plan.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute()
    plan.sync()
    t1 = time.perf_counter()
    st = plan.stats()
    coeffs = plan.download()
    elapsed = t1 - t0
    probes_per_s = P * args.steps / elapsed
    rays_per_step = cnt["primary_rays"] + cnt["shadow_rays"]
    grays = rays_per_step * args.steps / elapsed / 1e9
    roof = trace_roofline("k_trace" if is_fused(variant) else "k_rays + k_shade", variant, cnt["node_visits"],
                          cnt["tri_tests"], st["trace_ms"] / 1e3, 1, load_traffic("bake_c2_traffic.json"))

    orc = oracle_leg(args)
    err_line, cpu, hits = None, None, None
    if orc is not None:
        rng = np.random.default_rng(0)
        sub = np.sort(rng.choice(P, size=64, replace=False))
        psub = vlb.config_c2_params(scene, seed=2)
        psub.probes = params.probes[sub].copy()
        oscene = orc.Scene.from_dict(scene)
        ref = oscene.bake(psub, mode64=True, nthreads=CPU_THREADS)["coeffs"]
        err = rel_err(coeffs[sub], ref)
        err_line = {"max": float(err.max()), "p99": float(np.percentile(err, 99)), "probes": len(sub)}
        # hit parity over every primary ray of the bake
        hits = _hit_parity(ctx, vlb, [scene], [sid], lambda i: oscene,
                           [(0, vlb.config_c2_params(scene, seed=2))], variant)
        # the FP32 C oracle on the same scene, probes and rays: whole c2
        # bakes repeated until >= cpu_seconds of wall time on the host's cores
        pc = vlb.config_c2_params(scene, seed=2)
        reps = 0
        c0 = time.perf_counter()
        while True:
            oscene.bake(pc, mode64=False, nthreads=CPU_THREADS)
            reps += 1
            if time.perf_counter() - c0 >= args.cpu_seconds:
                break
        c1 = time.perf_counter()
        cpu = cpu_line(round(reps * P / (c1 - c0), 2), "probes/s",
                       f"FP32 C oracle: {reps} full bakes of the same {P} probes x {R} rays "
                       f"(same scene, BVH-independent results), {CPU_THREADS} threads, {c1 - c0:.1f} s")
    line = {
        "metric": METRIC,
        "value": round(probes_per_s, 2),
        "unit": "probes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else args.precision,
        "data": "synthetic (seeded procedural room, Fibonacci rays)",
        "config": {"workload": "c2 plain single-scene bake", "scene": "room seed 2",
                   "triangles": info["num_tris"], "bvh_depth": info["depth"], "probes": P,
                   "rays_per_probe": R, "sh_order": 2, "lights": "1 directional + sky",
                   "precision": args.precision},
        "grays_per_s": round(grays, 3),
        "probe_rays_per_s": round(P * R * args.steps / elapsed, 1),
        "shadow_rays_per_step": cnt["shadow_rays"],
        "per_ray": {"inner_nodes": round(cnt["node_visits"] / rays_per_step, 2),
                    "triangles": round(cnt["tri_tests"] / rays_per_step, 2)},
        "trace_variant": variant,
        "kernel_ms": {"trace": round(st["trace_ms"], 4), "project": round(st["project_ms"], 4),
                      "merge": round(st["merge_ms"], 4), "timed_runs": st["timed_runs"]},
        "sh_rel_err_vs_fp64": err_line,
        "hit_parity": hits,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    plan.destroy()
    ctx.release_scene(sid)
    ctx.close()
    return line


def bench_c1(args):
    """Skybox SH: 1024 procedural 6x128x128 RGBA32F cubemaps -> L2 per step."""
    import numpy as np

    import vlb

    N, B = 128, 1024
    rng = np.random.default_rng(1)
    base = np.stack([vlb.procedural_cubemap(N, vlb.sun_direction(rng)) for _ in range(8)])
    ctx = vlb.Context(0)
    nbytes = B * 6 * N * N * 16
    d_in = ctx.alloc(nbytes)
    d_out = ctx.alloc(B * 27 * 4)
    for i in range(B // 8):
        ctx.upload(d_in + i * base.nbytes, base)
    for _ in range(args.warmup):
        ctx.skybox_sh_device(d_in, N, B, 2, d_out)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.skybox_sh_device(d_in, N, B, 2, d_out)
    ctx.sync()
    t1 = time.perf_counter()
    # kernel duration: events of the last step, then a few event-timed steps
    proj = [ctx.skybox_last_ms()[0]]
    for _ in range(4):
        ctx.skybox_sh_device(d_in, N, B, 2, d_out)
        proj.append(ctx.skybox_last_ms()[0])
    out = np.zeros((B, 9, 3), np.float32)
    ctx.download(d_out, out)
    # batch = 1: one skybox per call, synchronised (launch-latency bound)
    lat, k1, m1 = [], [], []
    for i in range(50):
        c0 = time.perf_counter()
        ctx.skybox_sh_device(d_in, N, 1, 2, d_out)
        ctx.sync()
        lat.append(time.perf_counter() - c0)
        pm = ctx.skybox_last_ms()
        k1.append(pm[0])
        m1.append(pm[1])
    orc = oracle_leg(args)
    err, cpu = None, None
    if orc is not None:
        err = rel_err(out[0], orc.skybox_sh(base[0], 9)).max()
        # CPU leg: the oracle's projection (FP64) of the same cubemaps, one
        # cubemap per thread at a time on every usable core (the ctypes call
        # releases the GIL)
        import threading

        counts = [0] * CPU_THREADS
        c0 = time.perf_counter()

        def worker(k):
            while time.perf_counter() - c0 < args.cpu_seconds / 2:
                orc.skybox_sh(base[(counts[k] + k) % 8], 9)
                counts[k] += 1

        ths = [threading.Thread(target=worker, args=(k,)) for k in range(CPU_THREADS)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        c1 = time.perf_counter()
        n_cpu = sum(counts)
        cpu = cpu_line(round(n_cpu / (c1 - c0), 1), "skyboxes/s",
                       f"C oracle (FP64, exact solid angles), {n_cpu} of the same 6x128x128 cubemaps, "
                       f"{CPU_THREADS} threads each projecting whole cubemaps, {c1 - c0:.1f} s")
    kms = float(np.mean(proj))
    traffic = load_traffic("skybox_c1_traffic.json")
    achieved = (nbytes + B * 27 * 4) / (kms / 1e3) / 1e9
    # attainable read bandwidth on this device: a read-only streaming pass
    # over the same cubemap buffer, at the skybox's own grid (3 workgroups
    # per cubemap) and at 16 workgroups per CU; the faster one is the ceiling
    ceil = {}
    for grid in (3 * B, 0):
        ctx.device_read(d_in, nbytes, grid)
        ms = [ctx.device_read(d_in, nbytes, grid) for _ in range(5)]
        ceil[grid] = nbytes / (float(np.mean(ms)) / 1e3) / 1e9
    ceil_grid = max(ceil, key=ceil.get)
    line = {"metric": "skyboxes/s (6x128x128 RGBA32F -> L2)", "value": round(B * args.steps / (t1 - t0), 1),
            "unit": "skyboxes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * (t1 - t0) / args.steps, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "c1 skybox SH", "batch": B, "face": N},
            "sh_rel_err_vs_fp64": None if err is None else float(err),
            "batch1": {"wall_us_per_skybox": round(1e6 * float(np.median(lat)), 1),
                       "kernel_us": round(1e3 * float(np.median(k1)), 2),
                       "merge_us": round(1e3 * float(np.median(m1)), 2), "calls": len(lat)},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": traffic.get("hbm_bytes_per_launch") if traffic else None,
                         "attainable_read_gbs": round(ceil[ceil_grid], 1),
                         "attainable_grid": ceil_grid or "16 per CU",
                         "frac_of_attainable": round(achieved / ceil[ceil_grid], 4)},
            "cpu_baseline": cpu}
    ctx.close()
    return line


def sweep_c2(args):
    """Trace-kernel variants on the c2 bake: one JSON line per variant."""
    import vlb

    scene = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    ctx = vlb.Context(0)
    sid = ctx.upload_scene_dict(scene)
    rows = []
    for v in range(1, 37):
        params = vlb.config_c2_params(scene, seed=2)
        params.trace_variant = v
        plan = ctx.plan(sid, params)
        for _ in range(3):
            plan.execute()
        plan.set_timing(True)
        for _ in range(args.steps):
            plan.execute()
        plan.sync()
        st = plan.stats()
        kind = "fused" if v <= 12 else ("wave" if v <= 24 else "fused far-first")
        b = (v - 1) % 12
        rows.append({"variant": v, "kind": kind, "map": b % 4,
                     "occ_or_refill": ([16, 32, 8][b // 4] if kind == "wave" else [0, 6, 8][b // 4]),
                     "trace_ms": round(st["trace_ms"], 4), "project_ms": round(st["project_ms"], 4),
                     "total_ms": round(st["total_ms"], 4), "grid": st["grid"]})
        print(json.dumps(rows[-1]), flush=True)
        plan.destroy()
    ctx.close()
    best = min(rows, key=lambda r: r["total_ms"])
    return {"sweep": "c2 trace variants", "best": best}


def bench_c3(args):
    """Precision form: terrain + 50 icospheres (1.154 M tris, seed 3), 32x16x32
    probes x 16384 Fibonacci rays, L3 in FP32 / BF16 / FP16 projection; per-band
    error of every projection form against the FP64 oracle on a probe subset."""
    import numpy as np

    import vlb

    scene = vlb.generate_scene(vlb.SCENE_TERRAIN, 3)
    ctx = vlb.Context(0)
    sid = ctx.upload_scene_dict(scene)
    info = ctx.scene_info(sid)
    lo, hi = scene["bounds_min"], scene["bounds_max"]
    probes, _ = vlb.probe_grid([4.0, lo[1] + 1.0, 4.0], [252.0, hi[1] + 20.0, 252.0], [32, 16, 32])
    rng = np.random.default_rng(3)
    sun = vlb.sun_direction(rng)

    def params(prec, order, pr=probes):
        return vlb.Params(probes=pr, rays=16384, sh_order=order, precision=prec,
                          shade_mode=vlb.SHADE_LAMBERT, sun_dir=sun.tolist(),
                          lights=[(vlb.LIGHT_DIRECTIONAL, (-sun).tolist(), (3.0, 3.0, 3.0))],
                          shadow_bias=2e-3)

    runs = {}
    cnt = None
    for name, prec in (("fp32", 0), ("bf16", 1), ("fp16", 2)):
        prm = params(prec, 3)
        prm.trace_variant = args.variant
        plan = ctx.plan(sid, prm)
        variant = plan.autotune(1) if args.variant == 0 else args.variant
        if cnt is None:
            plan.count()
            cnt = plan.stats()
        plan.execute()
        plan.sync()
        plan.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.execute()
        plan.sync()
        t1 = time.perf_counter()
        st = plan.stats()
        runs[name] = {"probes_per_s": round(len(probes) * args.steps / (t1 - t0), 1),
                      "ms_per_step": round(1e3 * (t1 - t0) / args.steps, 2),
                      "trace_ms": round(st["trace_ms"], 3), "project_ms": round(st["project_ms"], 4),
                      "merge_ms": round(st["merge_ms"], 4), "variant": variant}
        plan.destroy()
    # per-band error vs FP64 on a seeded probe subset (the oracle is the checker)
    sub = np.sort(rng.choice(len(probes), size=24, replace=False))
    orc = oracle_leg(args)
    osc = orc.Scene.from_dict(scene) if orc is not None else None
    cpu = None
    if osc is not None:
        # CPU leg: the FP32 oracle on a fixed seeded subset of 256 probes, L3
        spp, n = _cpu_seconds_per_probe(vlb, osc, params(0, 3), 256, args.cpu_seconds / 2, seed=3)
        cpu = cpu_line(round(1.0 / spp, 2), "probes/s",
                       f"FP32 C oracle, a fixed seeded subset of 256 of the 16,384 probes baked repeatedly "
                       f"({n} probe bakes of 16,384 rays, L3, same scene, BVH-independent results), "
                       f"{CPU_THREADS} threads, {spp * n:.1f} s")
    errs = {}
    for order in ((2, 3) if osc is not None else ()):
        ref = osc.bake(params(0, order, probes[sub].copy()), mode64=True,
                       nthreads=CPU_THREADS)["coeffs"]
        K = 9 if order == 2 else 16
        for name, prec in (("fp32", 0), ("bf16", 1), ("fp16", 2), ("bf16_fast", 3), ("fp16_fast", 4)):
            got = ctx.bake(sid, params(prec, order, probes[sub].copy()))
            band = {}
            for l in range(order + 1):
                ks = slice(l * l, (l + 1) * (l + 1))
                band[f"l{l}"] = float(rel_err(got[:, ks], ref[:, ks]).max())
            band["all"] = float(rel_err(got, ref).max())
            errs[f"L{order}_{name}"] = band
    # hit parity over every primary ray of 8 of the seeded probes
    hits = (_hit_parity(ctx, vlb, [scene], [sid], lambda i: osc, [(0, params(0, 2, probes[sub[:8]].copy()))],
                        variant) if osc is not None else None)
    rays = cnt["primary_rays"] + cnt["shadow_rays"]
    tr = runs["bf16"]["trace_ms"] / 1e3
    line = {"metric": METRIC, "value": runs["bf16"]["probes_per_s"], "unit": "probes/s", "n_gpus": 1,
            "steps": args.steps, "warmup": 1, "ms_per_step": runs["bf16"]["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
            "data": "synthetic (seeded terrain + icospheres, Fibonacci rays)",
            "config": {"workload": "c3 precision form", "triangles": info["num_tris"],
                       "bvh_depth": info["depth"], "bvh_build_ms": round(info["build_ms"], 2),
                       "probes": len(probes), "rays_per_probe": 16384, "sh_order": 3},
            "runs": runs, "grays_per_s": round(rays / tr / 1e9, 3),
            "per_ray": {"inner_nodes": round(cnt["node_visits"] / rays, 2),
                        "triangles": round(cnt["tri_tests"] / rays, 2)},
            "sh_rel_err_vs_fp64_by_band": errs or None, "error_probes": len(sub) if errs else 0,
            "hit_parity": hits,
            "cpu_baseline": cpu,
            "roofline": trace_roofline("k_trace" if is_fused(variant) else "k_rays + k_shade", variant,
                                       cnt["node_visits"], cnt["tri_tests"], tr, 1,
                                       load_traffic("bake_c3_traffic.json"))}
    ctx.close()
    return line


def bench_reference(args):
    """The reference's own bake settings (LightBaker's defaults,
    light_baker.cpp:38 and 65, env_map_generator.cpp:24-28, env_map.rchit):
    7x7x7 probes in the reference's BFS order over the scene bounds (origin
    included, scene_manager.cpp:497-507), 3141 x 1000 equirect rays per
    probe, the REFERENCE shading with its shadow ray and sRGB + UNORM8
    conversion, L3 (16 coefficients, the literal constants of sh_common.h),
    FP32, on the c2 room (the reference's glTF scenes are not in its
    snapshot).  A step is one whole bake of 343 probes x 3.141 M rays."""
    import numpy as np

    import vlb

    scene = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    lo = np.minimum(0.0, np.asarray(scene["bounds_min"], np.float64))
    hi = np.maximum(0.0, np.asarray(scene["bounds_max"], np.float64))
    probes, _ = vlb.probe_grid(lo.tolist(), hi.tolist(), [7, 7, 7], vlb.ORDER_REFERENCE_BFS)

    def params(pr, prec):
        return vlb.Params(probes=pr, dir_mode=vlb.DIRS_EQUIRECT, equirect_w=3141, equirect_h=1000,
                          sh_order=3, sh_constants=vlb.SH_LITERAL, precision=prec,
                          shade_mode=vlb.SHADE_REFERENCE, trace_variant=args.variant)

    prm = params(probes, PREC[args.precision])
    P, R = probes.shape[0], 3141 * 1000
    ctx = vlb.Context(0)
    sid = ctx.upload_scene_dict(scene)
    plan = ctx.plan(sid, prm)
    variant = plan.autotune(2) if args.variant == 0 else args.variant
    plan.count()
    cnt = plan.stats()
    for _ in range(max(1, args.warmup)):
        plan.execute()
    plan.sync()
    plan.set_timing(True)
    plan.sync()
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute()
    plan.sync()
    el = (time.perf_counter() - t0) / steps
    st = plan.stats()
    got = plan.download()
    plan.destroy()
    ctx.close()
    # CPU leg: two probes against the FP64 oracle, and the FP32 oracle's rate
    orc = oracle_leg(args)
    err, cpu = None, None
    if orc is not None:
        osc = orc.Scene.from_dict(scene)
        sub = np.array([0, P // 2])
        nth = CPU_THREADS
        ref = osc.bake(params(probes[sub].copy(), PREC["fp32"]), mode64=True, nthreads=nth)["coeffs"]
        e = rel_err(got[sub], ref)
        err = {"max": float(e.max()), "probes": int(sub.size)}
        c0 = time.perf_counter()
        osc.bake(params(probes[sub].copy(), PREC["fp32"]), mode64=False, nthreads=nth)
        spp = (time.perf_counter() - c0) / sub.size
        cpu = cpu_line(round(1.0 / spp, 3), "probes/s",
                       f"FP32 C oracle, {sub.size} of the {P} probes (3.141 M rays each, per-probe cost "
                       f"uniform in rays), {nth} threads, {spp * sub.size:.1f} s")
    rays = cnt["primary_rays"] + cnt["shadow_rays"]
    return {
        "metric": METRIC, "value": round(P / el, 1), "unit": "probes/s", "n_gpus": 1,
        "steps": steps, "warmup": max(1, args.warmup), "ms_per_step": round(el * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else args.precision,
        "data": "synthetic (seeded procedural room, the reference's equirect ray grid)",
        "config": {"workload": "reference bake settings", "scene": "room seed 2", "probes": P,
                   "rays_per_probe": R, "sh_order": 3, "shading": "reference (env_map.rchit)",
                   "precision": args.precision, "trace_variant": variant},
        "grays_per_s": round(rays / el / 1e9, 3), "shadow_rays_per_bake": cnt["shadow_rays"],
        "kernel_ms": {"trace": round(st["trace_ms"], 3), "project": round(st["project_ms"], 3),
                      "merge": round(st["merge_ms"], 3)},
        "sh_rel_err_vs_fp64": err, "cpu_baseline": cpu}


def bench_bounce(args):
    """Multi-bounce on the c2 room: a direct bake, then bounces over the
    plan's own 16x8x16 probe grid (irradiance lookup at every primary hit),
    once without and once with the 8 corner visibility rays per hit.  A step
    is one bounce (one execute, the source grid refreshed on the device)."""
    import numpy as np

    import vlb

    scene = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    params = vlb.config_c2_params(scene, seed=2, precision=PREC[args.precision])
    params.trace_variant = args.variant
    P, R = params.probes.shape[0], params.rays
    b0, b1, cnt3 = [0.3125, 0.25, 0.3125], [9.6875, 3.75, 9.6875], [16, 8, 16]
    probes, step = vlb.probe_grid(b0, b1, cnt3)
    assert np.array_equal(probes, params.probes)
    g = vlb.grid_desc(b0, step, cnt3, num_coefs=9)
    ctx = vlb.Context(0)
    sid = ctx.upload_scene_dict(scene)
    orc = oracle_leg(args)
    oscene = orc.Scene.from_dict(scene) if orc is not None else None
    sub = np.sort(np.random.default_rng(0).choice(P, size=32, replace=False))
    psub = vlb.config_c2_params(scene, seed=2)
    psub.probes = params.probes[sub].copy()
    legs = {}
    for vis in (False, True):
        lp = vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, vis, 2e-3)
        plan = ctx.plan(sid, params)
        plan.execute()
        direct = plan.download()
        plan.set_indirect(g, lp)
        for _ in range(args.warmup):
            plan.execute()
        plan.sync()
        plan.set_timing(True)
        plan.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.execute()
        plan.sync()
        t1 = time.perf_counter()
        st = plan.stats()
        prev = plan.download()
        plan.execute()
        cur = plan.download()
        plan.count()
        c = plan.stats()
        err_line = None
        if oscene is not None:   # CPU leg: this bounce against the FP64 oracle fed the same grid
            ref = oscene.bake_indirect(psub, g, prev, lp, 1.0, mode64=True,
                                       nthreads=CPU_THREADS)["coeffs"]
            err = rel_err(cur[sub], ref)
            err_line = {"max": float(err.max()), "p99": float(np.percentile(err, 99)),
                        "probes": int(sub.size)}
        legs[vis] = {"probes_per_s": round(P * args.steps / (t1 - t0), 1),
                     "ms_per_bounce": round(1e3 * (t1 - t0) / args.steps, 4),
                     "trace_ms": round(st["trace_ms"], 4),
                     "shadow_rays_per_bounce": c["shadow_rays"],
                     "grays_per_s": round((c["primary_rays"] + c["shadow_rays"]) * args.steps / (t1 - t0) / 1e9, 3),
                     "c00_gain_vs_direct": round(float(np.abs(cur[:, 0, :]).sum() / np.abs(direct[:, 0, :]).sum()), 4),
                     "sh_rel_err_vs_fp64": err_line}
        plan.destroy()
    ctx.close()
    main_leg = legs[True]
    return {
        "metric": METRIC, "value": main_leg["probes_per_s"], "unit": "probes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": main_leg["ms_per_bounce"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else args.precision,
        "data": "synthetic (seeded procedural room, Fibonacci rays)",
        "config": {"workload": "c2 multi-bounce (one step = one bounce)", "scene": "room seed 2",
                   "probes": P, "rays_per_probe": R, "grid": cnt3, "lookup": "irradiance",
                   "visibility_rays": True, "precision": args.precision},
        "with_visibility": legs[True], "without_visibility": legs[False]}


def bench_copy(args):
    """Streaming copy of 1 GiB (k_copy): attainable HBM bandwidth, and the
    reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import vlb

    ctx = vlb.Context(0)
    nbytes = 1 << 30
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    for _ in range(args.warmup):
        ctx.device_copy(dst, src, nbytes)
    ms = [ctx.device_copy(dst, src, nbytes) for _ in range(args.steps)]
    best = min(ms)
    mean = sum(ms) / len(ms)
    ctx.device_read(src, nbytes)
    rd = [ctx.device_read(src, nbytes) for _ in range(args.steps)]
    read_gbs = nbytes / (sum(rd) / len(rd) / 1e3) / 1e9
    line = {"metric": "streaming copy bandwidth (read+write)", "value": round(2 * nbytes / (mean / 1e3) / 1e9, 1),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(mean, 4), "best_ms": round(best, 4), "bytes_copied": nbytes,
            "higher_is_better": True, "dtype": "u8", "config": {"workload": "k_copy 1 GiB"},
            "read_only_gbs": round(read_gbs, 1)}
    ctx.close()
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c4",
                    choices=["c4", "c2", "c1", "c3", "c5", "copy", "sweep", "bounce", "reference"])
    ap.add_argument("--rate", type=float, default=40.0, help="served form: request arrivals per second")
    ap.add_argument("--policy", default="srpt", choices=sorted(SCHED), help="served form: step scheduling")
    ap.add_argument("--requests", type=int, default=200,
                    help="served form: number of requests (0 skips the served leg of the default line)")
    ap.add_argument("--state-budget", type=int, default=256, help="served form: state arena (MiB)")
    ap.add_argument("--step-rays", type=int, default=1 << 23,
                    help="served form: primary rays per step (2^23: half the median latency of 2^25 at "
                         "the same p99 when arrival-bound)")
    ap.add_argument("--age-ms", type=int, default=0,
                    help="served form: SRPT age bound (ms since submit; 0 = none)")
    ap.add_argument("--plain-passes", type=int, default=2,
                    help="batched form: timed passes of the plain per-job baker (0 skips it)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline: seconds of oracle work")
    ap.add_argument("--precision", default="fp32", choices=sorted(PREC))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant", type=int, default=0, help="trace kernel variant (0 = autotuned / default)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the probe baker has no multi-GPU path (probes would shard "
              "by probe group, each GPU holding a BVH replica).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    line = {"c4": bench_c4, "c2": bench_c2, "c1": bench_c1, "c3": bench_c3, "c5": bench_c5,
            "copy": bench_copy, "sweep": sweep_c2, "bounce": bench_bounce,
            "reference": bench_reference}[args.config](args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
