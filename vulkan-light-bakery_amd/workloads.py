"""Seeded synthetic workloads of the benchmark configurations.

Shared by bench.py and the GPU parity tests, so that what is measured is
what is checked.  Every workload is a pure function of its seed:

* precision form: 256 x 256 fBm terrain + 50 icospheres at subdivision 5
  (1,154,050 triangles, seed 3), 32 x 16 x 32 probes x 16384 Fibonacci rays,
  L2 and L3;
* batched form: 64 jobs over four resident scenes (rooms of 11k and 99k
  triangles, terrains of 0.5M and 1.15M), probes per job log-uniform in
  [64, 32768] in a random axis-aligned region, rays per probe from
  {1024, 4096, 16384}, L2, one directional light + sky (seed 4);
* serving form: a 400-building city block (~2M triangles) plus the four
  scenes above; Poisson arrivals; each request moves the sun or adds a point
  light (intensity U[10, 500]) and rebakes 256..65536 probes (log-uniform) of
  a random region at 4096 rays per probe (seed 5).

The reference bakes one glTF scene per process (src/baker/light_baker.cpp:
38-66); these scenes and jobs run its bake loop (light_baker.cpp:287-328) at
the sizes the workloads name.
"""
from __future__ import annotations

import numpy as np

import vlb

# (label, kind, seed, param0, param1): ~10k, ~100k, ~500k, ~1.1M triangles
C4_SCENES = (
    ("room 11k", vlb.SCENE_ROOM, 40, 0, 0),
    ("room 99k", vlb.SCENE_ROOM, 41, 64, 77),
    ("terrain 499k", vlb.SCENE_TERRAIN, 42, 256, 18),
    ("terrain 1.15M", vlb.SCENE_TERRAIN, 43, 256, 50),
)
C5_SCENES = (("city 2M", vlb.SCENE_CITY, 5, 0, 0),) + C4_SCENES
C4_JOBS = 64
C4_RAYS = (1024, 4096, 16384)
C5_RAYS = 4096


def region_probes(rng, scene, n):
    """n probes uniformly in a random axis-aligned region of the scene bounds."""
    lo = np.asarray(scene["bounds_min"], np.float64)
    hi = np.asarray(scene["bounds_max"], np.float64)
    ext = hi - lo
    size = ext * rng.uniform(0.25, 1.0, 3)
    start = lo + rng.uniform(0.0, 1.0, 3) * (ext - size)
    return (start + rng.uniform(0.0, 1.0, (n, 3)) * size).astype(np.float32)


def sun_params(probes, rays, precision, sun, lights=None, sh_order=2):
    """LAMBERT bake of `probes` under a directional light from `sun` (3.0
    irradiance) and the procedural sky whose sun lobe points at `sun`."""
    return vlb.Params(probes=probes, rays=rays, sh_order=sh_order, precision=precision,
                      shade_mode=vlb.SHADE_LAMBERT, sun_dir=np.asarray(sun).tolist(),
                      lights=lights if lights is not None else
                      [(vlb.LIGHT_DIRECTIONAL, (-np.asarray(sun)).tolist(), (3.0, 3.0, 3.0))],
                      shadow_bias=2e-3)


def scenes(specs):
    """Host scenes of (label, kind, seed, p0, p1) specs."""
    return [vlb.generate_scene(kind, seed, p0, p1) for _, kind, seed, p0, p1 in specs]


# ----------------------------------------------------------- precision form
def c3_scene():
    return vlb.generate_scene(vlb.SCENE_TERRAIN, 3)


def c3_probes(scene):
    """32 x 16 x 32 probes over the terrain (linear order)."""
    lo, hi = scene["bounds_min"], scene["bounds_max"]
    probes, _ = vlb.probe_grid([4.0, lo[1] + 1.0, 4.0], [252.0, hi[1] + 20.0, 252.0], [32, 16, 32])
    return probes


def c3_sun():
    return vlb.sun_direction(np.random.default_rng(3))


def c3_params(probes, precision, sh_order, sun):
    return sun_params(probes, 16384, precision, sun, sh_order=sh_order)


# ------------------------------------------------------------ batched form
def c4_jobs(host_scenes, seed=4, njobs=C4_JOBS):
    """The jobs: (scene index, probes float32[P, 3], rays per probe), and
    the sun direction they share."""
    rng = np.random.default_rng(seed)
    sun = vlb.sun_direction(rng)
    spec = []
    for _ in range(njobs):
        si = int(rng.integers(0, len(host_scenes)))
        P = int(round(np.exp(rng.uniform(np.log(64), np.log(32768)))))
        R = int(rng.choice(C4_RAYS))
        spec.append((si, region_probes(rng, host_scenes[si], P), R))
    return spec, sun


# ------------------------------------------------------------ serving form
def c5_requests(host_scenes, n, rate, seed=5):
    """n requests over host_scenes (the city first) arriving as a Poisson
    process of `rate` per second: arrival times (s, the first at 0) and
    (scene index, probes, sun, lights) per request."""
    rng = np.random.default_rng(seed)
    sun0 = vlb.sun_direction(rng)
    gaps = rng.exponential(1.0 / rate, n)
    arrive = np.cumsum(gaps) - gaps[0]
    reqs = []
    for _ in range(n):
        si = int(rng.integers(0, len(host_scenes)))
        P = int(round(np.exp(rng.uniform(np.log(256), np.log(65536)))))
        pr = region_probes(rng, host_scenes[si], P)
        if rng.uniform() < 0.5:   # the sun moves
            sun = sun0 + rng.normal(0.0, 0.2, 3)
            sun[1] = abs(sun[1]) + 0.1
            sun /= np.linalg.norm(sun)
            lights = [(vlb.LIGHT_DIRECTIONAL, (-sun).tolist(), (3.0, 3.0, 3.0))]
        else:                     # a point light is added
            sun = sun0
            pos = pr[int(rng.integers(0, P))] + rng.normal(0.0, 1.0, 3).astype(np.float32)
            lights = [(vlb.LIGHT_DIRECTIONAL, (-sun).tolist(), (3.0, 3.0, 3.0)),
                      (vlb.LIGHT_POINT, pos.tolist(), (float(rng.uniform(10, 500)),) * 3)]
        reqs.append((si, pr, sun, lights))
    return arrive, reqs, sun0


def c5_params(req, precision):
    si, pr, sun, lights = req
    return sun_params(pr, C5_RAYS, precision, sun, lights)
