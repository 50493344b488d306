"""ctypes binding of libvlb.so (include/vlb_bake.h) for Python callers.

The product is the C ABI in libvlb.so (HIP kernels for gfx950 + C++ host);
this module only marshals numpy arrays across it.  There is no CPU fallback:
if the library is missing, importing raises, and on a machine without a GPU
``Context()`` raises ``VlbError`` with status ``NO_DEVICE``.

Names follow the reference baker: ``probe_grid`` is
LightBaker::probePositionsFromBoudingBox (src/baker/light_baker.cpp:80-101),
``Context.bake`` is LightBaker::bake (light_baker.cpp:287-328) with
per-probe reset, ``base64_encode`` is light_baker.cpp:330-373.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VLB_LIB selects another build of the same sources (the sanitizer build of
# the CPU tests, vulkan-light-bakery_amd/Makefile asan-test)
LIB_PATH = os.environ.get("VLB_LIB") or os.path.join(_HERE, "libvlb.so")

# ---------------------------------------------------------------- constants
OK, ERR_INVALID, ERR_DEVICE, ERR_OOM, ERR_BUDGET, ERR_NOT_FOUND, ERR_BUSY, \
    ERR_UNSUPPORTED, ERR_TIMEOUT, ERR_CANCELLED, ERR_INJECTED, ERR_NO_DEVICE = range(12)
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "DEVICE", 3: "OOM", 4: "BUDGET", 5: "NOT_FOUND",
                6: "BUSY", 7: "UNSUPPORTED", 8: "TIMEOUT", 9: "CANCELLED", 10: "INJECTED",
                11: "NO_DEVICE"}
PREC_FP32, PREC_BF16, PREC_FP16, PREC_BF16_FAST, PREC_FP16_FAST = 0, 1, 2, 3, 4
SHADE_LAMBERT, SHADE_REFERENCE = 0, 1
DIRS_FIBONACCI, DIRS_EQUIRECT = 0, 1
SH_LITERAL, SH_EXACT = 0, 1
ORDER_LINEAR, ORDER_REFERENCE_BFS = 0, 1
LOOKUP_REFERENCE, LOOKUP_IRRADIANCE = 0, 1
LIGHT_DIRECTIONAL, LIGHT_POINT = 0, 1
REF_SRGB, REF_UNORM8 = 1, 2
EQUIRECT_SH_COMP, EQUIRECT_SKYBOX_COMP = 0, 1
SCENE_ROOM, SCENE_TERRAIN, SCENE_CITY, SCENE_CUBE, SCENE_SPHERES = range(5)
SCHED_SRPT, SCHED_SHARE, SCHED_FIFO = 0, 1, 2
INJECT_PLAN_EXECUTE, INJECT_BATCH_EXECUTE, INJECT_SERVER_STEP = 0, 1, 2
MAX_LIGHTS = 8


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("vec", C.c_float * 3), ("rgb", C.c_float * 3)]


class Sky(C.Structure):
    _fields_ = [("zenith_rgb", C.c_float * 3), ("horizon_rgb", C.c_float * 3),
                ("sun_dir", C.c_float * 3), ("sun_rgb", C.c_float * 3),
                ("sun_sharpness", C.c_float)]


class RefShading(C.Structure):
    _fields_ = [("light_position", C.c_float * 3), ("shadow_bias", C.c_float),
                ("ambient", C.c_float), ("c_diffuse", C.c_float), ("c_specular", C.c_float),
                ("glossiness", C.c_float), ("flags", C.c_uint32)]


class BakeParams(C.Structure):
    _fields_ = [("num_probes", C.c_uint32), ("probe_xyz", C.POINTER(C.c_float)),
                ("dir_mode", C.c_int32), ("rays_per_probe", C.c_uint32),
                ("equirect_w", C.c_uint32), ("equirect_h", C.c_uint32),
                ("t_min", C.c_float), ("t_max", C.c_float),
                ("sh_order", C.c_uint32), ("sh_constants", C.c_int32), ("precision", C.c_int32),
                ("shade_mode", C.c_int32), ("num_lights", C.c_uint32),
                ("lights", C.POINTER(Light)), ("shadow_bias", C.c_float),
                ("sky", Sky), ("ref", RefShading),
                ("rays_per_slice", C.c_uint32), ("trace_variant", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class SceneInfo(C.Structure):
    _fields_ = [("num_tris", C.c_uint32), ("num_nodes", C.c_uint32), ("depth", C.c_uint32),
                ("reserved", C.c_uint32), ("device_bytes", C.c_uint64), ("build_ms", C.c_double),
                ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3)]


class BakeStats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("trace_ms", C.c_double), ("project_ms", C.c_double), ("merge_ms", C.c_double),
                ("total_ms", C.c_double),
                ("tiles", C.c_uint32), ("grid", C.c_uint32), ("rays_per_slice", C.c_uint32),
                ("slices", C.c_uint32), ("timed_runs", C.c_uint32), ("reserved", C.c_uint32)]


class BatchJob(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("reserved", C.c_uint32), ("params", C.POINTER(BakeParams))]


class BatchOptions(C.Structure):
    _fields_ = [("max_round_rays", C.c_uint64), ("trace_variant", C.c_uint32),
                ("chunk", C.c_uint32), ("keep_f64", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class BatchInfo(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("projection_tiles", C.c_uint64), ("device_bytes", C.c_uint64),
                ("total_ms", C.c_double),
                ("jobs", C.c_uint32), ("rounds", C.c_uint32), ("grid", C.c_uint32),
                ("trace_variant", C.c_uint32), ("chunk", C.c_uint32), ("timed_runs", C.c_uint32),
                ("trace_ms", C.c_double), ("project_ms", C.c_double), ("merge_ms", C.c_double),
                ("kernel_timed_runs", C.c_uint32), ("reserved", C.c_uint32)]


class ServerOptions(C.Structure):
    _fields_ = [("state_budget", C.c_uint64), ("step_rays", C.c_uint64),
                ("result_bytes", C.c_uint64), ("upload_bytes", C.c_uint64),
                ("max_active", C.c_uint32), ("max_lights", C.c_uint32), ("precision", C.c_int32),
                ("trace_variant", C.c_uint32), ("chunk", C.c_uint32), ("max_depth", C.c_uint32),
                ("policy", C.c_uint32), ("age_ms", C.c_uint32)]


class RequestInfo(C.Structure):
    _fields_ = [("submit_ms", C.c_double), ("admit_ms", C.c_double), ("start_ms", C.c_double),
                ("done_ms", C.c_double), ("primary_rays", C.c_uint64), ("steps", C.c_uint32),
                ("status", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("steps", C.c_uint64),
                ("primary_rays", C.c_uint64), ("probes", C.c_uint64),
                ("state_bytes_peak", C.c_uint64), ("busy_ms", C.c_double),
                ("clock_ms", C.c_double), ("active_peak", C.c_uint32), ("queued", C.c_uint32),
                ("status", C.c_int32), ("trace_variant", C.c_uint32)]


class ProbeGridDesc(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("step", C.c_float * 3), ("count", C.c_uint32 * 3),
                ("order", C.c_int32), ("num_coefs", C.c_uint32), ("sh_constants", C.c_int32)]


class LookupParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("visibility", C.c_uint32), ("bias", C.c_float),
                ("reserved", C.c_uint32)]


def grid_desc(origin, step, count, order=None, num_coefs=9, sh_constants=None) -> ProbeGridDesc:
    g = ProbeGridDesc()
    g.origin[:] = [float(x) for x in origin]
    g.step[:] = [float(x) for x in step]
    g.count[:] = [int(x) for x in count]
    g.order = ORDER_LINEAR if order is None else order
    g.num_coefs = num_coefs
    g.sh_constants = SH_EXACT if sh_constants is None else sh_constants
    return g


def lookup_params(mode=0, visibility=True, bias=1e-3) -> LookupParams:
    lp = LookupParams()
    lp.mode, lp.visibility, lp.bias = mode, int(visibility), bias
    return lp


class IndirectParams(C.Structure):
    _fields_ = [("grid", ProbeGridDesc), ("lookup", LookupParams), ("scale", C.c_float),
                ("reserved", C.c_uint32 * 3)]


class HostScene(C.Structure):
    _fields_ = [("nverts", C.c_uint32), ("ntris", C.c_uint32), ("xyz", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("idx", C.POINTER(C.c_uint32)),
                ("albedo", C.POINTER(C.c_float)), ("bounds_min", C.c_float * 3),
                ("bounds_max", C.c_float * 3)]


# ------------------------------------------------------------------ loading
def _load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"libvlb.so not built at {path}: run __graft_entry__.build()")
    lib = C.CDLL(path)
    P, VP, F, U32, I32, U64, SZ = (C.POINTER, C.c_void_p, C.POINTER(C.c_float), C.c_uint32,
                                   C.c_int32, C.c_uint64, C.c_size_t)
    sig = {
        "vlb_abi_version": (I32, []),
        "vlb_ctx_create": (I32, [I32, U64, P(VP)]),
        "vlb_ctx_destroy": (I32, [VP]),
        "vlb_last_error": (C.c_char_p, [VP]),
        "vlb_device_info": (I32, [VP, P(I32), C.c_char_p, SZ]),
        "vlb_debug_inject": (I32, [VP, I32, I32]),
        "vlb_scene_upload": (I32, [VP, F, U32, P(U32), U32, F, F, P(U32)]),
        "vlb_scene_release": (I32, [VP, U32]),
        "vlb_scene_info_get": (I32, [VP, U32, P(SceneInfo)]),
        "vlb_scene_export_bvh": (I32, [VP, U32, F, SZ, P(I32), SZ]),
        "vlb_trace_closest": (I32, [VP, U32, F, F, U32, C.c_float, C.c_float, P(I32), F, F, F]),
        "vlb_trace_occluded": (I32, [VP, U32, F, F, F, U32, C.c_float, P(C.c_uint8)]),
        "vlb_bake": (I32, [VP, U32, P(BakeParams), F, SZ]),
        "vlb_plan_create": (I32, [VP, U32, P(BakeParams), P(VP)]),
        "vlb_plan_execute": (I32, [VP]),
        "vlb_plan_sync": (I32, [VP]),
        "vlb_plan_download": (I32, [VP, F, SZ]),
        "vlb_plan_download_f64": (I32, [VP, P(C.c_double), SZ]),
        "vlb_plan_set_timing": (I32, [VP, I32]),
        "vlb_plan_use_graph": (I32, [VP, I32]),
        "vlb_plan_count": (I32, [VP]),
        "vlb_plan_stats": (I32, [VP, P(BakeStats)]),
        "vlb_plan_autotune": (I32, [VP, U32, P(U32)]),
        "vlb_plan_destroy": (I32, [VP]),
        "vlb_bake_radiance": (I32, [VP, U32, P(BakeParams), F, SZ, P(I32), SZ]),
        "vlb_batch_create": (I32, [VP, P(BatchJob), U32, P(BatchOptions), P(VP)]),
        "vlb_batch_execute": (I32, [VP]),
        "vlb_batch_sync": (I32, [VP]),
        "vlb_batch_use_graph": (I32, [VP, I32]),
        "vlb_batch_set_timing": (I32, [VP, I32]),
        "vlb_batch_download": (I32, [VP, U32, F, SZ]),
        "vlb_batch_download_f64": (I32, [VP, U32, P(C.c_double), SZ]),
        "vlb_batch_autotune": (I32, [VP, U32, P(U32), P(U32)]),
        "vlb_batch_count": (I32, [VP]),
        "vlb_batch_info_get": (I32, [VP, P(BatchInfo)]),
        "vlb_batch_destroy": (I32, [VP]),
        "vlb_server_create": (I32, [VP, P(ServerOptions), P(VP)]),
        "vlb_submit": (I32, [VP, U32, P(BakeParams), P(U64)]),
        "vlb_poll": (I32, [VP, U64, P(I32)]),
        "vlb_wait": (I32, [VP, U64, F, SZ, P(RequestInfo)]),
        "vlb_wait_for": (I32, [VP, U64, C.c_double, F, SZ, P(RequestInfo)]),
        "vlb_cancel": (I32, [VP, U64]),
        "vlb_server_hold": (I32, [VP, I32]),
        "vlb_server_stats_get": (I32, [VP, P(ServerStats)]),
        "vlb_server_destroy": (I32, [VP]),
        "vlb_skybox_sh": (I32, [VP, F, U32, U32, U32, F, SZ]),
        "vlb_skybox_sh_d": (I32, [VP, VP, U32, U32, U32, VP]),
        "vlb_skybox_sh_noalloc": (I32, [VP, VP, U32, U32, U32, VP]),
        "vlb_skybox_last_ms": (I32, [VP, P(C.c_double), P(C.c_double)]),
        "vlb_project_equirect": (I32, [VP, P(C.c_uint8), U32, U32, I32, F]),
        "vlb_sh_reconstruct": (I32, [VP, F, U32, U32, P(C.c_uint8)]),
        "vlb_probe_lookup": (I32, [VP, U32, P(ProbeGridDesc), F, P(LookupParams), F, F, U32, F]),
        "vlb_plan_set_indirect": (I32, [VP, P(IndirectParams), F]),
        "vlb_device_alloc": (I32, [VP, SZ, P(VP)]),
        "vlb_device_free": (I32, [VP, VP]),
        "vlb_upload": (I32, [VP, VP, VP, SZ]),
        "vlb_download": (I32, [VP, VP, VP, SZ]),
        "vlb_ctx_sync": (I32, [VP]),
        "vlb_device_copy": (I32, [VP, VP, VP, SZ, P(C.c_double)]),
        "vlb_device_read": (I32, [VP, VP, SZ, U32, P(C.c_double)]),
        "vlb_probe_grid": (I32, [F, F, P(U32), I32, F, F]),
        "vlb_directions": (I32, [I32, U32, U32, U32, F, F, SZ]),
        "vlb_base64_encode": (I32, [P(C.c_uint8), SZ, C.c_char_p, SZ, P(SZ)]),
        "vlb_gltf_bake_json": (I32, [C.c_char_p, F, U32, U32, F, C.c_char_p, SZ, P(SZ)]),
        "vlb_scene_generate": (I32, [I32, U64, U32, U32, P(P(HostScene))]),
        "vlb_host_scene_free": (None, [P(HostScene)]),
        "vlb_gltf_load": (I32, [C.c_char_p, P(P(HostScene)), P(C.c_void_p)]),
        "vlb_text_free": (None, [C.c_void_p]),
        "vlb_gltf_read_bake": (I32, [C.c_char_p, F, P(U32), F, SZ, P(SZ)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class VlbError(RuntimeError):
    def __init__(self, status: int, msg: str = ""):
        self.status = status
        super().__init__(f"vlb status {status} ({STATUS_NAMES.get(status, '?')}): {msg}")


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, shape_last=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    return a


# ------------------------------------------------------------- host helpers
def probe_grid(bmin, bmax, counts, order=ORDER_LINEAR):
    bmin = _f32(bmin).reshape(3)
    bmax = _f32(bmax).reshape(3)
    cnt = np.ascontiguousarray(np.asarray(counts, dtype=np.uint32).reshape(3))
    P = int(np.prod(cnt.astype(np.int64)))
    out = np.zeros((P, 3), np.float32)
    step = np.zeros(3, np.float32)
    st = lib.vlb_probe_grid(_fptr(bmin), _fptr(bmax), cnt.ctypes.data_as(C.POINTER(C.c_uint32)),
                            order, _fptr(out), _fptr(step))
    if st:
        raise VlbError(st, "probe_grid")
    return out, step


def directions(dir_mode=DIRS_FIBONACCI, rays=0, w=0, h=0):
    n = rays if dir_mode == DIRS_FIBONACCI else w * h
    xyz = np.zeros((n, 3), np.float32)
    wt = np.zeros(n, np.float32)
    st = lib.vlb_directions(dir_mode, rays, w, h, _fptr(xyz), _fptr(wt), n)
    if st:
        raise VlbError(st, "directions")
    return xyz, wt


def base64_encode(data: bytes) -> str:
    buf = (C.c_uint8 * max(1, len(data))).from_buffer_copy(data or b"\0")
    need = C.c_size_t(0)
    lib.vlb_base64_encode(buf, len(data), None, 0, C.byref(need))
    out = C.create_string_buffer(need.value + 1)
    st = lib.vlb_base64_encode(buf, len(data), out, need.value + 1, C.byref(need))
    if st:
        raise VlbError(st, "base64")
    return out.value.decode()


def gltf_bake_json(gltf_text: str, coeffs: np.ndarray, grid_step) -> str:
    c = np.ascontiguousarray(coeffs, dtype=np.float32)
    P, K = c.shape[0], c.shape[1]
    gs = _f32(grid_step).reshape(3)
    src = gltf_text.encode()
    need = C.c_size_t(0)
    st = lib.vlb_gltf_bake_json(src, _fptr(c), P, K, _fptr(gs), None, 0, C.byref(need))
    if st:
        raise VlbError(st, "gltf_bake_json")
    out = C.create_string_buffer(need.value + 1)
    st = lib.vlb_gltf_bake_json(src, _fptr(c), P, K, _fptr(gs), out, need.value + 1, C.byref(need))
    if st:
        raise VlbError(st, "gltf_bake_json")
    return out.value.decode()


def _host_scene_dict(ptr) -> dict:
    s = ptr.contents
    nv, nt = s.nverts, s.ntris
    out = {
        "xyz": np.ctypeslib.as_array(s.xyz, (max(1, nv * 3),))[:nv * 3].reshape(nv, 3).copy(),
        "normals": np.ctypeslib.as_array(s.normals, (max(1, nv * 3),))[:nv * 3].reshape(nv, 3).copy(),
        "idx": np.ctypeslib.as_array(s.idx, (max(1, nt * 3),))[:nt * 3].reshape(nt, 3).copy(),
        "albedo": np.ctypeslib.as_array(s.albedo, (max(1, nt * 3),))[:nt * 3].reshape(nt, 3).copy(),
        "bounds_min": np.array(s.bounds_min[:], np.float32),
        "bounds_max": np.array(s.bounds_max[:], np.float32),
    }
    lib.vlb_host_scene_free(ptr)
    return out


def load_gltf(path: str):
    """glTF / GLB scene -> (scene dict, glTF JSON text)."""
    ptr = C.POINTER(HostScene)()
    txt = C.c_void_p()
    st = lib.vlb_gltf_load(str(path).encode(), C.byref(ptr), C.byref(txt))
    if st:
        raise VlbError(st, f"gltf_load({path})")
    text = C.string_at(txt.value).decode() if txt.value else ""
    lib.vlb_text_free(txt)
    return _host_scene_dict(ptr), text


def read_baked_gltf(text: str):
    """Baked-light glTF -> (coefficients [n] float32, grid step, coefficient count)."""
    step = np.zeros(3, np.float32)
    k = C.c_uint32(0)
    nf = C.c_size_t(0)
    src = text.encode()
    st = lib.vlb_gltf_read_bake(src, _fptr(step), C.byref(k), None, 0, C.byref(nf))
    if st:
        raise VlbError(st, "gltf_read_bake")
    out = np.zeros(nf.value, np.float32)
    st = lib.vlb_gltf_read_bake(src, _fptr(step), C.byref(k), _fptr(out), out.size, C.byref(nf))
    if st:
        raise VlbError(st, "gltf_read_bake")
    return out, step, k.value


def generate_scene(kind: int, seed: int, p0: int = 0, p1: int = 0) -> dict:
    ptr = C.POINTER(HostScene)()
    st = lib.vlb_scene_generate(kind, seed, p0, p1, C.byref(ptr))
    if st:
        raise VlbError(st, "scene_generate")
    s = ptr.contents
    nv, nt = s.nverts, s.ntris
    out = {
        "xyz": np.ctypeslib.as_array(s.xyz, (nv * 3,)).reshape(nv, 3).copy(),
        "normals": np.ctypeslib.as_array(s.normals, (nv * 3,)).reshape(nv, 3).copy(),
        "idx": np.ctypeslib.as_array(s.idx, (nt * 3,)).reshape(nt, 3).copy(),
        "albedo": np.ctypeslib.as_array(s.albedo, (nt * 3,)).reshape(nt, 3).copy(),
        "bounds_min": np.array(s.bounds_min[:], np.float32),
        "bounds_max": np.array(s.bounds_max[:], np.float32),
    }
    lib.vlb_host_scene_free(ptr)
    return out


# ----------------------------------------------------------------- params
@dataclass
class Params:
    """Python mirror of vlb_bake_params; ``to_c()`` keeps the arrays alive."""
    probes: np.ndarray
    rays: int = 4096
    dir_mode: int = DIRS_FIBONACCI
    equirect_w: int = 0
    equirect_h: int = 0
    t_min: float = 0.001
    t_max: float = 10000.0
    sh_order: int = 2
    sh_constants: int = SH_EXACT
    precision: int = PREC_FP32
    shade_mode: int = SHADE_LAMBERT
    lights: Sequence[tuple] = ()          # (type, vec3, rgb3)
    shadow_bias: float = 1e-3
    sky_zenith: Sequence[float] = (2.0, 2.0, 2.0)
    sky_horizon: Sequence[float] = (0.5, 0.5, 0.5)
    sun_dir: Sequence[float] = (0.0, 1.0, 0.0)
    sun_rgb: Sequence[float] = (50.0, 50.0, 50.0)
    sun_sharpness: float = 200.0
    ref_light: Sequence[float] = (1.0, 10.0, 1.0)
    ref_bias: float = 0.005
    ref_ambient: float = 0.0
    ref_cd: float = 0.5
    ref_cs: float = 0.5
    ref_gloss: float = 16.0
    ref_flags: int = REF_SRGB | REF_UNORM8
    rays_per_slice: int = 0
    trace_variant: int = 0
    _keep: list = field(default_factory=list, repr=False)

    @property
    def K(self) -> int:
        return 9 if self.sh_order == 2 else 16

    @property
    def R(self) -> int:
        return self.rays if self.dir_mode == DIRS_FIBONACCI else self.equirect_w * self.equirect_h

    def to_c(self) -> BakeParams:
        pr = _f32(self.probes).reshape(-1, 3)
        p = BakeParams()
        p.num_probes = pr.shape[0]
        p.probe_xyz = _fptr(pr)
        p.dir_mode = self.dir_mode
        p.rays_per_probe = self.rays
        p.equirect_w = self.equirect_w
        p.equirect_h = self.equirect_h
        p.t_min = self.t_min
        p.t_max = self.t_max
        p.sh_order = self.sh_order
        p.sh_constants = self.sh_constants
        p.precision = self.precision
        p.shade_mode = self.shade_mode
        lights = (Light * max(1, len(self.lights)))()
        for i, (ty, vec, rgb) in enumerate(self.lights):
            lights[i].type = ty
            lights[i].vec[:] = [float(x) for x in vec]
            lights[i].rgb[:] = [float(x) for x in rgb]
        p.num_lights = len(self.lights)
        p.lights = C.cast(lights, C.POINTER(Light))
        p.shadow_bias = self.shadow_bias
        p.sky.zenith_rgb[:] = list(self.sky_zenith)
        p.sky.horizon_rgb[:] = list(self.sky_horizon)
        p.sky.sun_dir[:] = list(self.sun_dir)
        p.sky.sun_rgb[:] = list(self.sun_rgb)
        p.sky.sun_sharpness = self.sun_sharpness
        p.ref.light_position[:] = list(self.ref_light)
        p.ref.shadow_bias = self.ref_bias
        p.ref.ambient = self.ref_ambient
        p.ref.c_diffuse = self.ref_cd
        p.ref.c_specular = self.ref_cs
        p.ref.glossiness = self.ref_gloss
        p.ref.flags = self.ref_flags
        p.rays_per_slice = self.rays_per_slice
        p.trace_variant = self.trace_variant
        self._keep[:] = [pr, lights, p]
        return p


# ---------------------------------------------------------------- context
class Plan:
    def __init__(self, ctx: "Context", handle: C.c_void_p, P: int, K: int):
        self.ctx, self.h, self.P, self.K = ctx, handle, P, K

    def _chk(self, st, what):
        if st:
            raise VlbError(st, f"{what}: {self.ctx.last_error()}")

    def execute(self):
        self._chk(lib.vlb_plan_execute(self.h), "plan_execute")

    def sync(self):
        self._chk(lib.vlb_plan_sync(self.h), "plan_sync")

    def download(self) -> np.ndarray:
        out = np.zeros((self.P, self.K, 3), np.float32)
        self._chk(lib.vlb_plan_download(self.h, _fptr(out), out.size), "plan_download")
        return out

    def download_f64(self) -> np.ndarray:
        out = np.zeros((self.P, self.K, 3), np.float64)
        self._chk(lib.vlb_plan_download_f64(self.h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                            out.size), "plan_download_f64")
        return out

    def set_timing(self, on=True):
        lib.vlb_plan_set_timing(self.h, int(on))

    def use_graph(self, on=True):
        lib.vlb_plan_use_graph(self.h, int(on))

    def count(self):
        self._chk(lib.vlb_plan_count(self.h), "plan_count")

    def autotune(self, iters: int = 5) -> int:
        v = C.c_uint32(0)
        self._chk(lib.vlb_plan_autotune(self.h, iters, C.byref(v)), "plan_autotune")
        return v.value

    def stats(self) -> dict:
        s = BakeStats()
        lib.vlb_plan_stats(self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in BakeStats._fields_}

    def set_indirect(self, grid: "ProbeGridDesc" = None, lp: "LookupParams" = None, scale=1.0,
                     coeffs=None):
        """Multi-bounce (vlb_plan_set_indirect): from now on every execute adds
        albedo x scale x the grid lookup at each primary hit, reading the
        previous execute's result (or `coeffs` for the first one).  grid None:
        back to direct light only."""
        if grid is None:
            self._chk(lib.vlb_plan_set_indirect(self.h, None, None), "plan_set_indirect")
            return
        ip = IndirectParams()
        ip.grid = grid
        ip.lookup = lp if lp is not None else lookup_params(LOOKUP_IRRADIANCE, False)
        ip.scale = scale
        c = None
        if coeffs is not None:
            c = np.ascontiguousarray(coeffs, dtype=np.float32)
            if c.size != self.P * self.K * 3:
                raise ValueError("coeffs must hold P x K x 3 floats")
        self._chk(lib.vlb_plan_set_indirect(self.h, C.byref(ip), None if c is None else _fptr(c)),
                  "plan_set_indirect")

    def destroy(self):
        if self.h:
            lib.vlb_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """Many bake jobs run as one unit (vlb_batch_*)."""

    def __init__(self, ctx: "Context", handle: C.c_void_p, shapes):
        self.ctx, self.h, self.shapes = ctx, handle, shapes

    def _chk(self, st, what):
        if st:
            raise VlbError(st, f"{what}: {self.ctx.last_error()}")

    def execute(self):
        self._chk(lib.vlb_batch_execute(self.h), "batch_execute")

    def sync(self):
        self._chk(lib.vlb_batch_sync(self.h), "batch_sync")

    def use_graph(self, on=True):
        self._chk(lib.vlb_batch_use_graph(self.h, int(on)), "batch_use_graph")

    def set_timing(self, on=True):
        """Per-kernel hipEvent timing of every later execute (trace_ms,
        project_ms, merge_ms in info())."""
        self._chk(lib.vlb_batch_set_timing(self.h, int(on)), "batch_set_timing")

    def download(self, job: int) -> np.ndarray:
        out = np.zeros(self.shapes[job], np.float32)
        self._chk(lib.vlb_batch_download(self.h, job, _fptr(out), out.size), "batch_download")
        return out

    def download_f64(self, job: int) -> np.ndarray:
        out = np.zeros(self.shapes[job], np.float64)
        self._chk(lib.vlb_batch_download_f64(self.h, job, out.ctypes.data_as(C.POINTER(C.c_double)),
                                             out.size), "batch_download_f64")
        return out

    def autotune(self, iters: int = 3):
        v, c = C.c_uint32(0), C.c_uint32(0)
        self._chk(lib.vlb_batch_autotune(self.h, iters, C.byref(v), C.byref(c)), "batch_autotune")
        return v.value, c.value

    def count(self):
        self._chk(lib.vlb_batch_count(self.h), "batch_count")

    def info(self) -> dict:
        s = BatchInfo()
        self._chk(lib.vlb_batch_info_get(self.h, C.byref(s)), "batch_info")
        return {k: getattr(s, k) for k, _ in BatchInfo._fields_}

    def destroy(self):
        if self.h:
            lib.vlb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Server:
    """Rebake request server (vlb_server_*): continuous batching on the device."""

    def __init__(self, ctx: "Context", precision: int = PREC_FP32, state_budget: int = 0,
                 step_rays: int = 0, max_active: int = 0, max_lights: int = 0,
                 trace_variant: int = 0, chunk: int = 0, result_bytes: int = 0,
                 upload_bytes: int = 0, policy: int = 0, age_ms: int = 0):
        self.ctx = ctx
        o = ServerOptions()
        o.state_budget, o.step_rays, o.max_active = state_budget, step_rays, max_active
        o.max_lights, o.precision, o.trace_variant, o.chunk = max_lights, precision, trace_variant, chunk
        o.result_bytes, o.upload_bytes = result_bytes, upload_bytes
        o.policy, o.age_ms = policy, age_ms
        self.h = C.c_void_p()
        st = lib.vlb_server_create(ctx.h, C.byref(o), C.byref(self.h))
        if st:
            raise VlbError(st, f"server_create: {ctx.last_error()}")
        self._shape = {}
        ctx._adopt(self)

    def submit(self, sid: int, params: Params) -> int:
        p = params.to_c()
        rid = C.c_uint64(0)
        st = lib.vlb_submit(self.h, sid, C.byref(p), C.byref(rid))
        if st:
            raise VlbError(st, f"submit: {self.ctx.last_error()}")
        self._shape[rid.value] = (p.num_probes, params.K, 3)
        return rid.value

    def poll(self, rid: int) -> bool:
        d = C.c_int32(0)
        st = lib.vlb_poll(self.h, rid, C.byref(d))
        if st:
            raise VlbError(st, "poll")
        return bool(d.value)

    def wait(self, rid: int, timeout_ms: Optional[float] = None):
        """Coefficients and info of a finished request.  With timeout_ms,
        raises VlbError(ERR_TIMEOUT) if it is not done by then (the request
        stays and can be waited for again).  A request that did not finish
        raises VlbError with its status (ERR_CANCELLED, ERR_INJECTED, ...)."""
        out = np.zeros(self._shape[rid], np.float32)
        info = RequestInfo()
        if timeout_ms is None:
            st = lib.vlb_wait(self.h, rid, _fptr(out), out.size, C.byref(info))
        else:
            st = lib.vlb_wait_for(self.h, rid, float(timeout_ms), _fptr(out), out.size, C.byref(info))
            if st == ERR_TIMEOUT:
                raise VlbError(st, "wait_for: deadline passed")
        del self._shape[rid]
        if st:
            raise VlbError(st, f"wait: {self.ctx.last_error()}")
        return out, {k: getattr(info, k) for k, _ in RequestInfo._fields_}

    def cancel(self, rid: int):
        st = lib.vlb_cancel(self.h, rid)
        if st:
            raise VlbError(st, "cancel")

    def hold(self, on=True):
        """Admission hold: queued requests stay queued while on."""
        st = lib.vlb_server_hold(self.h, int(on))
        if st:
            raise VlbError(st, "server_hold")

    def stats(self) -> dict:
        s = ServerStats()
        lib.vlb_server_stats_get(self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def destroy(self):
        if self.h:
            lib.vlb_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        st = lib.vlb_ctx_create(device, hbm_budget, C.byref(h))
        if st:
            raise VlbError(st, "vlb_ctx_create")
        self.h = h
        # plans, batches and servers still alive: vlb_ctx_destroy refuses
        # while any is, so close() destroys them first
        self._children = weakref.WeakSet()

    def _adopt(self, child):
        self._children.add(child)
        return child

    def close(self):
        if getattr(self, "h", None):
            for ch in list(self._children):
                ch.destroy()
            self._chk(lib.vlb_ctx_destroy(self.h), "ctx_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def last_error(self) -> str:
        m = lib.vlb_last_error(self.h)
        return m.decode() if m else ""

    def _chk(self, st, what):
        if st:
            raise VlbError(st, f"{what}: {self.last_error()}")

    def debug_inject(self, site: int, count: int = 1):
        """The next `count` calls at `site` (INJECT_*) fail with ERR_INJECTED."""
        self._chk(lib.vlb_debug_inject(self.h, site, count), "debug_inject")

    def device_info(self):
        n = C.c_int32(0)
        name = C.create_string_buffer(128)
        self._chk(lib.vlb_device_info(self.h, C.byref(n), name, 128), "device_info")
        return {"num_cus": n.value, "name": name.value.decode()}

    # scenes
    def upload_scene(self, xyz, idx, normals=None, albedo=None) -> int:
        xyz = _f32(xyz).reshape(-1, 3)
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.uint32).reshape(-1, 3))
        nrm = None if normals is None else _f32(normals).reshape(-1, 3)
        alb = None if albedo is None else _f32(albedo).reshape(-1, 3)
        sid = C.c_uint32(0)
        st = lib.vlb_scene_upload(self.h, _fptr(xyz), xyz.shape[0],
                                  idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.shape[0],
                                  None if nrm is None else _fptr(nrm),
                                  None if alb is None else _fptr(alb), C.byref(sid))
        self._chk(st, "scene_upload")
        return sid.value

    def upload_scene_dict(self, s: dict) -> int:
        return self.upload_scene(s["xyz"], s["idx"], s.get("normals"), s.get("albedo"))

    def release_scene(self, sid: int):
        self._chk(lib.vlb_scene_release(self.h, sid), "scene_release")

    def scene_info(self, sid: int) -> dict:
        i = SceneInfo()
        self._chk(lib.vlb_scene_info_get(self.h, sid, C.byref(i)), "scene_info")
        d = {k: getattr(i, k) for k, _ in SceneInfo._fields_}
        d["bounds_min"] = list(i.bounds_min)
        d["bounds_max"] = list(i.bounds_max)
        return d

    def export_bvh(self, sid: int):
        info = self.scene_info(sid)
        nodes = np.zeros((max(1, info["num_nodes"]), 16), np.float32)
        leaf = np.zeros(max(1, info["num_tris"]), np.int32)
        self._chk(lib.vlb_scene_export_bvh(self.h, sid, _fptr(nodes), nodes.size,
                                           leaf.ctypes.data_as(C.POINTER(C.c_int32)), leaf.size),
                  "export_bvh")
        return nodes[: info["num_nodes"]], leaf[: info["num_tris"]]

    # ray queries
    def trace_closest(self, sid, org, dirs, t_min=0.001, t_max=10000.0):
        org = _f32(org).reshape(-1, 3)
        dirs = _f32(dirs).reshape(-1, 3)
        n = org.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        self._chk(lib.vlb_trace_closest(self.h, sid, _fptr(org), _fptr(dirs), n, t_min, t_max,
                                        prim.ctypes.data_as(C.POINTER(C.c_int32)), _fptr(t),
                                        _fptr(u), _fptr(v)), "trace_closest")
        return prim, t, u, v

    def trace_occluded(self, sid, org, dirs, t_max, t_min=0.0):
        org = _f32(org).reshape(-1, 3)
        dirs = _f32(dirs).reshape(-1, 3)
        n = org.shape[0]
        tm = _f32(np.broadcast_to(np.asarray(t_max, np.float32), (n,)))
        occ = np.zeros(n, np.uint8)
        self._chk(lib.vlb_trace_occluded(self.h, sid, _fptr(org), _fptr(dirs), _fptr(tm), n, t_min,
                                         occ.ctypes.data_as(C.POINTER(C.c_uint8))), "trace_occluded")
        return occ

    # baking
    def bake(self, sid: int, params: Params) -> np.ndarray:
        p = params.to_c()
        out = np.zeros((p.num_probes, params.K, 3), np.float32)
        self._chk(lib.vlb_bake(self.h, sid, C.byref(p), _fptr(out), out.size), "bake")
        return out

    def plan(self, sid: int, params: Params) -> Plan:
        p = params.to_c()
        h = C.c_void_p()
        self._chk(lib.vlb_plan_create(self.h, sid, C.byref(p), C.byref(h)), "plan_create")
        return self._adopt(Plan(self, h, p.num_probes, params.K))

    def bake_bounces(self, sid: int, params: Params, grid: "ProbeGridDesc", bounces: int,
                     lp: "LookupParams" = None, scale=1.0) -> list:
        """Direct bake, then `bounces` indirect passes over the plan's own grid;
        returns the coefficients after each pass (direct first)."""
        plan = self.plan(sid, params)
        try:
            plan.execute()
            out = [plan.download()]
            if bounces > 0:
                plan.set_indirect(grid, lp, scale)
                for _ in range(bounces):
                    plan.execute()
                    out.append(plan.download())
            return out
        finally:
            plan.destroy()

    def batch(self, jobs, max_round_rays: int = 0, trace_variant: int = 0, chunk: int = 0,
              keep_f64: bool = False) -> Batch:
        """jobs: sequence of (scene id, Params)."""
        cps = [prm.to_c() for _, prm in jobs]
        arr = (BatchJob * len(jobs))()
        for i, ((sid, _), cp) in enumerate(zip(jobs, cps)):
            arr[i].scene = sid
            arr[i].params = C.pointer(cp)
        opt = BatchOptions()
        opt.max_round_rays = max_round_rays
        opt.trace_variant = trace_variant
        opt.chunk = chunk
        opt.keep_f64 = int(keep_f64)
        h = C.c_void_p()
        self._chk(lib.vlb_batch_create(self.h, arr, len(jobs), C.byref(opt), C.byref(h)),
                  "batch_create")
        shapes = [(cp.num_probes, prm.K, 3) for (_, prm), cp in zip(jobs, cps)]
        return self._adopt(Batch(self, h, shapes))

    def bake_radiance(self, sid: int, params: Params, want_radiance=True, want_hits=True):
        p = params.to_c()
        P, R = p.num_probes, params.R
        rad = np.zeros((P, R, 3), np.float32) if want_radiance else None
        hit = np.zeros((P, R), np.int32) if want_hits else None
        self._chk(lib.vlb_bake_radiance(
            self.h, sid, C.byref(p), None if rad is None else _fptr(rad),
            0 if rad is None else rad.size,
            None if hit is None else hit.ctypes.data_as(C.POINTER(C.c_int32)),
            0 if hit is None else hit.size), "bake_radiance")
        return rad, hit

    def probe_lookup(self, sid: int, grid: "ProbeGridDesc", coeffs, lp: "LookupParams", points,
                     normals) -> np.ndarray:
        c = _f32(coeffs)
        pts = _f32(points).reshape(-1, 3)
        nrm = _f32(normals).reshape(-1, 3)
        out = np.zeros((pts.shape[0], 3), np.float32)
        self._chk(lib.vlb_probe_lookup(self.h, sid, C.byref(grid), _fptr(c), C.byref(lp), _fptr(pts),
                                       _fptr(nrm), pts.shape[0], _fptr(out)), "probe_lookup")
        return out

    # skies and images
    def skybox_sh(self, cubes: np.ndarray, sh_order: int = 2) -> np.ndarray:
        cubes = _f32(cubes)
        if cubes.ndim == 4:
            cubes = cubes[None]
        B, six, N, N2, four = cubes.shape
        assert six == 6 and N == N2 and four == 4
        K = 9 if sh_order == 2 else 16
        out = np.zeros((B, K, 3), np.float32)
        self._chk(lib.vlb_skybox_sh(self.h, _fptr(cubes), N, B, sh_order, _fptr(out), out.size),
                  "skybox_sh")
        return out

    def skybox_sh_device(self, d_cubes: int, face_size: int, batch: int, sh_order: int, d_out: int):
        self._chk(lib.vlb_skybox_sh_d(self.h, C.c_void_p(d_cubes), face_size, batch, sh_order,
                                      C.c_void_p(d_out)), "skybox_sh_d")

    def skybox_sh_noalloc(self, d_cubes: int, face_size: int, batch: int, sh_order: int, d_out: int):
        """vlb_skybox_sh_noalloc: vlb_skybox_sh_d with no device allocation
        (fixed per-GPU scratch; calls that overlap on one GPU race)."""
        self._chk(lib.vlb_skybox_sh_noalloc(self.h, C.c_void_p(d_cubes), face_size, batch, sh_order,
                                            C.c_void_p(d_out)), "skybox_sh_noalloc")

    def skybox_last_ms(self):
        a, b = C.c_double(0), C.c_double(0)
        self._chk(lib.vlb_skybox_last_ms(self.h, C.byref(a), C.byref(b)), "skybox_last_ms")
        return a.value, b.value

    def project_equirect(self, img: np.ndarray, variant: int = EQUIRECT_SH_COMP) -> np.ndarray:
        img = np.ascontiguousarray(img, dtype=np.uint8)
        H, W, _ = img.shape
        out = np.zeros((16, 3), np.float32)
        self._chk(lib.vlb_project_equirect(self.h, img.ctypes.data_as(C.POINTER(C.c_uint8)), W, H,
                                           variant, _fptr(out)), "project_equirect")
        return out

    def sh_reconstruct(self, coeffs16x3: np.ndarray, W: int, H: int) -> np.ndarray:
        c = _f32(coeffs16x3).reshape(16, 3)
        out = np.zeros((H, W, 4), np.uint8)
        self._chk(lib.vlb_sh_reconstruct(self.h, _fptr(c), W, H,
                                         out.ctypes.data_as(C.POINTER(C.c_uint8))), "sh_reconstruct")
        return out

    # device memory
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._chk(lib.vlb_device_alloc(self.h, nbytes, C.byref(p)), "device_alloc")
        return p.value

    def free(self, dptr: int):
        self._chk(lib.vlb_device_free(self.h, C.c_void_p(dptr)), "device_free")

    def upload(self, dptr: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        self._chk(lib.vlb_upload(self.h, C.c_void_p(dptr), a.ctypes.data_as(C.c_void_p), a.nbytes),
                  "upload")

    def download(self, dptr: int, arr: np.ndarray):
        self._chk(lib.vlb_download(self.h, arr.ctypes.data_as(C.c_void_p), C.c_void_p(dptr),
                                   arr.nbytes), "download")

    def sync(self):
        self._chk(lib.vlb_ctx_sync(self.h), "sync")

    def device_copy(self, dst: int, src: int, nbytes: int) -> float:
        ms = C.c_double(0)
        self._chk(lib.vlb_device_copy(self.h, C.c_void_p(dst), C.c_void_p(src), nbytes,
                                      C.byref(ms)), "device_copy")
        return ms.value

    def device_read(self, src: int, nbytes: int, grid: int = 0) -> float:
        """Event-timed read-only streaming pass over nbytes at src (ms)."""
        ms = C.c_double(0)
        self._chk(lib.vlb_device_read(self.h, C.c_void_p(src), nbytes, grid, C.byref(ms)),
                  "device_read")
        return ms.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -------------------------------------------------- benchmark configurations
def config_c2_params(scene: dict, seed: int = 2, precision: int = PREC_FP32) -> Params:
    """Single-scene plain bake: 16x8x16 probes inside the 10x4x10 m room,
    4096 Fibonacci rays, one directional light + procedural sky, L2."""
    rng = np.random.default_rng(seed)
    probes, _ = probe_grid([0.3125, 0.25, 0.3125], [9.6875, 3.75, 9.6875], [16, 8, 16])
    sun = sun_direction(rng)
    return Params(probes=probes, rays=4096, sh_order=2, precision=precision,
                  shade_mode=SHADE_LAMBERT,
                  lights=[(LIGHT_DIRECTIONAL, (-sun).tolist(), (3.0, 3.0, 3.0))],
                  sun_dir=sun.tolist(), shadow_bias=1e-3)


def procedural_cubemap(N: int, sun, zenith=2.0, horizon=0.5, sun_rgb=50.0, k=200.0,
                       tint=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Synthetic input: 6 x N x N RGBA32F cubemap of the benchmark sky (horizon to
    zenith gradient plus a sun lobe), Vulkan face order +X, -X, +Y, -Y, +Z, -Z."""
    c = (np.arange(N) + 0.5) * (2.0 / N) - 1.0
    v, u = np.meshgrid(c, c, indexing="ij")
    one = np.ones_like(u)
    faces = [(one, -v, -u), (-one, -v, u), (u, one, v), (u, -one, -v), (u, -v, one), (-u, -v, -one)]
    out = np.zeros((6, N, N, 4), np.float32)
    s = np.asarray(sun, np.float64)
    for f, (x, y, z) in enumerate(faces):
        l = np.sqrt(x * x + y * y + z * z)
        x, y, z = x / l, y / l, z / l
        g = np.maximum(y, 0.0)
        base = horizon + (zenith - horizon) * g
        lobe = sun_rgb * np.exp(k * (x * s[0] + y * s[1] + z * s[2] - 1.0))
        for ch in range(3):
            out[f, :, :, ch] = (base + lobe) * tint[ch]
        out[f, :, :, 3] = 1.0
    return out


def sun_direction(rng) -> np.ndarray:
    el = rng.uniform(0.35, 1.1)
    az = rng.uniform(0.0, 2 * np.pi)
    d = np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)], np.float64)
    return (d / np.linalg.norm(d)).astype(np.float32)
