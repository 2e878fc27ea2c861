"""TEST INFRASTRUCTURE ONLY — ctypes view of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker or the timed CPU baseline. The product
library never calls it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
NO_HIT = 0xFFFFFFFF

_lib = None


class BuildInfo(ctypes.Structure):
    _fields_ = [("num_treelets", c_uint64), ("top_nodes", c_uint64), ("total_nodes", c_uint64),
                ("max_treelet", c_uint64)]


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} missing; run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    p = c_void_p
    sig = {
        "orc_morton_codes": (c_int, [p, c_uint64, p]),
        "orc_morton_sort": (c_int, [p, c_uint64, p, p]),
        "orc_bvh_build": (c_int, [p, c_uint64, POINTER(p), POINTER(BuildInfo)]),
        "orc_bvh_build_threads": (c_int, [p, c_uint64, c_int, POINTER(p), POINTER(BuildInfo)]),
        "orc_bvh_error": (c_char_p, [p]),
        "orc_bvh_free": (None, [p]),
        "orc_bvh_export_canonical": (c_int, [p, p, POINTER(c_uint64)]),
        "orc_bvh_intersect": (c_int, [p, p, c_uint64, c_int, c_int, p, p, p, p]),
        "orc_brute_intersect": (c_int, [p, c_uint64, p, c_uint64, p, p]),
        "orc_intersect_triangle": (c_int, [p, p, c_float, p]),
        "orc_ao_bounce": (c_int, [p, p, p, p, c_uint64, p]),
        "orc_pcg32_sequence": (c_uint32, [c_uint64, c_uint64, c_int, c_int64, p, c_uint64]),
        "orc_murmur64a": (c_uint64, [p, c_uint64, c_uint64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def morton_codes(tri: np.ndarray) -> np.ndarray:
    out = np.empty(tri.shape[0], np.uint32)
    load().orc_morton_codes(_p(tri), tri.shape[0], _p(out))
    return out


def morton_sort(tri: np.ndarray):
    n = tri.shape[0]
    codes, order = np.empty(n, np.uint32), np.empty(n, np.uint32)
    load().orc_morton_sort(_p(tri), n, _p(codes), _p(order))
    return codes, order


class OracleBvh:
    def __init__(self, tri: np.ndarray, threads: int = 1):
        self.tri = np.ascontiguousarray(tri, np.float32)
        self.lib = load()
        self.h = c_void_p()
        self.info = BuildInfo()
        rc = self.lib.orc_bvh_build_threads(_p(self.tri), self.tri.shape[0], int(threads), ctypes.byref(self.h),
                                            ctypes.byref(self.info))
        if rc != 0:
            msg = self.lib.orc_bvh_error(self.h).decode()
            self.lib.orc_bvh_free(self.h)
            self.h = None
            raise RuntimeError(f"oracle build failed ({rc}): {msg}")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_bvh_free(self.h)
            self.h = None

    def export_canonical(self) -> bytes:
        ln = c_uint64(0)
        self.lib.orc_bvh_export_canonical(self.h, None, ctypes.byref(ln))
        buf = np.empty(ln.value, np.uint8)
        self.lib.orc_bvh_export_canonical(self.h, _p(buf), ctypes.byref(ln))
        return buf.tobytes()

    def intersect(self, rays: np.ndarray, any_hit=False, threads=8):
        n = rays.shape[0]
        prim = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        bary = np.empty((n, 3), np.float32)
        stats = np.zeros(2, np.uint64)
        rc = self.lib.orc_bvh_intersect(self.h, _p(rays), n, int(any_hit), threads, _p(prim), _p(t), _p(bary),
                                        _p(stats))
        if rc != 0:
            raise RuntimeError("oracle traversal stack overflow")
        return prim, t, bary, stats


def brute_intersect(tri: np.ndarray, rays: np.ndarray):
    n = rays.shape[0]
    prim = np.empty(n, np.uint32)
    t = np.empty(n, np.float32)
    load().orc_brute_intersect(_p(tri), tri.shape[0], _p(rays), n, _p(prim), _p(t))
    return prim, t


def intersect_triangle(ray: np.ndarray, tri9: np.ndarray, t_max: float):
    out = np.zeros(4, np.float32)
    hit = load().orc_intersect_triangle(_p(np.ascontiguousarray(ray, np.float32)),
                                        _p(np.ascontiguousarray(tri9, np.float32)), t_max, _p(out))
    return bool(hit), out


def ao_bounce(tri: np.ndarray, rays: np.ndarray, prim: np.ndarray, bary: np.ndarray) -> np.ndarray:
    n = rays.shape[0]
    out = np.empty((n, 8), np.float32)
    load().orc_ao_bounce(_p(tri), _p(rays), _p(prim), _p(np.ascontiguousarray(bary, np.float32)), n, _p(out))
    return out


def pcg32(count: int, seq: int = 0, offset: int = 0, mode: int = 1, advance: int = 0) -> np.ndarray:
    out = np.empty(count, np.uint32)
    load().orc_pcg32_sequence(seq, offset, mode, advance, _p(out), count)
    return out


def murmur64a(data: bytes, seed: int = 0) -> int:
    buf = np.frombuffer(data, np.uint8).copy() if data else np.zeros(1, np.uint8)
    return int(load().orc_murmur64a(_p(buf), len(data), seed))


# ----------------------------------------------------------------- path tracer
class OrcMaterial(ctypes.Structure):
    _fields_ = [("type", c_uint32), ("rgb", c_float * 3), ("metal", c_uint32), ("uroughness", c_float),
                ("vroughness", c_float), ("remap_roughness", c_uint32), ("eta", c_float), ("eta_named", c_uint32)]


class OrcLight(ctypes.Structure):
    _fields_ = [("rgb", c_float * 3), ("scale", c_float), ("two_sided", c_uint32)]


class OrcInfiniteLight(ctypes.Structure):
    _fields_ = [("rgb", c_float * 3), ("has_rgb", c_uint32), ("scale", c_float), ("illuminance", c_float),
                ("light_index", c_uint32)]


class OrcShape(ctypes.Structure):
    _fields_ = [("type", c_uint32), ("xform", c_uint32), ("radius", c_float), ("zmin", c_float), ("zmax", c_float),
                ("phimax", c_float), ("height", c_float), ("innerradius", c_float)]


class OrcScene(ctypes.Structure):
    _fields_ = [("eye", c_float * 3), ("look", c_float * 3), ("up", c_float * 3), ("fov", c_float),
                ("width", c_uint32), ("height", c_uint32), ("spp", c_uint32), ("max_depth", c_uint32),
                ("pool_size", c_uint32), ("num_tris", c_uint64), ("tri", c_void_p), ("tri_material", c_void_p),
                ("tri_light", c_void_p), ("tri_reverse", c_void_p), ("num_materials", c_uint32),
                ("materials", POINTER(OrcMaterial)), ("num_lights", c_uint32), ("lights", POINTER(OrcLight)),
                ("num_xforms", c_uint32), ("xforms", c_void_p), ("tri_xform", c_void_p), ("select_scale", c_void_p),
                ("num_infinite_lights", c_uint32), ("infinite_lights", POINTER(OrcInfiniteLight)),
                ("xforms_inv", c_void_p), ("num_shapes", c_uint32), ("shapes", POINTER(OrcShape)),
                ("tri_shape", c_void_p)]


SHAPE_KINDS = {"sphere": 1, "disk": 2, "instance": 3}


METALS = {"Ag": 0, "Al": 1, "Au": 2, "Cu": 3}


class OrcBxdf(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("r", c_float * 4), ("eta4", c_float * 4), ("k4", c_float * 4),
                ("eta", c_float), ("ax", c_float), ("ay", c_float)]


def bxdf(kind, r=(0, 0, 0, 0), eta4=(0, 0, 0, 0), k4=(0, 0, 0, 0), eta=1.0, ax=0.0, ay=None):
    """A BxDF of the oracle in its local frame: kind 'diffuse', 'conductor' or 'dielectric'."""
    t = {"diffuse": 0, "conductor": 1, "dielectric": 2}[kind]
    return OrcBxdf(t, (c_float * 4)(*r), (c_float * 4)(*eta4), (c_float * 4)(*k4), eta, ax,
                   ax if ay is None else ay)


def _dirs(a):
    return np.ascontiguousarray(np.broadcast_to(a, np.shape(a)), np.float32).reshape(-1, 3)


def bxdf_flags(b):
    return int(_pt().orc_bxdf_flags(ctypes.byref(b)))


def bxdf_eval(b, wo, wi):
    """BSDF::f (4 wavelengths) and BSDF::pdf for direction pairs [n][3]."""
    wo, wi = _dirs(wo), _dirs(wi)
    n = len(wo)
    f = np.zeros((n, 4), np.float32)
    pdf = np.zeros(n, np.float32)
    _pt().orc_bxdf_eval(ctypes.byref(b), n, _p(wo), _p(wi), _p(f), _p(pdf))
    return f, pdf


def bxdf_sample(b, wo, u3):
    """BSDF::sample_f: f [n][4], wi [n][3], pdf, eta, flags (1 ok, 2 specular, 4 transmission)."""
    wo = _dirs(wo)
    u3 = np.ascontiguousarray(u3, np.float32).reshape(-1, 3)
    n = len(wo)
    f = np.zeros((n, 4), np.float32)
    wi = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    eta = np.zeros(n, np.float32)
    flags = np.zeros(n, np.uint32)
    _pt().orc_bxdf_sample(ctypes.byref(b), n, _p(wo), _p(u3), _p(f), _p(wi), _p(pdf), _p(eta), _p(flags))
    return f, wi, pdf, eta, flags


def fr_dielectric(cos_i, eta):
    return float(_pt().orc_fr_dielectric(float(cos_i), float(eta)))


def refract(wi, n, eta):
    """(transmitted direction, relative IOR), or None at total internal reflection."""
    wi = np.ascontiguousarray(wi, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    wt = np.zeros(3, np.float32)
    etap = c_float(0)
    if not _pt().orc_refract(_p(wi), _p(n), float(eta), ctypes.byref(etap), _p(wt)):
        return None
    return wt, etap.value


def tr_eval(ax, ay, w, wm):
    """Trowbridge-Reitz [n][4]: D(wm), Lambda(w), G1(w), D_w(wm)."""
    w, wm = _dirs(w), _dirs(wm)
    out = np.zeros((len(w), 4), np.float32)
    _pt().orc_tr_eval(float(ax), float(ay), len(w), _p(w), _p(wm), _p(out))
    return out


def tr_sample(ax, ay, w, u2):
    w = _dirs(w)
    u2 = np.ascontiguousarray(u2, np.float32).reshape(-1, 2)
    out = np.zeros((len(w), 3), np.float32)
    _pt().orc_tr_sample(float(ax), float(ay), len(w), _p(w), _p(u2), _p(out))
    return out


def light_pmf(sc):
    """The power light sampler's pmf in light order, and each light's triangle (-1: infinite)."""
    s, keep = _scene_struct(sc, 1)
    cap = 4096
    pmf = np.zeros(cap, np.float32)
    tri = np.zeros(cap, np.int32)
    cnt = ctypes.c_uint32(0)
    if _pt().orc_light_pmf(ctypes.byref(s), _p(pmf), _p(tri), cap, ctypes.byref(cnt)) != 0:
        raise RuntimeError("light_pmf")
    return pmf[:cnt.value].copy(), tri[:cnt.value].copy()


def _pt_sigs(lib):
    p = c_void_p
    sig = {
        "orc_render": (c_int, [POINTER(OrcScene), c_int, p, p, p]),
        "orc_render_paths": (c_int, [POINTER(OrcScene), c_int, p, p, p]),
        "orc_render_path_L": (c_int, [POINTER(OrcScene), c_int, p, p]),
        "orc_render_megakernel": (c_int, [POINTER(OrcScene), c_int, c_int, p, p, p, p]),
        "orc_rgb_to_sigmoid_jacobian": (c_int, [p, p]),
        "orc_render_grad": (c_int, [POINTER(OrcScene), c_int, p, p, POINTER(ctypes.c_double), p, p, p, c_int]),
        "orc_rgb_to_sigmoid": (c_int, [p, p]),
        "orc_srgb_matrices": (c_int, [p, p, p]),
        "orc_filter_sample": (c_int, [p, p]),
        "orc_spectrum_eval": (c_float, [c_int, c_float]),
        "orc_trace_ray": (c_int, [p, p, p, c_float, c_int, p, p, p]),
        "orc_sample_visible": (c_int, [c_float, p, p]),
        "orc_sampler_sequence": (c_int, [c_uint32, c_uint32, c_uint32, p, c_int, p]),
        "orc_camera_ray": (c_int, [POINTER(OrcScene), c_float, c_float, p]),
        "orc_spherical_triangle": (c_int, [p, p, p, p, POINTER(c_float)]),
        "orc_invert_spherical_triangle": (c_int, [p, p, p, p]),
        "orc_fr_complex": (c_float, [c_float, c_float, c_float]),
        "orc_srgb8": (c_int, [p, c_uint64, p]),
        "orc_bxdf_flags": (ctypes.c_uint32, [POINTER(OrcBxdf)]),
        "orc_bxdf_eval": (c_int, [POINTER(OrcBxdf), c_uint64, p, p, p, p]),
        "orc_bxdf_sample": (c_int, [POINTER(OrcBxdf), c_uint64, p, p, p, p, p, p, p]),
        "orc_fr_dielectric": (c_float, [c_float, c_float]),
        "orc_refract": (c_int, [p, p, c_float, POINTER(c_float), p]),
        "orc_tr_eval": (c_int, [c_float, c_float, c_uint64, p, p, p]),
        "orc_tr_sample": (c_int, [c_float, c_float, c_uint64, p, p, p]),
        "orc_light_pmf": (c_int, [POINTER(OrcScene), p, p, ctypes.c_uint32, POINTER(ctypes.c_uint32)]),
        "orc_shape_intersect": (c_int, [POINTER(OrcShape), p, p, c_int, p, c_uint64, p, p, p, p]),
        "orc_shape_sample": (c_int, [POINTER(OrcShape), p, p, c_int, p, p, c_uint64, p, p, p, p]),
        "orc_shape_pdf": (c_int, [POINTER(OrcShape), p, p, c_int, p, p, c_uint64, p]),
        "orc_directed": (c_int, [c_int, c_float, c_float, p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def _scene_struct(sc, pool_size):
    """Flatten a scene dict (see tests/scenes.py) into the C description; returns
    (struct, keepalive)."""
    tri = np.ascontiguousarray(sc["tri"], dtype=np.float32).reshape(-1, 9)
    mat = np.ascontiguousarray(sc["tri_material"], dtype=np.uint32)
    lig = np.ascontiguousarray(sc["tri_light"], dtype=np.int32)
    rev = np.ascontiguousarray(sc.get("tri_reverse", np.zeros(len(tri), np.uint8)), dtype=np.uint8)
    if sc.get("materials") is None:  # all diffuse, albedo [n][3]
        alb = np.ascontiguousarray(sc["albedo"], dtype=np.float32).reshape(-1, 3)
        marr = np.zeros(len(alb), dtype=np.dtype([("type", np.uint32), ("rgb", np.float32, 3), ("metal", np.uint32),
                                                  ("uroughness", np.float32), ("vroughness", np.float32),
                                                  ("remap_roughness", np.uint32), ("eta", np.float32),
                                                  ("eta_named", np.uint32)]))
        assert marr.dtype.itemsize == ctypes.sizeof(OrcMaterial)
        marr["rgb"] = alb
        mats = marr.ctypes.data_as(POINTER(OrcMaterial))
        materials = []
    else:
        marr = None
        materials = sc["materials"]
        mats = (OrcMaterial * len(materials))()
    for i, m in enumerate(materials):
        if m["type"] == "diffuse":
            mats[i].type = 0
            mats[i].rgb[:] = list(m["rgb"])
        else:
            mats[i].type = 1 if m["type"] == "conductor" else 2
            if m["type"] == "conductor":
                mats[i].metal = METALS[m["metal"]]
            else:
                eta = m.get("eta", 1.5)
                mats[i].eta_named = 1 if eta == "glass-BK7" else 0
                mats[i].eta = 0.0 if isinstance(eta, str) else float(eta)
            mats[i].uroughness = m.get("uroughness", m.get("roughness", 0.0))
            mats[i].vroughness = m.get("vroughness", m.get("roughness", 0.0))
            mats[i].remap_roughness = int(m.get("remap", True))
    lights = (OrcLight * max(1, len(sc["lights"])))()
    for i, l in enumerate(sc["lights"]):
        lights[i].rgb[:] = list(l["rgb"])
        lights[i].scale = l.get("scale", 1.0)
        lights[i].two_sided = int(l.get("two_sided", False))
    s = OrcScene()
    s.eye[:], s.look[:], s.up[:] = sc["eye"], sc["look"], sc["up"]
    s.fov = sc["fov"]
    s.width, s.height, s.spp, s.max_depth = sc["width"], sc["height"], sc["spp"], sc["max_depth"]
    s.pool_size = pool_size
    s.num_tris = len(tri)
    s.tri, s.tri_material, s.tri_light, s.tri_reverse = _p(tri), _p(mat), _p(lig), _p(rev)
    s.num_materials = len(materials) if marr is None else len(marr)
    s.materials = mats
    s.num_lights = len(sc["lights"])
    s.lights = lights
    xf = np.ascontiguousarray(sc.get("xforms", np.zeros((0, 4, 4))), dtype=np.float32).reshape(-1, 16)
    txf = sc.get("tri_xform")
    txf = None if txf is None else np.ascontiguousarray(txf, dtype=np.uint32)
    s.num_xforms = len(xf)
    s.xforms = _p(xf) if len(xf) else None
    s.tri_xform = _p(txf)
    sel = sc.get("select_scale")
    sel = None if sel is None else np.ascontiguousarray(sel, dtype=np.float32)
    s.select_scale = _p(sel)
    inf = sc.get("infinite_lights", [])
    infl = (OrcInfiniteLight * max(1, len(inf)))()
    for i, l in enumerate(inf):
        if l.get("rgb") is not None:
            infl[i].rgb[:] = list(l["rgb"])
            infl[i].has_rgb = 1
        infl[i].scale = l.get("scale", 1.0)
        infl[i].illuminance = l.get("illuminance", -1.0)
        infl[i].light_index = l.get("light_index", 0)
    s.num_infinite_lights = len(inf)
    s.infinite_lights = infl
    shp = sc.get("shapes") or []
    shapes = (OrcShape * max(1, len(shp)))()
    for i, q in enumerate(shp):
        r = q.get("radius", 1.0)
        shapes[i].type = SHAPE_KINDS[q["type"]]
        shapes[i].xform = 0xFFFFFFFF if q.get("xform") is None else q["xform"]
        shapes[i].radius = r
        shapes[i].zmin = q.get("zmin", -r)
        shapes[i].zmax = q.get("zmax", r)
        shapes[i].phimax = q.get("phimax", 360.0)
        shapes[i].height = q.get("height", 0.0)
        shapes[i].innerradius = q.get("innerradius", 0.0)
    tsh = sc.get("tri_shape")
    tsh = None if tsh is None else np.ascontiguousarray(tsh, dtype=np.uint32)
    xinv = sc.get("xforms_inv")
    xinv = None if xinv is None else np.ascontiguousarray(xinv, dtype=np.float32).reshape(-1, 16)
    s.num_shapes = len(shp)
    s.shapes = shapes
    s.tri_shape = _p(tsh)
    s.xforms_inv = _p(xinv)
    return s, (tri, mat, lig, rev, mats, lights, xf, txf, sel, marr, infl, shapes, tsh, xinv)


def render(sc, pool_size=2 * 1024 * 1024, threads=8):
    """Render a scene dict; returns (film_rgb [h][w][3], aux [h][w][4], stats dict)."""
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    s, keep = _scene_struct(sc, pool_size)
    h, w = sc["height"], sc["width"]
    rgb = np.zeros((h, w, 3), np.float32)
    aux = np.zeros((h, w, 4), np.float32)
    st = np.zeros(5, np.uint64)
    rc = lib.orc_render(ctypes.byref(s), threads, _p(rgb), _p(aux), _p(st))
    if rc != 0:
        raise RuntimeError(f"orc_render failed ({rc})")
    return rgb, aux, {"rounds": int(st[0]), "rays": int(st[1]), "shadow_rays": int(st[2]), "paths": int(st[3]),
                      "render_ms": int(st[4]) / 1e3}


INTEGRATORS = {"megakernelpath": 1, "ambientocclusion": 2, "surfacenormal": 3}


def render_megakernel(sc, integrator, threads=8, path_L=False):
    """Render a scene dict with a megakernel integrator ("megakernelpath",
    "ambientocclusion", "surfacenormal"); returns (film_rgb, aux, stats) and, with
    path_L, each path's spectral radiance [spp][h][w][4] as a fourth item."""
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    s, keep = _scene_struct(sc, 1)
    h, w = sc["height"], sc["width"]
    rgb = np.zeros((h, w, 3), np.float32)
    aux = np.zeros((h, w, 4), np.float32)
    pl = np.zeros((sc["spp"], h, w, 4), np.float32) if path_L else None
    st = np.zeros(5, np.uint64)
    rc = lib.orc_render_megakernel(ctypes.byref(s), INTEGRATORS[integrator], threads, _p(rgb), _p(aux), _p(pl),
                                   _p(st))
    if rc != 0:
        raise RuntimeError(f"orc_render_megakernel failed ({rc})")
    stats = {"rays": int(st[1]), "shadow_rays": int(st[2]), "paths": int(st[3]), "render_ms": int(st[4]) / 1e3}
    return (rgb, aux, stats, pl) if path_L else (rgb, aux, stats)


def render_path_L(sc, pool_size=2 * 1024 * 1024, threads=8):
    """The wavefront render's per-path spectral radiance and wavelength pdfs, each
    [spp][h][w][4] (path id = sample * W * H + pixel)."""
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    s, keep = _scene_struct(sc, pool_size)
    shape = (sc["spp"], sc["height"], sc["width"], 4)
    L = np.zeros(shape, np.float32)
    pdf = np.zeros(shape, np.float32)
    rc = lib.orc_render_path_L(ctypes.byref(s), threads, _p(L), _p(pdf))
    if rc != 0:
        raise RuntimeError(f"orc_render_path_L failed ({rc})")
    return L, pdf


def render_paths(sc, pool_size=2 * 1024 * 1024, threads=8):
    """Render a scene dict; returns per path id (sample * W * H + pixel) the round its
    first ray is traced in and the round it reaches the film, and the stats dict."""
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    s, keep = _scene_struct(sc, pool_size)
    n = sc["width"] * sc["height"] * sc["spp"]
    start = np.zeros(n, np.uint32)
    finish = np.zeros(n, np.uint32)
    st = np.zeros(5, np.uint64)
    rc = lib.orc_render_paths(ctypes.byref(s), threads, _p(start), _p(finish), _p(st))
    if rc != 0:
        raise RuntimeError(f"orc_render_paths failed ({rc})")
    return start, finish, {"rounds": int(st[0]), "rays": int(st[1]), "shadow_rays": int(st[2]), "paths": int(st[3])}


def render_grad(sc, target, pool_size=2 * 1024 * 1024, threads=8, hold_selection=False):
    """Loss sum((film - target)^2) and its FP64 path-replay gradient: returns
    (loss, d_rgb [num_materials][3], d_scale [num_lights], film)."""
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    s, keep = _scene_struct(sc, pool_size)
    h, w = sc["height"], sc["width"]
    tgt = np.ascontiguousarray(target, np.float32)
    assert tgt.shape == (h, w, 3)
    film = np.zeros((h, w, 3), np.float32)
    loss = ctypes.c_double(0)
    d_rgb = np.zeros((s.num_materials, 3), np.float64)
    d_scale = np.zeros(len(sc["lights"]), np.float64)
    st = np.zeros(5, np.uint64)
    rc = lib.orc_render_grad(ctypes.byref(s), threads, _p(tgt), _p(film), ctypes.byref(loss), _p(d_rgb),
                             _p(d_scale), _p(st), int(hold_selection))
    if rc != 0:
        raise RuntimeError(f"orc_render_grad failed ({rc})")
    return loss.value, d_rgb, d_scale, film


def rgb_to_sigmoid_jacobian(rgb):
    """d(c0, c1, c2)/d(r, g, b) of the RGB-to-spectrum table lookup, [3][3]."""
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    a = np.asarray(rgb, np.float32)
    out = np.zeros(9, np.float64)
    lib.orc_rgb_to_sigmoid_jacobian(_p(a), _p(out))
    return out.reshape(3, 3)


def rgb_to_sigmoid(rgb):
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    a = np.asarray(rgb, np.float32)
    out = np.zeros(3, np.float32)
    lib.orc_rgb_to_sigmoid(_p(a), _p(out))
    return out


def srgb_matrices():
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    a, b, w = np.zeros(9, np.float32), np.zeros(9, np.float32), np.zeros(2, np.float32)
    lib.orc_srgb_matrices(_p(a), _p(b), _p(w))
    return a.reshape(3, 3), b.reshape(3, 3), w


def filter_sample(u0, u1):
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    u = np.array([u0, u1], np.float32)
    out = np.zeros(3, np.float32)
    lib.orc_filter_sample(_p(u), _p(out))
    return out


def spectrum_eval(which, lam):
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    return float(lib.orc_spectrum_eval(int(which), float(lam)))


def _pt():
    lib = load()
    if not hasattr(lib, "_pt_ready"):
        _pt_sigs(lib)
        lib._pt_ready = True
    return lib


def sample_visible(u):
    lam, pdf = np.zeros(4, np.float32), np.zeros(4, np.float32)
    _pt().orc_sample_visible(float(u), _p(lam), _p(pdf))
    return lam, pdf


def sampler_sequence(spd, pixel, sample, kinds):
    k = np.asarray(kinds, np.int32)
    out = np.zeros(2 * len(k), np.float32)
    n = _pt().orc_sampler_sequence(spd, pixel, sample, _p(k), len(k), _p(out))
    return out[:n]


def camera_ray(sc, px, py):
    s, keep = _scene_struct(sc, 1)
    out = np.zeros(6, np.float32)
    if _pt().orc_camera_ray(ctypes.byref(s), float(px), float(py), _p(out)) != 0:
        raise RuntimeError("camera")
    return out[:3], out[3:]


def spherical_triangle(v, p, u):
    v = np.ascontiguousarray(v, np.float32).reshape(9)
    p = np.ascontiguousarray(p, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    b = np.zeros(3, np.float32)
    pdf = c_float(0)
    _pt().orc_spherical_triangle(_p(v), _p(p), _p(u), _p(b), ctypes.byref(pdf))
    return b, pdf.value


def invert_spherical_triangle(v, p, w):
    v = np.ascontiguousarray(v, np.float32).reshape(9)
    p = np.ascontiguousarray(p, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    u = np.zeros(2, np.float32)
    _pt().orc_invert_spherical_triangle(_p(v), _p(p), _p(w), _p(u))
    return u


def _shape(kind, radius=1.0, zmin=None, zmax=None, phimax=360.0, height=0.0, innerradius=0.0):
    q = OrcShape()
    q.type = SHAPE_KINDS[kind]
    q.radius = radius
    q.zmin = -radius if zmin is None else zmin
    q.zmax = radius if zmax is None else zmax
    q.phimax, q.height, q.innerradius = phimax, height, innerradius
    return q


def _xf(m):
    m = np.ascontiguousarray(m, np.float32).reshape(4, 4)
    return m, np.ascontiguousarray(np.linalg.inv(m.astype(np.float64)), np.float32)


def shape_intersect(kind, m, rays, reverse=False, inv=None, **params):
    """Closest hits of rays [n][8] with one shape placed by render_from_object m
    (inverse `inv`, default the FP64 inverse rounded): (hit, t, point, normal)."""
    q = _shape(kind, **params)
    m, inv0 = _xf(m)
    inv = inv0 if inv is None else np.ascontiguousarray(inv, np.float32)
    r = np.ascontiguousarray(rays, np.float32)
    n = len(r)
    t, pp, nn, hit = np.zeros(n, np.float32), np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), \
        np.zeros(n, np.uint8)
    _pt().orc_shape_intersect(ctypes.byref(q), _p(m), _p(inv), int(reverse), _p(r), n, _p(t), _p(pp), _p(nn), _p(hit))
    return hit.astype(bool), t, pp, nn


def shape_sample(kind, m, ctx, u, reverse=False, **params):
    """Shape::sample(ctx, u): ctx [n][9] = (p, n, ns), u [n][2] -> (ok, point, normal, pdf)."""
    q = _shape(kind, **params)
    m, inv = _xf(m)
    c = np.ascontiguousarray(ctx, np.float32)
    uu = np.ascontiguousarray(u, np.float32)
    n = len(c)
    pp, nn, pdf, ok = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.float32), \
        np.zeros(n, np.uint8)
    _pt().orc_shape_sample(ctypes.byref(q), _p(m), _p(inv), int(reverse), _p(c), _p(uu), n, _p(pp), _p(nn), _p(pdf),
                           _p(ok))
    return ok.astype(bool), pp, nn, pdf


def shape_pdf(kind, m, ctx, wi, reverse=False, **params):
    q = _shape(kind, **params)
    m, inv = _xf(m)
    c = np.ascontiguousarray(ctx, np.float32)
    w = np.ascontiguousarray(wi, np.float32)
    pdf = np.zeros(len(c), np.float32)
    _pt().orc_shape_pdf(ctypes.byref(q), _p(m), _p(inv), int(reverse), _p(c), _p(w), len(c), _p(pdf))
    return pdf


def directed(op, a, b=0.0):
    """[round down, round up] of a*b ("mul"), a/b ("div"), sqrt(a) ("sqrt"), a+b ("add")."""
    out = np.zeros(2, np.float32)
    _pt().orc_directed({"mul": 0, "div": 1, "sqrt": 2, "add": 3}[op], float(a), float(b), _p(out))
    return float(out[0]), float(out[1])


def fr_complex(cos_i, eta, k):
    return float(_pt().orc_fr_complex(float(cos_i), float(eta), float(k)))


def srgb8(rgb):
    a = np.ascontiguousarray(rgb, np.float32)
    out = np.zeros(a.shape[:-1] + (4,), np.uint8)
    _pt().orc_srgb8(_p(a), a.size // 3, _p(out))
    return out
