"""ctypes binding of libpbrtm.so (include/pbrtm.h) for tests and the benchmark.

The library is the product: every call below runs the HIP kernels for gfx950.
There is no CPU fallback; loading fails loudly when libpbrtm.so is missing.
Device-pointer calls take torch tensors (``.data_ptr()``) allocated on the
context's device.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libpbrtm.so")

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_DEVICE = -3
ERR_NOMEM = -4
ERR_UNSUPPORTED = -5
DEVICE_PTRS = 1
TRACE_BINARY = 2
NO_HIT = 0xFFFFFFFF

# Every symbol include/pbrtm.h declares.
EXPORTS = (
    "pbrtm_create", "pbrtm_destroy", "pbrtm_last_error", "pbrtm_device_count", "pbrtm_stream",
    "pbrtm_morton_sort", "pbrtm_bvh_build", "pbrtm_bvh_free", "pbrtm_bvh_bounds",
    "pbrtm_bvh_export_canonical", "pbrtm_bvh_intersect", "pbrtm_ao_bounce", "pbrtm_gen_soup",
    "pbrtm_gen_camera_rays", "pbrtm_scene_create", "pbrtm_scene_free", "pbrtm_render", "pbrtm_rgb_to_sigmoid",
    "pbrtm_srgb_matrices", "pbrtm_srgb8", "pbrtm_write_png", "pbrtm_pbrt_parse", "pbrtm_pbrt_desc",
    "pbrtm_pbrt_render_params", "pbrtm_pbrt_output", "pbrtm_pbrt_free", "pbrtm_render_batch",
    "pbrtm_render_grad", "pbrtm_integrator_from_name", "pbrtm_write_pfm", "pbrtm_bvh_intersect_noalloc",
)
BATCH_NO_GRAPH = 2
BATCH_AUTOTUNE = 4
RENDER_COUNT = 0x100
# pbrtm_render_params.flags: the integrator
INTEGRATORS = {"path": 0x00, "volpath": 0x00, "megakernelpath": 0x10, "ambientocclusion": 0x20,
               "surfacenormal": 0x30}


class BuildStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", c_double), ("ms_sort", c_double), ("ms_treelets", c_double), ("ms_top", c_double),
        ("ms_bottom", c_double), ("ms_layout", c_double), ("num_prims", c_uint64),
        ("num_treelets", c_uint64), ("top_levels", c_uint64), ("bottom_levels", c_uint64),
        ("small_top", c_uint64), ("small_bottom", c_uint64), ("wide_nodes", c_uint64), ("wide_depth", c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Material(ctypes.Structure):
    _fields_ = [("type", c_uint32), ("rgb", c_float * 3), ("metal", c_uint32), ("uroughness", c_float),
                ("vroughness", c_float), ("remap_roughness", c_uint32), ("eta", c_float), ("eta_named", c_uint32)]


class AreaLight(ctypes.Structure):
    _fields_ = [("rgb", c_float * 3), ("scale", c_float), ("two_sided", c_uint32)]


class InfiniteLight(ctypes.Structure):
    _fields_ = [("rgb", c_float * 3), ("has_rgb", c_uint32), ("scale", c_float), ("illuminance", c_float),
                ("light_index", c_uint32)]


class Shape(ctypes.Structure):
    _fields_ = [("type", c_uint32), ("xform", c_uint32), ("radius", c_float), ("zmin", c_float), ("zmax", c_float),
                ("phimax", c_float), ("height", c_float), ("innerradius", c_float)]


SHAPES = {"sphere": 1, "disk": 2, "instance": 3}


class SceneDesc(ctypes.Structure):
    _fields_ = [("eye", c_float * 3), ("look", c_float * 3), ("up", c_float * 3), ("fov", c_float),
                ("width", c_uint32), ("height", c_uint32), ("num_tris", c_uint64), ("tri", c_void_p),
                ("tri_material", c_void_p), ("tri_light", c_void_p), ("tri_reverse", c_void_p),
                ("num_materials", c_uint32), ("materials", POINTER(Material)), ("num_lights", c_uint32),
                ("lights", POINTER(AreaLight)), ("num_xforms", c_uint32), ("xforms", c_void_p),
                ("tri_xform", c_void_p), ("num_infinite_lights", c_uint32),
                ("infinite_lights", POINTER(InfiniteLight)), ("xforms_inv", c_void_p), ("num_shapes", c_uint32),
                ("shapes", POINTER(Shape)), ("tri_shape", c_void_p)]


class RenderParams(ctypes.Structure):
    _fields_ = [("spp", c_uint32), ("max_depth", c_uint32), ("pool_size", c_uint32), ("flags", c_uint32)]


class RenderStats(ctypes.Structure):
    _fields_ = [("ms_total", c_double), ("ms_trace", c_double), ("ms_shadow", c_double), ("ms_shade", c_double),
                ("ms_film", c_double), ("rounds", c_uint64), ("rays", c_uint64), ("shadow_rays", c_uint64),
                ("paths", c_uint64), ("ray_order", c_uint32), ("grid_quarters", c_uint32),
                ("closest_counts", c_uint64 * 3), ("shadow_counts", c_uint64 * 3)]

    def as_dict(self):
        out = {name: getattr(self, name) for name, _ in self._fields_}
        out["closest_counts"] = list(self.closest_counts)
        out["shadow_counts"] = list(self.shadow_counts)
        return out


class GradStats(ctypes.Structure):
    _fields_ = [("ms_forward", c_double), ("ms_adjoint", c_double), ("ms_replay", c_double), ("ms_total", c_double),
                ("rounds", c_uint64), ("rays", c_uint64), ("shadow_rays", c_uint64), ("paths", c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Job(ctypes.Structure):
    _fields_ = [("scene", c_void_p), ("params", RenderParams), ("film_rgb", c_void_p), ("film_aux", c_void_p),
                ("rounds", c_uint32), ("reserved", c_uint32)]


class BatchParams(ctypes.Structure):
    _fields_ = [("pool_size", c_uint32), ("rounds_per_launch", c_uint32), ("flags", c_uint32),
                ("solo_paths", c_uint32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("ms_total", c_double), ("ms_rounds", c_double), ("ms_schedule", c_double), ("ms_film", c_double),
                ("rounds", c_uint64), ("rays", c_uint64), ("shadow_rays", c_uint64), ("paths", c_uint64),
                ("pool_size", c_uint32), ("launches", c_uint32), ("solo_jobs", c_uint32), ("grid_quarters", c_uint32),
                ("tune_state", c_uint32), ("reserved", c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


METALS = {"Ag": 0, "Al": 1, "Au": 2, "Cu": 3}


class PbrtmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pbrtm error {code}: {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH):
    """Load libpbrtm.so and declare the C signatures (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    p = c_void_p
    sig = {
        "pbrtm_create": (c_int, [POINTER(p), c_int]),
        "pbrtm_destroy": (None, [p]),
        "pbrtm_last_error": (c_char_p, [p]),
        "pbrtm_device_count": (c_int, [POINTER(c_int)]),
        "pbrtm_stream": (p, [p]),
        "pbrtm_morton_sort": (c_int, [p, p, c_uint64, c_uint32, p, p]),
        "pbrtm_bvh_build": (c_int, [p, p, c_uint64, p, c_uint64, c_uint32, POINTER(p), POINTER(BuildStats)]),
        "pbrtm_bvh_free": (None, [p]),
        "pbrtm_bvh_bounds": (c_int, [p, p]),
        "pbrtm_bvh_export_canonical": (c_int, [p, p, POINTER(c_uint64)]),
        "pbrtm_bvh_intersect": (c_int, [p, p, c_uint64, c_int, c_uint32, p, p, p, p]),
        "pbrtm_bvh_intersect_noalloc": (c_int, [p, p, c_uint64, c_int, p, p, p]),
        "pbrtm_ao_bounce": (c_int, [p, p, p, p, c_uint64, c_uint32, p]),
        "pbrtm_gen_soup": (c_int, [c_uint64, c_uint64, c_float, c_float, c_int, p]),
        "pbrtm_gen_camera_rays": (c_int, [c_uint32, c_uint32, p, p, p, c_float, p]),
        "pbrtm_scene_create": (c_int, [p, POINTER(SceneDesc), POINTER(p)]),
        "pbrtm_scene_free": (None, [p]),
        "pbrtm_render": (c_int, [p, p, POINTER(RenderParams), p, p, POINTER(RenderStats)]),
        "pbrtm_integrator_from_name": (c_int, [ctypes.c_char_p, POINTER(c_uint32)]),
        "pbrtm_rgb_to_sigmoid": (c_int, [p, p]),
        "pbrtm_srgb_matrices": (c_int, [p, p, p]),
        "pbrtm_srgb8": (c_int, [p, c_uint64, p]),
        "pbrtm_write_png": (c_int, [p, c_uint32, c_uint32, c_char_p]),
        "pbrtm_write_pfm": (c_int, [p, c_uint32, c_uint32, c_char_p]),
        "pbrtm_pbrt_parse": (c_int, [c_char_p, POINTER(p), c_char_p, c_uint64]),
        "pbrtm_pbrt_desc": (POINTER(SceneDesc), [p]),
        "pbrtm_pbrt_render_params": (c_int, [p, POINTER(RenderParams)]),
        "pbrtm_pbrt_output": (c_char_p, [p]),
        "pbrtm_pbrt_free": (None, [p]),
        "pbrtm_render_batch": (c_int, [p, POINTER(Job), c_uint32, POINTER(BatchParams), POINTER(BatchStats)]),
        "pbrtm_render_grad": (c_int, [p, p, POINTER(RenderParams), p, p, POINTER(c_double), p, p,
                                      POINTER(GradStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data_as(c_void_p)
    return c_void_p(int(a.data_ptr()))  # torch tensor on the device


def gen_soup(n: int, seed: int = 0x5EED, mu: float = -6.0, sigma: float = 0.5, threads: int = 8) -> np.ndarray:
    """Seeded triangle soup of the HLBVH microbenchmark, [n][3][3] float32."""
    lib = load()
    out = np.empty((n, 3, 3), np.float32)
    rc = lib.pbrtm_gen_soup(seed, n, mu, sigma, threads, _ptr(out))
    if rc != OK:
        raise PbrtmError(rc, "gen_soup")
    return out


def gen_camera_rays(width: int, height: int, eye=(0.5, 0.5, -1.5), look=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0),
                    fov_deg: float = 45.0) -> np.ndarray:
    lib = load()
    out = np.empty((height * width, 8), np.float32)
    e, l, u = (np.asarray(v, np.float32) for v in (eye, look, up))
    rc = lib.pbrtm_gen_camera_rays(width, height, _ptr(e), _ptr(l), _ptr(u), fov_deg, _ptr(out))
    if rc != OK:
        raise PbrtmError(rc, "gen_camera_rays")
    return out


def device_count() -> int:
    c = c_int(0)
    load().pbrtm_device_count(ctypes.byref(c))
    return c.value


class Context:
    """One HIP device and stream (pbrtm_create)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = c_void_p()
        rc = self.lib.pbrtm_create(ctypes.byref(h), device)
        if rc != OK:
            raise PbrtmError(rc, f"cannot create a context on device {device}")
        self.h = h
        self._handles = weakref.WeakSet()

    def close(self):
        """Frees every aggregate built on this context, then the context."""
        if self.h:
            for b in list(self._handles):
                b.free()
            self.lib.pbrtm_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return self.lib.pbrtm_stream(self.h) or 0

    def check(self, rc):
        if rc != OK:
            raise PbrtmError(rc, self.lib.pbrtm_last_error(self.h).decode())

    def morton_sort(self, tri, device_ptrs=False, codes=None, order=None):
        n = tri.shape[0]
        if not device_ptrs:
            codes = np.empty(n, np.uint32)
            order = np.empty(n, np.uint32)
        self.check(self.lib.pbrtm_morton_sort(self.h, _ptr(tri), n, DEVICE_PTRS if device_ptrs else 0,
                                              _ptr(codes), _ptr(order)))
        return codes, order

    def scene(self, sc) -> "Scene":
        return Scene(self, sc)

    def render_batch(self, jobs, pool_size=0, rounds_per_launch=0, graph=True, aux=False, solo_paths=0,
                     autotune=False):
        """jobs: [(Scene, spp, max_depth, pool_size)]. Returns ([film_rgb], [film_aux or None],
        [one-by-one round count per job], stats dict). autotune=True: the library picks the
        shared pool size and traversal grid from measured time across calls of the same batch."""
        arr = (Job * len(jobs))()
        rgbs, auxs = [], []
        for k, (scene, spp, depth, vpool) in enumerate(jobs):
            rgb = np.zeros((scene.height, scene.width, 3), np.float32)
            ax = np.zeros((scene.height, scene.width, 4), np.float32) if aux else None
            rgbs.append(rgb)
            auxs.append(ax)
            arr[k].scene = scene.h
            arr[k].params = RenderParams(spp, depth, vpool, 0)
            arr[k].film_rgb = _ptr(rgb)
            arr[k].film_aux = _ptr(ax)
        bp = BatchParams(pool_size, rounds_per_launch, (0 if graph else BATCH_NO_GRAPH) | (BATCH_AUTOTUNE if autotune else 0),
                         solo_paths)
        st = BatchStats()
        self.check(self.lib.pbrtm_render_batch(self.h, arr, len(jobs), ctypes.byref(bp), ctypes.byref(st)))
        return rgbs, auxs, [arr[k].rounds for k in range(len(jobs))], st.as_dict()

    def scene_from_file(self, f: "PbrtFile") -> "Scene":
        """A scene straight from a parsed .pbrt file's description (no Python copy)."""
        return Scene(self, f)

    def build(self, verts, idx=None, device_ptrs=False) -> "Bvh":
        stats = BuildStats()
        h = c_void_p()
        if idx is None:
            nt = verts.shape[0]
            nv = 3 * nt
        else:
            nt = idx.shape[0]
            nv = verts.shape[0]
        self.check(self.lib.pbrtm_bvh_build(self.h, _ptr(verts), nv, _ptr(idx), nt,
                                            DEVICE_PTRS if device_ptrs else 0, ctypes.byref(h),
                                            ctypes.byref(stats)))
        bvh = Bvh(self, h, nt, stats, keep=verts if device_ptrs else None)
        self._handles.add(bvh)
        return bvh


def rgb_to_sigmoid(rgb) -> np.ndarray:
    """Sigmoid-polynomial coefficients (c0, c1, c2) of an RGB (host only, no GPU needed)."""
    lib = load()
    a = np.ascontiguousarray(rgb, np.float32)
    out = np.zeros(3, np.float32)
    rc = lib.pbrtm_rgb_to_sigmoid(_ptr(a), _ptr(out))
    if rc != OK:
        raise PbrtmError(rc, "rgb_to_sigmoid")
    return out


def srgb8(rgb) -> np.ndarray:
    """Linear RGB [..][3] -> sRGB RGBA8 [..][4] (the film's PNG encoding; host only)."""
    lib = load()
    a = np.ascontiguousarray(rgb, np.float32)
    n = a.size // 3
    out = np.zeros(a.shape[:-1] + (4,), np.uint8)
    rc = lib.pbrtm_srgb8(_ptr(a), n, _ptr(out))
    if rc != OK:
        raise PbrtmError(rc, "srgb8")
    return out


def write_pfm(rgb, path):
    """Write a linear RGB film [h][w][3] as a little-endian float PFM (host only)."""
    lib = load()
    a = np.ascontiguousarray(rgb, np.float32)
    h, w = a.shape[:2]
    rc = lib.pbrtm_write_pfm(_ptr(a), w, h, os.fsencode(path))
    if rc != OK:
        raise PbrtmError(rc, f"write_pfm {path}")


def write_png(rgb, path):
    """Write a linear RGB film [h][w][3] as an sRGB RGBA8 PNG (host only)."""
    lib = load()
    a = np.ascontiguousarray(rgb, np.float32)
    h, w = a.shape[:2]
    rc = lib.pbrtm_write_png(_ptr(a), w, h, os.fsencode(path))
    if rc != OK:
        raise PbrtmError(rc, f"write_png {path}")


def srgb_matrices():
    """(xyz_from_rgb, rgb_from_xyz, white xy) of the sRGB colour space (host only)."""
    lib = load()
    a, b, w = np.zeros(9, np.float32), np.zeros(9, np.float32), np.zeros(2, np.float32)
    lib.pbrtm_srgb_matrices(_ptr(a), _ptr(b), _ptr(w))
    return a.reshape(3, 3), b.reshape(3, 3), w


def scene_desc(sc):
    """Flatten a scene dict {eye, look, up, fov, width, height, tri [n][9], tri_material, tri_light,
    [tri_reverse], materials [{type: diffuse|conductor, ...}], lights [{rgb, scale, two_sided}]} into
    pbrtm_scene_desc; returns (desc, keepalive)."""
    tri = np.ascontiguousarray(sc["tri"], dtype=np.float32).reshape(-1, 9)
    mat = np.ascontiguousarray(sc["tri_material"], dtype=np.uint32)
    lig = np.ascontiguousarray(sc["tri_light"], dtype=np.int32)
    rev = np.ascontiguousarray(sc.get("tri_reverse", np.zeros(len(tri), np.uint8)), dtype=np.uint8)
    if sc.get("materials") is None:  # all diffuse, albedo [n][3]
        alb = np.ascontiguousarray(sc["albedo"], dtype=np.float32).reshape(-1, 3)
        arr = np.zeros(len(alb), dtype=_MATERIAL_DTYPE)
        arr["rgb"] = alb
        mats = arr.ctypes.data_as(POINTER(Material))
        nmat = len(alb)
    else:
        mats, nmat, arr = _materials(sc["materials"]), len(sc["materials"]), None
    lights = (AreaLight * max(1, len(sc["lights"])))()
    for i, l in enumerate(sc["lights"]):
        lights[i].rgb[:] = list(l["rgb"])
        lights[i].scale = l.get("scale", 1.0)
        lights[i].two_sided = int(l.get("two_sided", False))
    d = SceneDesc()
    d.eye[:], d.look[:], d.up[:] = sc["eye"], sc["look"], sc["up"]
    d.fov = sc["fov"]
    d.width, d.height = sc["width"], sc["height"]
    d.num_tris = len(tri)
    d.tri, d.tri_material, d.tri_light, d.tri_reverse = _ptr(tri), _ptr(mat), _ptr(lig), _ptr(rev)
    d.num_materials = nmat
    d.materials = mats
    d.num_lights = len(sc["lights"])
    d.lights = lights
    return _scene_desc_tail(sc, d, (tri, mat, lig, rev, mats, lights, arr))


_MATERIAL_DTYPE = np.dtype([("type", np.uint32), ("rgb", np.float32, 3), ("metal", np.uint32),
                            ("uroughness", np.float32), ("vroughness", np.float32), ("remap_roughness", np.uint32),
                            ("eta", np.float32), ("eta_named", np.uint32)])
assert _MATERIAL_DTYPE.itemsize == ctypes.sizeof(Material)


def _materials(materials):
    mats = (Material * max(1, len(materials)))()
    for i, m in enumerate(materials):
        if m["type"] == "diffuse":
            mats[i].type = 0
            mats[i].rgb[:] = list(m["rgb"])
        elif m["type"] in ("conductor", "dielectric"):
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
        else:
            raise ValueError(f"unknown material type {m['type']}")
    return mats


def _scene_desc_tail(sc, d, keep):
    xf = np.ascontiguousarray(sc.get("xforms", np.zeros((0, 4, 4))), dtype=np.float32).reshape(-1, 16)
    txf = sc.get("tri_xform")
    txf = None if txf is None else np.ascontiguousarray(txf, dtype=np.uint32)
    d.num_xforms = len(xf)
    d.xforms = _ptr(xf) if len(xf) else None
    d.tri_xform = _ptr(txf)
    # "infinite_lights": [{"rgb": (r, g, b) or None (the colour space illuminant), "scale",
    # "illuminance", "light_index": emissive triangles created before it}]
    inf = sc.get("infinite_lights", [])
    infl = (InfiniteLight * max(1, len(inf)))()
    for i, l in enumerate(inf):
        if l.get("rgb") is not None:
            infl[i].rgb[:] = list(l["rgb"])
            infl[i].has_rgb = 1
        infl[i].scale = l.get("scale", 1.0)
        infl[i].illuminance = l.get("illuminance", -1.0)
        infl[i].light_index = l.get("light_index", 0)
    d.num_infinite_lights = len(inf)
    d.infinite_lights = infl
    # analytic shapes: "shapes": [{"type": "sphere" | "disk" | "instance", "xform": index or None,
    # "radius", "zmin", "zmax", "phimax" (degrees), "height", "innerradius"}], "tri_shape": [n]
    # (0xFFFFFFFF = a triangle), "xforms_inv": [num_xforms][4][4] (optional)
    shp = sc.get("shapes") or []
    shapes = (Shape * max(1, len(shp)))()
    for i, q in enumerate(shp):
        r = q.get("radius", 1.0)
        shapes[i].type = SHAPES[q["type"]]
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
    d.num_shapes = len(shp)
    d.shapes = shapes
    d.tri_shape = _ptr(tsh)
    d.xforms_inv = _ptr(xinv)
    return d, keep + (xf, txf, infl, shapes, tsh, xinv)


class PbrtFile:
    """A parsed .pbrt file (pbrtm_pbrt_parse): its scene description, render
    parameters and output name. to_dict() gives the same dict form Scene takes."""

    def __init__(self, path):
        self.lib = load()
        h = c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = self.lib.pbrtm_pbrt_parse(os.fsencode(path), ctypes.byref(h), err, 512)
        if rc != OK:
            raise PbrtmError(rc, err.value.decode())
        self.h = h
        p = RenderParams()
        self.lib.pbrtm_pbrt_render_params(h, ctypes.byref(p))
        self.spp, self.max_depth = p.spp, p.max_depth
        flags = p.flags & 0xF0
        self.integrator = next(k for k, v in INTEGRATORS.items() if v == flags and k != "volpath")
        self.output = self.lib.pbrtm_pbrt_output(h).decode()

    def free(self):
        if self.h:
            self.lib.pbrtm_pbrt_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def to_dict(self):
        d = self.lib.pbrtm_pbrt_desc(self.h).contents
        n = d.num_tris

        def arr(ptr, ctype, count, dtype):
            if not ptr or count == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ctypes.cast(ptr, POINTER(ctype)), shape=(count,)).astype(dtype).copy()

        mats = []
        names = {v: k for k, v in METALS.items()}
        for i in range(d.num_materials):
            m = d.materials[i]
            if m.type == 0:
                mats.append({"type": "diffuse", "rgb": tuple(m.rgb)})
            elif m.type == 1:
                mats.append({"type": "conductor", "metal": names[m.metal], "uroughness": m.uroughness,
                             "vroughness": m.vroughness, "remap": bool(m.remap_roughness)})
            else:
                mats.append({"type": "dielectric", "eta": "glass-BK7" if m.eta_named else m.eta,
                             "uroughness": m.uroughness, "vroughness": m.vroughness, "remap": bool(m.remap_roughness)})
        lights = [{"rgb": tuple(d.lights[i].rgb), "scale": d.lights[i].scale, "two_sided": bool(d.lights[i].two_sided)}
                  for i in range(d.num_lights)]
        out = {
            "eye": tuple(d.eye), "look": tuple(d.look), "up": tuple(d.up), "fov": d.fov,
            "width": d.width, "height": d.height, "spp": self.spp, "max_depth": self.max_depth,
            "tri": arr(d.tri, c_float, 9 * n, np.float32).reshape(-1, 9),
            "tri_material": arr(d.tri_material, c_uint32, n, np.uint32),
            "tri_light": arr(d.tri_light, ctypes.c_int32, n, np.int32),
            "tri_reverse": arr(d.tri_reverse, ctypes.c_uint8, n, np.uint8),
            "materials": mats, "lights": lights,
        }
        if d.num_xforms:
            out["xforms"] = arr(d.xforms, c_float, 16 * d.num_xforms, np.float32).reshape(-1, 4, 4)
            out["tri_xform"] = arr(d.tri_xform, c_uint32, n, np.uint32)
            if d.xforms_inv:
                out["xforms_inv"] = arr(d.xforms_inv, c_float, 16 * d.num_xforms, np.float32).reshape(-1, 4, 4)
        if d.num_shapes:
            kinds = {v: k for k, v in SHAPES.items()}
            out["shapes"] = [
                {"type": kinds[q.type], "xform": None if q.xform == 0xFFFFFFFF else q.xform, "radius": q.radius,
                 "zmin": q.zmin, "zmax": q.zmax, "phimax": q.phimax, "height": q.height,
                 "innerradius": q.innerradius}
                for q in (d.shapes[i] for i in range(d.num_shapes))]
            out["tri_shape"] = arr(d.tri_shape, c_uint32, n, np.uint32)
        if d.num_infinite_lights:
            out["infinite_lights"] = [
                {"rgb": tuple(l.rgb) if l.has_rgb else None, "scale": l.scale, "illuminance": l.illuminance,
                 "light_index": l.light_index}
                for l in (d.infinite_lights[i] for i in range(d.num_infinite_lights))]
        return out


class Scene:
    """A scene on a context (pbrtm_scene_create); render() runs the wavefront path integrator."""

    def __init__(self, ctx: "Context", sc):
        self.ctx = ctx
        h = c_void_p()
        if isinstance(sc, PbrtFile):
            dp = ctx.lib.pbrtm_pbrt_desc(sc.h)
            self.width, self.height = dp.contents.width, dp.contents.height
            self.num_materials, self.num_lights = dp.contents.num_materials, dp.contents.num_lights
            ctx.check(ctx.lib.pbrtm_scene_create(ctx.h, dp, ctypes.byref(h)))
        else:
            self.width, self.height = sc["width"], sc["height"]
            d, keep = scene_desc(sc)
            self.num_materials, self.num_lights = d.num_materials, d.num_lights
            ctx.check(ctx.lib.pbrtm_scene_create(ctx.h, ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        ctx._handles.add(self)

    def free(self):
        if self.h:
            self.ctx.lib.pbrtm_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def render_grad(self, target, spp=4, max_depth=5, pool_size=0, integrator="path"):
        """Loss sum((film - target)^2) and its gradient (pbrtm_render_grad): returns
        (loss, d_albedo [num_materials][3], d_scale [num_lights], film, stats dict)."""
        tgt = np.ascontiguousarray(target, np.float32)
        assert tgt.shape == (self.height, self.width, 3)
        p = RenderParams(spp, max_depth, pool_size, INTEGRATORS[integrator])
        film = np.zeros((self.height, self.width, 3), np.float32)
        d_alb = np.zeros((self.num_materials, 3), np.float32)
        d_sc = np.zeros(max(1, self.num_lights), np.float32)
        loss = c_double(0)
        st = GradStats()
        self.ctx.check(self.ctx.lib.pbrtm_render_grad(self.ctx.h, self.h, ctypes.byref(p), _ptr(tgt), _ptr(film),
                                                      ctypes.byref(loss), _ptr(d_alb), _ptr(d_sc), ctypes.byref(st)))
        return loss.value, d_alb, d_sc[:self.num_lights], film, st.as_dict()

    def render(self, spp=4, max_depth=5, pool_size=0, film_rgb=None, film_aux=None, device_ptrs=False,
               integrator="path", count=False):
        """Returns (film_rgb [h][w][3], film_aux [h][w][4], stats dict). integrator: "path"
        (wavefront), "megakernelpath", "ambientocclusion" or "surfacenormal". count=True
        traces with the counting traversal: stats["closest_counts"] / ["shadow_counts"]
        then hold [nodes visited, box tests, primitive tests]."""
        p = RenderParams(spp, max_depth, pool_size, (DEVICE_PTRS if device_ptrs else 0) | INTEGRATORS[integrator]
                         | (RENDER_COUNT if count else 0))
        st = RenderStats()
        if not device_ptrs:
            film_rgb = np.zeros((self.height, self.width, 3), np.float32)
            film_aux = np.zeros((self.height, self.width, 4), np.float32)
        self.ctx.check(self.ctx.lib.pbrtm_render(self.ctx.h, self.h, ctypes.byref(p), _ptr(film_rgb), _ptr(film_aux),
                                                 ctypes.byref(st)))
        return film_rgb, film_aux, st.as_dict()


class Bvh:
    """An aggregate handle (pbrtm_bvh_build); mirrors HLBVH's query surface."""

    def __init__(self, ctx: Context, h, n: int, stats: BuildStats, keep=None):
        self.ctx, self.h, self.n, self.stats = ctx, h, n, stats
        self._keep = keep  # caller's device triangles referenced by the handle

    def free(self):
        if self.h:
            self.ctx.lib.pbrtm_bvh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def bounds(self) -> np.ndarray:
        out = np.empty(6, np.float32)
        self.ctx.check(self.ctx.lib.pbrtm_bvh_bounds(self.h, _ptr(out)))
        return out

    def export_canonical(self) -> bytes:
        ln = c_uint64(0)
        self.ctx.check(self.ctx.lib.pbrtm_bvh_export_canonical(self.h, None, ctypes.byref(ln)))
        buf = np.empty(ln.value, np.uint8)
        self.ctx.check(self.ctx.lib.pbrtm_bvh_export_canonical(self.h, _ptr(buf), ctypes.byref(ln)))
        return buf.tobytes()

    def intersect(self, rays, any_hit=False, device_ptrs=False, prim=None, t=None, bary=None, want_bary=True,
                  stats=None, binary=False):
        """Closest (or any) hit per ray. `stats`, a uint64[3] numpy array, receives
        [nodes visited, box tests, triangle tests] when given. binary=True walks the
        binary build tree in the reference's order instead of the 8-wide layout."""
        n = rays.shape[0]
        if not device_ptrs:
            prim = np.empty(n, np.uint32)
            t = np.empty(n, np.float32)
            bary = np.empty((n, 3), np.float32) if want_bary else None
        flags = (DEVICE_PTRS if device_ptrs else 0) | (TRACE_BINARY if binary else 0)
        self.ctx.check(self.ctx.lib.pbrtm_bvh_intersect(self.h, _ptr(rays), n, int(any_hit), flags, _ptr(prim),
                                                        _ptr(t), _ptr(bary), _ptr(stats)))
        return prim, t, bary

    def intersect_noalloc(self, rays, prim, t=None, bary=None, any_hit=False):
        """pbrtm_bvh_intersect_noalloc: device tensors in and out, no device
        allocation inside the library (its fixed scratch array instead)."""
        self.ctx.check(self.ctx.lib.pbrtm_bvh_intersect_noalloc(self.h, _ptr(rays), rays.shape[0], int(any_hit),
                                                                _ptr(prim), _ptr(t), _ptr(bary)))
        return prim, t, bary

    def ao_bounce(self, rays, prim, bary, device_ptrs=False, out=None):
        n = rays.shape[0]
        if not device_ptrs:
            out = np.empty((n, 8), np.float32)
        self.ctx.check(self.ctx.lib.pbrtm_ao_bounce(self.h, _ptr(rays), _ptr(prim), _ptr(bary), n,
                                                    DEVICE_PTRS if device_ptrs else 0, _ptr(out)))
        return out
