"""CPU-side checks of the C ABI library: it loads, exports every symbol that
include/pbrtm.h declares, rejects bad arguments without a GPU, and its host
workload generators are deterministic."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import pbrtm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pbrtm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbrtm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = pbrtm.load()
    names = declared_symbols()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(pbrtm.EXPORTS)


def test_null_arguments_are_errors_not_crashes():
    lib = pbrtm.load()
    assert lib.pbrtm_create(None, 0) == pbrtm.ERR_INVALID
    assert lib.pbrtm_bvh_bounds(None, None) == pbrtm.ERR_INVALID
    ln = ctypes.c_uint64(0)
    assert lib.pbrtm_bvh_export_canonical(None, None, ctypes.byref(ln)) == pbrtm.ERR_INVALID
    assert lib.pbrtm_bvh_intersect(None, None, 0, 0, 0, None, None, None, None) == pbrtm.ERR_INVALID
    assert lib.pbrtm_last_error(None) == b"null context"


def test_gen_soup_deterministic_and_thread_independent():
    a = pbrtm.gen_soup(200000, seed=0x5EED, threads=1)
    b = pbrtm.gen_soup(200000, seed=0x5EED, threads=7)
    assert np.array_equal(a, b)
    assert not np.array_equal(a[:100], pbrtm.gen_soup(100, seed=1))
    c = a.mean(axis=1)
    assert np.all((c > -0.05) & (c < 1.05))
    # the centroid draw of triangle i is the first three floats of PCG32 draws 16i..
    u = O.pcg32(3, seq=0x5EED, mode=1, advance=16 * 5).astype(np.float64) * 2.0 ** -32
    edge = np.linalg.norm(a[:, 1] - a[:, 0], axis=1)
    assert 1e-4 < np.median(edge) < 1e-2
    lens = np.linalg.norm(a - a.mean(axis=1, keepdims=True), axis=2)
    assert lens.max() < 0.2
    # vertex = c + l*u with |u| = 1, so the mean of the three vertices is near c
    assert np.allclose(a[5].mean(axis=0), u, atol=0.05)


def test_gen_camera_rays_geometry():
    rays = pbrtm.gen_camera_rays(64, 36)
    assert rays.shape == (64 * 36, 8)
    d = rays[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.all(np.isinf(rays[:, 3]))
    centre = d.reshape(36, 64, 3)[17:19, 31:33].mean(axis=(0, 1))
    assert np.allclose(centre / np.linalg.norm(centre), [0, 0, 1], atol=1e-3)
    # vertical field of view 45 degrees on a pinhole: y/z of row 0 is (1 - 1/36) tan(22.5 deg)
    top = d.reshape(36, 64, 3)[0, 32]
    assert abs(top[1] / top[2] - (1 - 1 / 36) * np.tan(np.radians(22.5))) < 1e-6


def test_cpp_interface_compiles_and_fails_cleanly_without_gpu():
    """host/hlbvh.hpp compiles against the C ABI; without a device the example
    reports an error instead of crashing."""
    import subprocess

    exe = os.path.join(ROOT, "pbr-rust_amd", "hlbvh_example")
    assert os.path.exists(exe)
    if pbrtm.device_count() == 0:
        out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert out.returncode == 1 and "pbrtm_create failed" in out.stderr


def test_pfm_round_trip(tmp_path):
    """pbrtm_write_pfm: header "PF", width height, scale -1 (little-endian), rows
    bottom-up; the floats come back bit for bit."""
    import numpy as np
    rgb = np.random.default_rng(7).standard_normal((5, 7, 3)).astype(np.float32)
    rgb[0, 0] = [np.inf, -0.0, 1e-42]
    p = tmp_path / "f.pfm"
    pbrtm.write_pfm(rgb, str(p))
    data = p.read_bytes()
    head = b"PF\n7 5\n-1.0\n"
    assert data[:len(head)] == head and len(data) == len(head) + rgb.size * 4
    back = np.frombuffer(data[len(head):], "<f4").reshape(5, 7, 3)[::-1]
    assert np.array_equal(back.view(np.uint32), rgb.view(np.uint32))
    with pytest.raises(pbrtm.PbrtmError):
        pbrtm.write_pfm(rgb, str(tmp_path / "missing" / "f.pfm"))
