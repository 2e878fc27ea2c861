"""pbrtm_bvh_intersect_noalloc through the C ABI against pbrtm_bvh_intersect.

The no-alloc entry point keeps its ray queue, error word and stack spill in the
library's fixed scratch array; its hits must equal the allocating entry point's
bit for bit. Two calls in a row, the second with fewer rays (the scratch keeps
the first call's contents), closest and any hit, on a soup and on a tree deep
enough to spill the traversal stack into the scratch.
"""
import numpy as np
import pytest

import pbrtm

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


class DeviceArray:
    """Device memory from the HIP runtime libpbrtm.so itself uses (the same
    libamdhip64.so.7 in this process), for the device-pointer entry points."""

    _hip = None

    def __init__(self, shape, dtype, host=None):
        import ctypes

        pbrtm.load()
        if DeviceArray._hip is None:
            DeviceArray._hip = ctypes.CDLL("libamdhip64.so.7")
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        assert self._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.nbytes, 4))) == 0
        if host is not None:
            h = np.ascontiguousarray(host, self.dtype)
            assert self._hip.hipMemcpy(self.ptr, h.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes),
                                       1) == 0

    def data_ptr(self):
        return self.ptr.value

    def numpy(self):
        import ctypes

        out = np.empty(self.shape, self.dtype)
        assert self._hip.hipDeviceSynchronize() == 0
        assert self._hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes),
                                   2) == 0
        return out

    def free(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = None


def _both(bvh, rays_np, any_hit):
    n = rays_np.shape[0]
    rays = DeviceArray(rays_np.shape, np.float32, rays_np)
    outs = []
    for noalloc in (False, True):
        prim = DeviceArray((n,), np.uint32)
        t = DeviceArray((n,), np.float32)
        bary = DeviceArray((n, 3), np.float32)
        if noalloc:
            bvh.intersect_noalloc(rays, prim, t, bary, any_hit=any_hit)
        else:
            bvh.intersect(rays, any_hit=any_hit, device_ptrs=True, prim=prim, t=t, bary=bary)
        outs.append((prim.numpy(), t.numpy(), bary.numpy()))
        for a in (prim, t, bary):
            a.free()
    rays.free()
    return outs


@pytest.mark.parametrize("any_hit", [False, True])
def test_noalloc_equals_intersect_two_calls(ctx, any_hit):
    tri = pbrtm.gen_soup(200_000, seed=0x5EED)
    bvh = ctx.build(tri)
    rays = pbrtm.gen_camera_rays(480, 270)
    # the second call is smaller: 10,007 rays from the middle rows of the image
    for lo, hi in ((0, rays.shape[0]), (60_000, 70_007)):
        log(f"noalloc soup any_hit={any_hit} rays [{lo}, {hi})")
        (p0, t0, b0), (p1, t1, b1) = _both(bvh, np.ascontiguousarray(rays[lo:hi]), any_hit)
        assert (p0 != pbrtm.NO_HIT).mean() > 0.05
        assert np.array_equal(p0, p1)
        if not any_hit:
            assert np.array_equal(t0, t1) and np.array_equal(b0, b1)
    bvh.free()


def test_noalloc_deep_tree_spill(ctx):
    import scenes

    tri = scenes.deep_chain()
    bvh = ctx.build(tri)
    assert bvh.stats.wide_depth > 12
    for n, seed in ((50_000, 5), (3_001, 6)):
        log(f"noalloc deep chain n={n}")
        rays = scenes.chain_rays(n, seed=seed)
        (p0, t0, b0), (p1, t1, b1) = _both(bvh, rays, False)
        assert (p0 != pbrtm.NO_HIT).all()
        assert np.array_equal(p0, p1) and np.array_equal(t0, t1) and np.array_equal(b0, b1)
    bvh.free()


def test_noalloc_rejects_oversize(ctx):
    tri = pbrtm.gen_soup(1000, seed=1)
    bvh = ctx.build(tri)
    n = (1 << 22) + 1
    rays = DeviceArray((n, 8), np.float32)
    prim = DeviceArray((n,), np.uint32)
    with pytest.raises(pbrtm.PbrtmError) as e:
        bvh.intersect_noalloc(rays, prim)
    assert e.value.code == pbrtm.ERR_INVALID
    rays.free()
    prim.free()
    bvh.free()
