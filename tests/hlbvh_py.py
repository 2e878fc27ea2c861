"""An independent, deliberately naive Python/numpy-FP32 restatement of the
reference HLBVH build (accelerator/hlbvh.cu) for tiny inputs. It shares no code
with the C++ oracle: the Morton interleave is a bit loop, the sort is Python's
sorted(), splits are set comprehensions. Used to cross-check the oracle's
canonical bytes."""
from __future__ import annotations

import struct

import numpy as np

F = np.float32


def f2u_sat(f: np.float32) -> int:
    if not (f > 0):
        return 0
    if f >= F(4294967296.0):
        return 0xFFFFFFFF
    return int(f)


def interleave(x: int, y: int, z: int) -> int:
    code = 0
    for v, shift in ((x, 0), (y, 1), (z, 2)):
        if v == 1024:
            v = 1023
        for bit in range(10):
            if (v >> bit) & 1:
                code |= 1 << (3 * bit + shift)
    return code


def max_dim(lo, hi) -> int:
    d = [F(hi[k] - lo[k]) for k in range(3)]
    if d[0] > d[1] and d[0] > d[2]:
        return 0
    if d[1] > d[2]:
        return 1
    return 2


def area(lo, hi) -> np.float32:
    d = [F(hi[k] - lo[k]) for k in range(3)]
    return F(F(F(d[0] * d[1]) + F(d[0] * d[2])) + F(d[1] * d[2])) * F(2)


def centroid(lo, hi):
    return [F(F(0.5) * F(lo[k] + hi[k])) for k in range(3)]


def build_canonical(tri: np.ndarray) -> bytes:
    tri = tri.astype(np.float32)
    n = tri.shape[0]
    lo = [[F(min(tri[i, :, k])) for k in range(3)] for i in range(n)]
    hi = [[F(max(tri[i, :, k])) for k in range(3)] for i in range(n)]
    cen = [centroid(lo[i], hi[i]) for i in range(n)]
    clo = [min(c[k] for c in cen) for k in range(3)]
    chi = [max(c[k] for c in cen) for k in range(3)]
    md = max_dim(clo, chi)
    radius = F(F(chi[md] - clo[md]) / F(2))
    cc = centroid(clo, chi)
    alo, ahi = list(clo), list(chi)
    for k in range(3):
        if k != md:
            alo[k] = F(cc[k] - radius)
            ahi[k] = F(cc[k] + radius)
    codes = []
    for i in range(n):
        cell = []
        for k in range(3):
            o = F(cen[i][k] - alo[k])
            if ahi[k] > alo[k]:
                o = F(o / F(ahi[k] - alo[k]))
            cell.append(f2u_sat(F(o * F(1024))))
        codes.append(interleave(*cell))
    order = sorted(range(n), key=lambda i: (codes[i], i))
    treelets = []  # (prim list, lo, hi)
    for i in order:
        key = codes[i] >> 9
        if treelets and treelets[-1][0] == key:
            treelets[-1][1].append(i)
        else:
            treelets.append((key, [i]))

    def box_of(prims):
        return ([min(lo[i][k] for i in prims) for k in range(3)], [max(hi[i][k] for i in prims) for k in range(3)])

    out = bytearray()

    def emit(tag, blo, bhi, count, prims=None):
        out.extend(struct.pack("<I", tag))
        for v in list(blo) + list(bhi):
            out.extend(struct.pack("<f", F(v) + F(0)))
        out.extend(struct.pack("<I", count))
        if prims is not None:
            for p in sorted(prims):
                out.extend(struct.pack("<I", p))

    def bottom(prims):
        blo, bhi = box_of(prims)
        if len(prims) > 1:
            c_lo = [min(cen[i][k] for i in prims) for k in range(3)]
            c_hi = [max(cen[i][k] for i in prims) for k in range(3)]
            ax = max_dim(c_lo, c_hi)
            split = centroid(c_lo, c_hi)[ax]
            left = [i for i in prims if cen[i][ax] < split]
            right = [i for i in prims if not cen[i][ax] < split]
            if left and right:
                emit(ax << 8, blo, bhi, len(prims))
                bottom(left)
                bottom(right)
                return
        emit(1, blo, bhi, len(prims), prims)

    def top(ids):
        if len(ids) == 1:
            bottom(treelets[ids[0]][1])
            return
        tb = {t: box_of(treelets[t][1]) for t in ids}
        tc = {t: centroid(*tb[t]) for t in ids}
        c_lo = [min(tc[t][k] for t in ids) for k in range(3)]
        c_hi = [max(tc[t][k] for t in ids) for k in range(3)]
        ax = max_dim(c_lo, c_hi)
        assert c_lo[ax] != c_hi[ax], "degenerate treelet centroids"
        base, span = c_lo[ax], F(c_hi[ax] - c_lo[ax])
        bucket = {}
        for t in ids:
            b = f2u_sat(F(F(24) * F(F(tc[t][ax] - base) / span)))
            bucket[t] = min(b, 23)
        all_prims = [p for t in ids for p in treelets[t][1]]
        flo, fhi = box_of(all_prims)
        total = area(flo, fhi)
        best, split = None, 0
        for s in range(23):
            L = [t for t in ids if bucket[t] <= s]
            R = [t for t in ids if bucket[t] > s]
            def side(ts):
                if not ts:
                    return F(0), F(np.inf)  # empty side: 0 * inf = NaN, never chosen
                pr = [p for t in ts for p in treelets[t][1]]
                return F(len(pr)), area(*box_of(pr))
            cl, al = side(L)
            cr, ar = side(R)
            with np.errstate(invalid="ignore"):
                cost = F(F(0.125) + F(F(F(cl * al) + F(cr * ar)) / total))
            if best is None or cost < best:
                best, split = cost, s
        left = [t for t in ids if bucket[t] <= split]
        right = [t for t in ids if bucket[t] > split]
        if not left or not right:
            mid = centroid(c_lo, c_hi)[ax]
            left = [t for t in ids if tc[t][ax] < mid]
            right = [t for t in ids if not tc[t][ax] < mid]
        emit(ax << 8, flo, fhi, len(all_prims))
        top(left)
        top(right)

    if n:
        top(list(range(len(treelets))))
    return bytes(out)
