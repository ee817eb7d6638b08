"""BVH4 layout and a CPU emulation of the device traversal.

``emulate_closest`` follows the traversal kernel step by step (node decode,
empty-slot rejection, slab test with the 4-ulp inflated far distance,
near-to-far push order, leaf decode, watertight triangle test in float32)
and asserts that every node, slot and stack access stays in bounds. Its
hits must equal the oracle's brute-force closest hits bit for bit.
"""

import numpy as np
import pytest

import hdcodex_amd as hx
import oracle
from helpers import scene_rays

F = np.float32
LEAF = 0x80000000
EMPTY = 0xFFFFFFFF
MAX_STACK = 64


def _pick(v, k):
    return v[k]


def _prepare(o, d):
    a = np.abs(d)
    kz = 0
    m = a[0]
    if a[1] > m:
        kz, m = 1, a[1]
    if a[2] > m:
        kz = 2
    kx = 0 if kz + 1 == 3 else kz + 1
    ky = 0 if kx + 1 == 3 else kx + 1
    if d[kz] < 0:
        kx, ky = ky, kx
    sx, sy, sz = F(d[kx] / d[kz]), F(d[ky] / d[kz]), F(F(1) / d[kz])
    inv = np.array([F(1) / (x if abs(x) > F(1e-30) else np.copysign(F(1e-30), x)) for x in d], dtype=F)
    return kx, ky, kz, sx, sy, sz, inv


def _intersect(o, pre, p0, p1, p2, tmin, tmax):
    kx, ky, kz, sx, sy, sz, _ = pre
    a, b, c = (p0[:3] - o).astype(F), (p1[:3] - o).astype(F), (p2[:3] - o).astype(F)
    ak, bk, ck = a[kz], b[kz], c[kz]
    ax, ay = F(a[kx] - F(sx * ak)), F(a[ky] - F(sy * ak))
    bx, by = F(b[kx] - F(sx * bk)), F(b[ky] - F(sy * bk))
    cx, cy = F(c[kx] - F(sx * ck)), F(c[ky] - F(sy * ck))
    U, V, W = F(F(cx * by) - F(cy * bx)), F(F(ax * cy) - F(ay * cx)), F(F(bx * ay) - F(by * ax))
    if U == 0 or V == 0 or W == 0:
        U = F(float(cx) * float(by) - float(cy) * float(bx))
        V = F(float(ax) * float(cy) - float(ay) * float(cx))
        W = F(float(bx) * float(ay) - float(by) * float(ax))
    if (U < 0 or V < 0 or W < 0) and (U > 0 or V > 0 or W > 0):
        return None
    det = F(F(U + V) + W)
    if det == 0:
        return None
    T = F(F(F(U * F(sz * ak)) + F(V * F(sz * bk))) + F(W * F(sz * ck)))
    t = F(T / det)
    if not (tmin <= t <= tmax):
        return None
    return t, F(V / det), F(W / det)


def emulate_closest(nodes, slots, o, d, tmin=F(0.001), tmax=F(1e4)):
    o = o.astype(F)
    d = d.astype(F)
    pre = _prepare(o, d)
    inv = pre[6]
    best_t, best_p, best_uv = F(tmax), EMPTY, (F(0), F(0))
    found = False
    stack = []
    cur = 0
    while True:
        if cur & LEAF:
            first, count = (cur >> 3) & 0x0FFFFFFF, (cur & 7) + 1
            assert first + count <= len(slots), "leaf beyond the slot array"
            for j in range(count):
                tri = slots[first + j]
                prim = int(tri[0, 3:4].view(np.uint32)[0])
                h = _intersect(o, pre, tri[0], tri[1], tri[2], tmin, best_t)
                if h is not None:
                    t, u, v = h
                    if not found or t < best_t or (t == best_t and prim < best_p):
                        found, best_t, best_p, best_uv = True, t, prim, (u, v)
        else:
            assert cur < len(nodes), "inner reference out of range"
            n = nodes[cur]
            tn, ref = [], []
            for i in range(4):
                c = int(n["child"][i])
                t0 = (n["lox"][i] - o[0]) * inv[0], (n["loy"][i] - o[1]) * inv[1], (n["loz"][i] - o[2]) * inv[2]
                t1 = (n["hix"][i] - o[0]) * inv[0], (n["hiy"][i] - o[1]) * inv[1], (n["hiz"][i] - o[2]) * inv[2]
                near = max(max(min(t0[0], t1[0]), min(t0[1], t1[1])), max(min(t0[2], t1[2]), tmin))
                far = min(min(max(t0[0], t1[0]), max(t0[1], t1[1])), min(max(t0[2], t1[2]), best_t))
                hit = c != EMPTY and near <= F(far * F(1.00000048))
                tn.append(F(near) if hit else F(np.inf))
                ref.append(c)
            order = sorted(range(4), key=lambda i: (tn[i], i))
            hits = [ref[i] for i in order if tn[i] != np.inf]
            if hits:
                for r in reversed(hits[1:]):
                    stack.append(r)
                assert len(stack) <= MAX_STACK, "traversal stack overflow"
                cur = hits[0]
                continue
        if not stack:
            break
        cur = stack.pop()
    return (best_t, best_p, best_uv) if found else (F(0), EMPTY, (F(0), F(0)))


@pytest.mark.parametrize("name,args", [("cornell", (0,)), ("materials", (0,)), ("heightfield", (2, 48, 16))])
def test_host_bvh_structure(name, args):
    scene = hx.ProceduralScene(name, *args)
    info, nodes, slots = hx.build_bvh_host(scene)
    assert info["slot_count"] == scene.triangle_count
    assert info["max_stack"] <= MAX_STACK
    prims = np.sort(slots[:, 0, 3].view(np.uint32))
    assert np.array_equal(prims, np.arange(scene.triangle_count, dtype=np.uint32))
    for n in nodes:
        for i in range(4):
            if n["child"][i] == EMPTY:
                assert n["lox"][i] > n["hix"][i]


@pytest.mark.parametrize("name,args,count", [("cornell", (0,), 300), ("materials", (0,), 200),
                                              ("heightfield", (2, 48, 16), 200)])
def test_traversal_emulation_matches_oracle(name, args, count):
    scene = hx.ProceduralScene(name, *args)
    _, nodes, slots = hx.build_bvh_host(scene)
    cam = scene.camera(64, 64)
    a = scene.arrays()["positions"][:, :3]
    rays = scene_rays(cam, count, 5, inside=(a.min(axis=0), a.max(axis=0)))
    ref, _ = oracle.OracleScene(scene.view).trace_closest(rays, brute_force=True)
    for i, r in enumerate(rays):
        t, p, (u, v) = emulate_closest(nodes, slots, r[0:3], r[4:7], F(r[3]), F(r[7]))
        assert p == ref["primitive"][i]
        if p != EMPTY:
            assert t.view(np.uint32) == ref["t"][i].view(np.uint32)
            assert u.view(np.uint32) == ref["u"][i].view(np.uint32)
            assert v.view(np.uint32) == ref["v"][i].view(np.uint32)


def test_empty_slot_box_alone_would_pass_the_slab_test():
    """Why the kernel rejects empty slots by reference and not by box: an
    inverted box passes a min/max slab test."""
    inf = F(np.inf)
    o, inv = F(0.3), F(1 / 0.7)
    t0, t1 = (inf - o) * inv, (-inf - o) * inv
    assert max(min(t0, t1), F(0.001)) <= min(max(t0, t1), F(1e4))
