"""Host replay of the library's LBVH walks (test infrastructure).

It restates, step by step and with the same FP32 / FP64 arithmetic, the
device walks whose in-kernel counters bench.py turns into algorithmic bytes:
the KNN nearest-child-first walk with its packed k-list
(rtdl_amd/csrc/traverse.hpp: walk_nearest, collect_packed), the ray
any-hit walk (trace_ray, mode 0) and the fixed-radius count walk with its
distance histogram (rtdl_amd/csrc/engine.hip: nbr_count_query).  A replay returns the nodes fetched and the
primitives tested, which must equal the kernels' counters.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from rtdl_amd import mi355x_runtime as rt_native

F32 = np.float32
END = 0xFFFFFFFF
LEAF = 0x80000000
START_MASK = (1 << 28) - 1
NODE_DTYPE = np.dtype([("lo", "<f4", (3,)), ("link", "<u4"), ("hi", "<f4", (3,)), ("escape", "<u4")])
assert NODE_DTYPE.itemsize == 32


class _Export(ctypes.Structure):
    _fields_ = [("node_count", ctypes.c_uint64), ("prim_count", ctypes.c_uint64),
                ("instance_count", ctypes.c_uint64), ("prim_float4s", ctypes.c_uint32),
                ("exact_doubles", ctypes.c_uint32), ("m", ctypes.c_float), ("pad", ctypes.c_uint32)]


def export_scene(handle) -> dict:
    lib = rt_native._library()
    info = _Export()
    err = ctypes.create_string_buffer(1024)
    if lib.rtdl_mi355x_scene_export(handle, ctypes.byref(info), None, None, None, None, None, err, 1024):
        raise RuntimeError(err.value.decode())
    nodes = np.zeros(info.node_count, dtype=NODE_DTYPE)
    prims = np.zeros((info.prim_count * info.prim_float4s, 4), dtype=np.float32)
    exact = np.zeros(max(1, info.prim_count * info.exact_doubles), dtype=np.float64)
    roots = np.zeros(max(1, info.instance_count), dtype=np.uint32)
    ends = np.zeros(max(1, info.instance_count), dtype=np.uint32)
    if lib.rtdl_mi355x_scene_export(handle, ctypes.byref(info), nodes.ctypes.data, prims.ctypes.data,
                                    exact.ctypes.data if info.exact_doubles else None, roots.ctypes.data,
                                    ends.ctypes.data, err, 1024):
        raise RuntimeError(err.value.decode())
    out = {"nodes": nodes, "prims": prims, "exact": exact if info.exact_doubles else None,
           "roots": roots, "ends": ends, "m": F32(info.m), "prim_float4s": info.prim_float4s,
           "qnodes": None, "qframe": None}
    has = ctypes.c_uint32(0)
    if lib.rtdl_mi355x_scene_export_compact(handle, ctypes.byref(has), None, None, err, 1024):
        raise RuntimeError(err.value.decode())
    if has.value:
        qnodes = np.zeros((info.node_count, 4), dtype=np.uint32)
        frame = np.zeros(8, dtype=np.float32)
        if lib.rtdl_mi355x_scene_export_compact(handle, ctypes.byref(has), qnodes.ctypes.data, frame.ctypes.data,
                                                err, 1024):
            raise RuntimeError(err.value.decode())
        out["qnodes"], out["qframe"] = qnodes, frame
    return out


def compact_boxes(sc: dict):
    """Dequantised boxes (lo, hi: float32 [n, 3]) and words of a scene's
    compact nodes: box = origin + q * step per axis in FP32, as the count
    walk computes it (exact for the frame's steps)."""
    qn, fr = sc["qnodes"], sc["qframe"]
    q = np.stack([qn[:, 0] & 0xFFFF, qn[:, 0] >> 16, qn[:, 1] & 0xFFFF,
                  qn[:, 1] >> 16, qn[:, 2] & 0xFFFF, qn[:, 2] >> 16], 1).astype(F32)
    o, st = fr[:3].astype(F32), fr[3:6].astype(F32)
    lo = (o + q[:, :3] * st).astype(F32)
    hi = (o + q[:, 3:] * st).astype(F32)
    return lo, hi, qn[:, 3].copy()


# ------------------------------------------------------------ arithmetic --

def f32_ru(x: float) -> np.float32:
    """FP32 rounding of x towards +infinity."""
    f = F32(x)
    if float(f) < x:
        f = np.nextafter(f, F32(np.inf))
    return f


def fma32(a: np.float32, b: np.float32, c: np.float32) -> np.float32:
    """a * b + c rounded once to FP32.  The product of two FP32 values is
    exact in FP64; the FP64 sum s carries an exact error term e (TwoSum), and
    s + e rounds to FP32 like s unless s is exactly halfway between two FP32
    values, where the sign of e decides."""
    p = float(a) * float(b)
    cc = float(c)
    s = p + cc
    bb = s - p
    e = (p - (s - bb)) + (cc - bb)
    f = F32(s)
    if float(f) != s and e != 0.0:
        lo = f if float(f) < s else np.nextafter(f, F32(-np.inf))
        hi = f if float(f) > s else np.nextafter(f, F32(np.inf))
        if s == (float(lo) + float(hi)) / 2:
            f = hi if e > 0 else lo
    return f


def box_dist2(q, lo, hi) -> np.float32:
    d2 = F32(0.0)
    for a in range(3):
        g = max(max(F32(lo[a] - q[a]), F32(q[a] - hi[a])), F32(0.0))
        d2 = fma32(g, g, d2)
    return d2


def frn_threshold2(rho: float, m: np.float32) -> np.float32:
    if not rho < 3.0e38:
        return F32(np.inf)
    slack = float(m) * 4.76837158203125e-07
    t = (rho + slack) * 1.000003814697265625
    return f32_ru(t * t)


def coord_bound(m: np.float32, x: float, y: float, z: float) -> np.float32:
    out = F32(m)
    for v in (x, y, z):
        out = max(out, f32_ru(abs(v)))
    return out


def leaf_range(link: int) -> range:
    first = link & START_MASK
    return range(first, first + ((link >> 28) & 7) + 1)


# -------------------------------------------------------------------- KNN --

def knn_counts(sc: dict, q, k: int, *, instance: int = 0, stack_depth: int = 32) -> tuple[int, int]:
    """Nodes fetched and primitives tested by the batched KNN walk of query q
    (FP64 coordinates of FP32 values) for a k-nearest search."""
    nodes, prims, exact = sc["nodes"], sc["prims"], sc["exact"]
    root, end = int(sc["roots"][instance]), int(sc["ends"][instance])
    kk = 1
    while kk < k:
        kk <<= 1
    qx, qy, qz = (float(v) for v in q)
    qf = (F32(qx), F32(qy), F32(qz))
    m = coord_bound(sc["m"], qx, qy, qz)
    full_key = (1 << 64) - 1
    lst = [0] * (kk - k) + [full_key] * k
    state = {"thr2": F32(np.inf), "nodes": 0, "prims": 0}

    def visit(slot: int) -> None:
        if exact is not None:
            px, py, pz = exact[3 * slot], exact[3 * slot + 1], exact[3 * slot + 2]
        else:
            px, py, pz = (float(v) for v in prims[slot, :3])
        dx, dy, dz = px - qx, py - qy, pz - qz
        d2 = dx * dx + dy * dy + dz * dz
        key = (int(F32(d2).view(np.uint32)) << 32) | slot
        if key < lst[-1]:
            lst[-1] = key
            lst.sort()
            if lst[-1] != full_key:
                kth = float(np.uint32(lst[-1] >> 32).view(np.float32))
                state["thr2"] = frn_threshold2(math.sqrt(kth) * 1.0000002384185791015625, m)

    def leaf_tests(link: int, skip_in_list: bool) -> None:
        for slot in leaf_range(link):
            state["prims"] += 1
            p = prims[slot]
            fx, fy, fz = F32(p[0] - qf[0]), F32(p[1] - qf[1]), F32(p[2] - qf[2])
            d2f = fma32(fz, fz, fma32(fy, fy, F32(fx * fx)))
            if d2f <= state["thr2"]:
                if skip_in_list:
                    if exact is not None:
                        px, py, pz = exact[3 * slot], exact[3 * slot + 1], exact[3 * slot + 2]
                    else:
                        px, py, pz = (float(v) for v in p[:3])
                    dx, dy, dz = px - qx, py - qy, pz - qz
                    key = (int(F32(dx * dx + dy * dy + dz * dz).view(np.uint32)) << 32) | slot
                    if key in lst:
                        continue
                visit(slot)

    def box(n: int) -> np.float32:
        return box_dist2(qf, nodes[n]["lo"], nodes[n]["hi"])

    if root == END or root == end:
        return 0, 0
    complete = True
    stack: list[int] = []
    cur = root
    state["nodes"] += 1
    if box(cur) <= state["thr2"]:
        while True:
            leaf = True
            while not (int(nodes[cur]["link"]) & LEAF):
                # the left child is the next record, the right one the link
                left, right = cur + 1, int(nodes[cur]["link"])
                state["nodes"] += 2
                dl, dr = box(left), box(right)
                right_first = dr < dl
                dn, df = (dr, dl) if right_first else (dl, dr)
                if df <= state["thr2"]:
                    if len(stack) < stack_depth:
                        stack.append(left if right_first else right)
                    else:
                        complete = False
                if not (dn <= state["thr2"]):
                    leaf = False
                    break
                cur = right if right_first else left
            if leaf:
                leaf_tests(int(nodes[cur]["link"]), False)
            found = False
            while stack:
                cur = stack.pop()
                state["nodes"] += 1
                if box(cur) <= state["thr2"]:
                    found = True
                    break
            if not found:
                break
    if not complete:  # the device walks the tree again in rope preorder
        node = root
        while node != end:
            state["nodes"] += 1
            if box(node) <= state["thr2"]:
                link = int(nodes[node]["link"])
                if not link & LEAF:
                    node += 1
                    continue
                leaf_tests(link, True)
            node = int(nodes[node]["escape"])
    return state["nodes"], state["prims"]


# ----------------------------------------------------------- fixed radius --

NBR_BUCKETS = 32
NBR_CAND_MAX = 128
NBR_COUNT_MAX = 0xFFFF


def _nbr_bucket(dist: float, scale: float) -> int:
    x = dist * scale
    if x < NBR_BUCKETS:
        return int(x) if x >= 0.0 else 0
    return NBR_BUCKETS - 1


def _nbr_bucket_edge(b: int, scale: float) -> float:
    return float(b + 1) / scale * 1.0000000000009094947


def frn_counts(sc: dict, queries, radius: float, k: int, *, instance: int = 0) -> tuple[int, int, int]:
    """Nodes tested, primitives tested and the wavefront's walk iterations of
    the batched fixed-radius count walk of one work item: up to 64 queries
    (FP64 coordinates of FP32 values), one per lane of a wavefront, over all
    count rounds.

    The wavefront runs in lock step.  Each iteration every walking lane takes
    one node step of its rope preorder walk, and a second one when the first
    descends (to the child stored right after its parent, fetched with it),
    with the same threshold; the leaf it reaches joins a leaf queue (lane
    order).  While 64 or more leaves are queued (and, once no lane walks,
    while any are), a batch of up to 64 is taken in queue order and the
    points of its leaves are prefiltered in FP32 against their owners'
    current thresholds; the survivors join a point queue (point index within
    the leaf first, then position in the batch), and after each point index,
    while 64 or more points are queued (and, once no lane walks and the leaf
    queue is empty, while any are), a batch of up to 64 is taken in queue
    order: each point inside the owner's FP64 radius whose distance bucket
    does not lie beyond the owner's last counted bucket is counted; then
    every lane whose count reached its checkpoint (every max(k/8, 4)) shrinks
    its threshold to the edge of the bucket holding its k-th neighbour.
    Rounds are wavefront-synchronous: a lane whose buckets up to the k-th
    still hold more than 128 neighbours counts again over the finer buckets
    of that ball (at most four rounds)."""
    nodes, prims, exact = sc["nodes"], sc["prims"], sc["exact"]
    root, end = int(sc["roots"][instance]), int(sc["ends"][instance])
    # the walk tests the compact nodes' boxes when the scene's frame is usable
    if sc.get("qnodes") is not None and sc["qframe"][6] == 0:
        box_lo, box_hi, _ = compact_boxes(sc)
    else:
        box_lo, box_hi = nodes["lo"], nodes["hi"]
    lanes = len(queries)
    q = [tuple(float(v) for v in row) for row in queries]
    qf = [(F32(x), F32(y), F32(z)) for x, y, z in q]
    m = [coord_bound(sc["m"], *row) for row in q]
    rad = [radius] * lanes
    rad_sq = [radius * radius] * lanes
    step = k // 8 if k // 8 > 4 else 4
    going = [True] * lanes
    n_nodes = n_prims = n_steps = 0

    def d2_of(lane: int, slot: int) -> float:
        if exact is not None:
            px, py, pz = exact[3 * slot], exact[3 * slot + 1], exact[3 * slot + 2]
        else:
            px, py, pz = (float(v) for v in prims[slot, :3])
        dx, dy, dz = px - q[lane][0], py - q[lane][1], pz - q[lane][2]
        return dx * dx + dy * dy + dz * dz

    for rnd in range(4):
        if not any(going):
            break
        scale = [float(NBR_BUCKETS) / rad[i] for i in range(lanes)]
        thr2 = [frn_threshold2(rad[i], m[i]) for i in range(lanes)]
        top = [NBR_BUCKETS - 1] * lanes
        nxt = [k] * lanes
        shrunk = [False] * lanes
        hist = [[0] * NBR_BUCKETS for _ in range(lanes)]
        count = [0] * lanes
        node = [root] * lanes
        walking = [going[i] and root != end for i in range(lanes)]
        queue: list[tuple[int, int]] = []
        leaves: list[tuple[int, int, int]] = []  # (owner, first, count)
        while True:
            n_steps += 1
            leaf = [(0, 0)] * lanes
            for i in range(lanes):
                if not walking[i]:
                    continue
                for half in range(2):
                    cur = node[i]
                    n_nodes += 1
                    nd = nodes[cur]
                    link = int(nd["link"])
                    inside = box_dist2(qf[i], box_lo[cur], box_hi[cur]) <= thr2[i]
                    if inside and link & LEAF:
                        leaf[i] = (link & START_MASK, ((link >> 28) & 7) + 1)
                    descend = inside and not link & LEAF
                    # into the left child (the next record), else past the node
                    node[i] = cur + 1 if descend else int(nd["escape"])
                    walking[i] = node[i] != end
                    if not (walking[i] and descend):
                        break
            def drain() -> None:
                nonlocal queue
                batch, queue = queue[:64], queue[64:]
                tops = list(top)
                for o, slot in batch:
                    d2 = d2_of(o, slot)
                    if d2 > rad_sq[o]:
                        continue
                    b = _nbr_bucket(math.sqrt(d2), scale[o])
                    if b <= tops[o]:
                        hist[o][b] += 1
                        count[o] += 1
                for i in range(lanes):
                    if not going[i] or count[i] < nxt[i]:
                        continue
                    nxt[i] = count[i] + step
                    cum = 0
                    for c in range(top[i]):
                        cum += hist[i][c]
                        if cum >= k:
                            top[i] = c
                            shrunk[i] = True
                            thr2[i] = frn_threshold2(min(rad[i], _nbr_bucket_edge(c, scale[i])), m[i])
                            break

            def leaf_drain() -> None:
                nonlocal leaves, n_prims
                batch, leaves = leaves[:64], leaves[64:]
                limits = [thr2[o] for o, _, _ in batch]  # the owners' thresholds when taken
                for e in range(8):
                    if not any(e < cnt for _, _, cnt in batch):
                        break
                    for (o, first, cnt), lim in zip(batch, limits):
                        if e >= cnt:
                            continue
                        n_prims += 1
                        slot = first + e
                        p = prims[slot]
                        fx, fy, fz = F32(p[0] - qf[o][0]), F32(p[1] - qf[o][1]), F32(p[2] - qf[o][2])
                        if fma32(fz, fz, fma32(fy, fy, F32(fx * fx))) <= lim:
                            queue.append((o, slot))
                    while len(queue) >= 64:
                        drain()

            for i in range(lanes):
                if leaf[i][1]:
                    leaves.append((i, leaf[i][0], leaf[i][1]))
            while len(leaves) >= 64:
                leaf_drain()
            if not any(walking):
                while leaves:
                    leaf_drain()
                while queue:
                    drain()
                break
        for i in range(lanes):
            if not going[i]:
                continue
            going[i] = False
            heavy = shrunk[i] or count[i] > k
            if not heavy or count[i] > NBR_COUNT_MAX:
                continue
            bucket, n_cand, cum = top[i], 0, 0
            for c in range(top[i] + 1):
                cum += hist[i][c]
                if cum >= k:
                    bucket, n_cand = c, cum
                    break
            if n_cand > NBR_CAND_MAX and rnd < 3:
                rad[i] = min(rad[i], _nbr_bucket_edge(bucket, scale[i]))
                rad_sq[i] = min(rad_sq[i], rad[i] * rad[i])
                going[i] = True
    return n_nodes, n_prims, n_steps


# ------------------------------------------------------------------- rays --

_C = F32(9.5367431640625e-07)


def _segment_box(o, d, inv, lo, hi, pad, slim) -> bool:
    tn = F32(-9.5367431640625e-07)
    tf = F32(F32(slim * F32(1.00000095367431640625)) + _C)
    for a in range(3):
        lo_a, hi_a = F32(lo[a] - pad), F32(hi[a] + pad)
        if d[a] == 0.0:
            if o[a] < lo_a or o[a] > hi_a:
                return False
            continue
        t0 = F32(F32(lo_a - o[a]) * inv[a])
        t1 = F32(F32(hi_a - o[a]) * inv[a])
        a0, a1 = min(t0, t1), max(t0, t1)
        tn = max(tn, F32(a0 - F32(abs(a0) * _C)))
        tf = min(tf, F32(a1 + F32(abs(a1) * _C)))
    return tn <= tf


def _mt_hit(ox, oy, oz, dx, dy, dz, tmax, v) -> bool:
    e1x, e1y, e1z = v[3] - v[0], v[4] - v[1], v[5] - v[2]
    e2x, e2y, e2z = v[6] - v[0], v[7] - v[1], v[8] - v[2]
    px, py, pz = dy * e2z - dz * e2y, dz * e2x - dx * e2z, dx * e2y - dy * e2x
    det = e1x * px + e1y * py + e1z * pz
    if abs(det) <= 1.0e-8:
        return False
    inv = 1.0 / det
    tx, ty, tz = ox - v[0], oy - v[1], oz - v[2]
    u = (tx * px + ty * py + tz * pz) * inv
    if u < 0.0 or u > 1.0:
        return False
    qx, qy, qz = ty * e1z - tz * e1y, tz * e1x - tx * e1z, tx * e1y - ty * e1x
    vv = (dx * qx + dy * qy + dz * qz) * inv
    if vv < 0.0 or (u + vv) > 1.0:
        return False
    t = (e2x * qx + e2y * qy + e2z * qz) * inv
    return 0.0 <= t <= tmax


def ray_anyhit_counts(sc: dict, ray, *, instance: int = 0) -> tuple[int, int]:
    """Nodes fetched and primitives tested by the any-hit walk of one ray
    (ox, oy, oz, dx, dy, dz, tmax as FP64 values of FP32 inputs)."""
    nodes, prims, exact = sc["nodes"], sc["prims"], sc["exact"]
    root, end = int(sc["roots"][instance]), int(sc["ends"][instance])
    ox, oy, oz, dx, dy, dz, tmax = (float(v) for v in ray)
    if not tmax >= 0.0 or root == END:
        return 0, 0
    o32 = (F32(ox), F32(oy), F32(oz))
    d32, inv, pad = [F32(0)] * 3, [F32(0)] * 3, F32(np.inf)
    seg_t = tmax
    if not tmax < math.inf:
        # an unbounded ray walks 4 (M + |o|) / |d| of its length (max norms)
        dn = max(max(abs(dx), abs(dy)), abs(dz))
        on = max(max(abs(ox), abs(oy)), abs(oz))
        seg_t = 4.0 * ((float(sc["m"]) + on) / dn) if dn > 0 else math.inf
    if all(abs(v * seg_t) < math.inf for v in (dx, dy, dz)):
        dd = (dx * seg_t, dy * seg_t, dz * seg_t)
        omax = max(max(abs(ox), abs(oy)), abs(oz))
        dmax = max(max(abs(dd[0]), abs(dd[1])), abs(dd[2]))
        pad = f32_ru((float(sc["m"]) + omax + dmax) * 3.814697265625e-06)
        d32 = [F32(v) for v in dd]
        inv = [F32(F32(1.0) / v) if v != 0 else F32(0) for v in d32]
    slim = F32(1.0)
    n_nodes = n_prims = 0
    node, done = root, False
    while node != end and not done:
        nd = nodes[node]
        n_nodes += 1
        if not _segment_box(o32, d32, inv, nd["lo"], nd["hi"], pad, slim):
            node = int(nd["escape"])
            continue
        link = int(nd["link"])
        if not link & LEAF:
            node += 1  # the left child is the next record
            continue
        for slot in leaf_range(link):
            n_prims += 1
            if exact is not None:
                v = [float(x) for x in exact[9 * slot: 9 * slot + 9]]
            else:
                a, b, c = prims[3 * slot], prims[3 * slot + 1], prims[3 * slot + 2]
                v = [float(a[0]), float(a[1]), float(a[2]), float(b[0]), float(b[1]), float(b[2]),
                     float(c[0]), float(c[1]), float(c[2])]
            if _mt_hit(ox, oy, oz, dx, dy, dz, tmax, v):
                done = True
                break
        node = int(nd["escape"])
    return n_nodes, n_prims
