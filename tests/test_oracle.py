"""Pins the CPU oracle (the checker of the HIP step) without the reference.

The reference has no tests or golden vectors, so each oracle function is
checked against something independent of its own code: exhaustive or naive
restatements of the bit arithmetic, numpy brute force, closed-form physics,
and the reference's runtime invariants (colouring validators, permutation
checksum).  Runs on the CPU.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import beatbox
import scenes
from beatbox import (LBVH_NODE, MANIFOLD_TAP, TAP_EDGE_COLORS, TAP_ISLANDS, TAP_LBVH,
                     TAP_MANIFOLDS, TAP_MORTON, TAP_PAIRS, TAP_SORTED, TAP_BODY_COLORS, WorldDesc)
from oracle_bind import OracleWorld, collide, olib


# ------------------------------------------------------------------ Morton
def naive_spread(v: int, bits: int, stride: int) -> int:
    out = 0
    for b in range(bits):
        if (v >> b) & 1:
            out |= 1 << (b * stride)
    return out


def test_expand10_exhaustive():
    L = olib()
    for v in range(1024):
        assert L.oracle_expand10(v) == naive_spread(v, 10, 3)


def test_expand21_sampled():
    L = olib()
    rng = np.random.default_rng(1)
    vals = list(range(0, 64)) + [2 ** 21 - 1, 2 ** 20, 12345] + rng.integers(0, 2 ** 21, 2000).tolist()
    for v in vals:
        assert L.oracle_expand21(int(v)) == naive_spread(int(v), 21, 3)


def test_morton30_closed_form_grid():
    L = olib()
    rng = np.random.default_rng(2)
    for _ in range(3000):
        ix, iy, iz = (int(x) for x in rng.integers(0, 1024, 3))
        p = [(i + 0.5) / 1024.0 for i in (ix, iy, iz)]
        want = (naive_spread(ix, 10, 3) << 2) | (naive_spread(iy, 10, 3) << 1) | naive_spread(iz, 10, 3)
        assert L.oracle_morton30(*p) == want
    # clamping below 0 and at/above 1 (RB_sim.slang normalisation + LBVH.slang:21-26)
    assert L.oracle_morton30(-0.5, -3.0, -1e9) == 0
    top = (naive_spread(1023, 10, 3) << 2) | (naive_spread(1023, 10, 3) << 1) | naive_spread(1023, 10, 3)
    assert L.oracle_morton30(1.0, 7.0, 1e9) == top


def test_morton63_closed_form():
    L = olib()
    rng = np.random.default_rng(3)
    for _ in range(2000):
        ix, iy, iz = (int(x) for x in rng.integers(0, 2 ** 21, 3))
        p = [np.float32((i + 0.5) / 2 ** 21) for i in (ix, iy, iz)]
        q = [min(max(int(np.float32(x) * np.float32(2 ** 21)), 0), 2 ** 21 - 1) for x in p]
        want = (naive_spread(q[0], 21, 3) << 2) | (naive_spread(q[1], 21, 3) << 1) | naive_spread(q[2], 21, 3)
        assert L.oracle_morton63(*[float(x) for x in p]) == want


def test_order_map_monotone_roundtrip():
    L = olib()
    vals = np.array([-np.inf, -3e38, -1.0, -1e-30, -1e-45, -0.0, 0.0, 1e-45, 1e-30, 1.0, 3e38, np.inf],
                    np.float32)
    mapped = [L.oracle_order_map(float(v)) for v in vals]
    assert mapped == sorted(mapped) and len(set(mapped)) == len(mapped)
    rng = np.random.default_rng(4)
    for v in rng.standard_normal(2000).astype(np.float32) * 1e3:
        back = L.oracle_order_unmap(L.oracle_order_map(float(v)))
        assert np.float32(back).tobytes() == np.float32(v).tobytes()


def test_sincos_against_libm():
    L = olib()
    s, c = C.c_float(), C.c_float()
    xs = np.concatenate([np.linspace(-10, 10, 4001), np.linspace(-1e-3, 1e-3, 201)]).astype(np.float32)
    for x in xs:
        L.oracle_sincos(float(x), C.byref(s), C.byref(c))
        assert abs(s.value - math.sin(float(x))) <= 4e-7 * max(1.0, abs(math.sin(float(x)))) + 1e-7
        assert abs(c.value - math.cos(float(x))) <= 4e-7 * max(1.0, abs(math.cos(float(x)))) + 1e-7


def test_solve6_against_dense_solve():
    L = olib()
    rng = np.random.default_rng(5)
    for _ in range(200):
        A = rng.standard_normal((6, 6))
        H = (A @ A.T + 6 * np.eye(6)).astype(np.float32)
        b = rng.standard_normal(6).astype(np.float32)
        x = np.zeros(6, np.float32)
        L.oracle_solve6(H.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                        x.ctypes.data_as(C.c_void_p))
        ref = np.linalg.solve(H.astype(np.float64), b.astype(np.float64))
        assert np.allclose(x, ref, rtol=1e-4, atol=1e-5)


# ------------------------------------------------------ step-level checks
def _world(bodies, **kw):
    desc = WorldDesc.make(**scenes.world_desc(**kw))
    w = OracleWorld(desc)
    w.add_bodies(bodies)
    return w


def _quat_rotate(q, v):
    x, y, z, w = q
    qv = np.array([x, y, z], np.float64)
    t = 2.0 * np.cross(qv, v)
    return v + w * t + np.cross(qv, t)


def _aabb_np(body):
    pts = []
    mn, mx = body["minimum"], body["maximum"]
    for i in range(8):
        c = np.array([mx[0] if i & 1 else mn[0], mx[1] if i & 2 else mn[1], mx[2] if i & 4 else mn[2]],
                     np.float64)
        pts.append(_quat_rotate(body["rotation"].astype(np.float64), c) + body["position"])
    pts = np.array(pts)
    return pts.min(0), pts.max(0)


@pytest.mark.parametrize("bits", [30, 63])
def test_sort_is_stable_sort_of_keys(bits):
    bodies = scenes.random_boxes(0xC001, 300, extent=20.0, y0=0.6, y1=30.0)
    w = _world(bodies, morton_bits=bits, quirks=0 if bits == 63 else scenes.QUIRKS_PARITY)
    for _ in range(3):
        w.step(1)
        keys = w.tap(TAP_MORTON, np.uint32 if bits == 30 else np.uint64).astype(np.uint64)
        srt = w.tap(TAP_SORTED, np.uint32).reshape(-1, 4)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(srt[:, 2], order)
        got = srt[:, 0].astype(np.uint64) | (srt[:, 1].astype(np.uint64) << np.uint64(32))
        assert np.array_equal(got, keys[order])


def _subtree_leaves(nodes, n, k):
    if nodes["left"][k] < 0:
        return [k - (n - 1)]
    return _subtree_leaves(nodes, n, nodes["left"][k]) + _subtree_leaves(nodes, n, nodes["right"][k])


def test_lbvh_topology_and_bounds_brute_force():
    for seed, n in [(0xC002, 2), (0xC003, 3), (0xC004, 17), (0xC005, 64), (0xC006, 200)]:
        bodies = scenes.random_boxes(seed, n - 1, extent=3.0, y0=0.6, y1=6.0)
        w = _world(bodies)
        w.step(1)
        nodes = w.tap(TAP_LBVH, LBVH_NODE)
        srt = w.tap(TAP_SORTED, np.uint32).reshape(-1, 4)
        codes = srt[:, 0].astype(np.int64)
        assert len(nodes) == 2 * n - 1
        parents = np.zeros(2 * n - 1, int)
        for k in range(n - 1):
            parents[nodes["left"][k]] += 1
            parents[nodes["right"][k]] += 1
            leaves = sorted(_subtree_leaves(nodes, n, k))
            # contiguous sorted range
            assert leaves == list(range(leaves[0], leaves[-1] + 1))
            # bounds = union of leaf boxes (exact)
            lo = np.min([nodes["minimum"][n - 1 + i] for i in leaves], axis=0)
            hi = np.max([nodes["maximum"][n - 1 + i] for i in leaves], axis=0)
            assert np.array_equal(lo, nodes["minimum"][k]) and np.array_equal(hi, nodes["maximum"][k])
            # split between the two children at the highest differing bit
            # (duplicates split by index: Karras 2012)
            if nodes["left"][k] >= n - 1:
                sl = [nodes["left"][k] - (n - 1)]
            else:
                sl = _subtree_leaves(nodes, n, nodes["left"][k])
            split = max(sl)
            f, l = leaves[0], leaves[-1]
            if codes[f] != codes[l]:
                pref = 31 - int(codes[f] ^ codes[l]).bit_length() + 1
                for s in range(f, l):
                    common = 31 - int(codes[f] ^ codes[s + 1]).bit_length() + 1
                    if common <= pref:  # first position where the prefix breaks
                        assert s == split
                        break
        assert parents[0] == 0 and np.all(parents[1:] == 1)
        # leaf primitive = its row; leaf AABB of each row matches numpy
        rows = w.tap(beatbox.TAP_ROWS, beatbox.BODY_STATE)
        allb = np.concatenate([bodies])
        for i in range(n):
            leaf = nodes[n - 1 + i]
            assert leaf["primitive_idx"] == i
            b = allb[rows["id"][i]].copy()
            b["position"] = rows["position"][i]
            b["rotation"] = rows["rotation"][i]
        for i in range(n):
            assert nodes["left"][n - 1 + i] == -1


def test_pairs_equal_brute_force_overlaps():
    for seed, n in [(0xC010, 50), (0xC011, 400)]:
        bodies = scenes.random_boxes(seed, n, extent=4.0, y0=0.55, y1=6.0)
        w = _world(bodies, iterations=4)
        for _ in range(4):
            w.step(1)
            N = n + 1
            nodes = w.tap(TAP_LBVH, LBVH_NODE)
            leaves = nodes[N - 1:]
            mn, mx = leaves["minimum"], leaves["maximum"]
            ov = np.all((mn[:, None, :] <= mx[None, :, :]) & (mx[:, None, :] >= mn[None, :, :]), axis=2)
            iu = np.triu_indices(N, 1)
            want = np.stack([iu[0][ov[iu]], iu[1][ov[iu]]], 1)
            got = w.tap(TAP_PAIRS, np.uint32).reshape(-1, 2)
            assert np.array_equal(got, want)


def test_leaf_aabb_matches_numpy_rotation():
    bodies = scenes.random_boxes(0xC012, 30, extent=4.0, y0=0.6, y1=5.0)
    w = _world(bodies)
    w.step(1)
    rows = w.tap(beatbox.TAP_ROWS, beatbox.BODY_STATE)
    nodes = w.tap(TAP_LBVH, LBVH_NODE)
    N = len(bodies)
    before = bodies  # step 1 builds the tree from the initial poses
    srt = w.tap(TAP_SORTED, np.uint32).reshape(-1, 4)
    for i in range(N):
        src = before[srt[i, 2]]
        lo, hi = _aabb_np(src)
        assert np.allclose(nodes["minimum"][N - 1 + i], lo, atol=2e-5)
        assert np.allclose(nodes["maximum"][N - 1 + i], hi, atol=2e-5)


# --------------------------------------------------------------- contacts
def _box(pos, quat=(0, 0, 0, 1), half=0.5, mass=1.0, dynamic=True):
    b = np.zeros(1, scenes.BODY_DESC)
    b["shape"] = scenes.SHAPE_BOX
    b["flags"] = (scenes.BODY_DYNAMIC | scenes.BODY_GRAVITY) if dynamic else 0
    b["position"] = pos
    b["rotation"] = quat
    b["minimum"] = -half
    b["maximum"] = half
    b["mass"] = mass if dynamic else 0.0
    b["friction"] = 0.5
    return b


def _sphere(pos, r, mass=1.0):
    b = np.zeros(1, scenes.BODY_DESC)
    b["shape"] = scenes.SHAPE_SPHERE
    b["flags"] = scenes.BODY_DYNAMIC | scenes.BODY_GRAVITY
    b["position"] = pos
    b["rotation"] = (0, 0, 0, 1)
    b["radius"] = r
    b["mass"] = mass
    b["friction"] = 0.5
    return b


DESC = WorldDesc.make(**scenes.world_desc())


def test_sat_resting_cube_four_corners():
    g = np.zeros(1, scenes.BODY_DESC)
    scenes._ground(g[0])
    for d, (x, z) in [(0.01, (0.3, 0.2)), (0.002, (-4.0, 7.5)), (0.05, (1.0, -2.0))]:
        m = collide(DESC, g, _box((x, 0.5 - d, z)))
        assert m is not None and m["key"] == 1 and m["contact_count"] == 4
        assert np.allclose(m["normal"], (0, 1, 0))
        pts = m["contacts"]["position"][:4]
        want = {(x + sx * 0.5, z + sz * 0.5) for sx in (-1, 1) for sz in (-1, 1)}
        got = {(round(float(p[0]), 4), round(float(p[2]), 4)) for p in pts}
        assert got == {(round(a, 4), round(b, 4)) for a, b in want}
        assert np.allclose(m["contacts"]["penetration"][:4], -d, atol=2e-6)
        # tangent basis: unit, orthogonal to the normal (compute_basis)
        t = m["tangent"]
        assert np.allclose(np.linalg.norm(t, axis=1), 1, atol=1e-6)
        assert abs(t[0] @ m["normal"]) < 1e-6 and abs(t[1] @ m["normal"]) < 1e-6


def test_sat_deep_overlap_either_face():
    # at equal separations the 0.95 / 0.01 axis bias (collision_detection.slang:
    # 720-743) may pick the cube's bottom face as reference: then the ground's
    # top face is clipped to the cube footprint (contacts at y = 0)
    g = np.zeros(1, scenes.BODY_DESC)
    scenes._ground(g[0])
    m = collide(DESC, g, _box((0.0, 0.3, 0.0)))
    assert m["contact_count"] == 4 and np.allclose(m["normal"], (0, 1, 0))
    pts = m["contacts"]["position"][:4]
    assert np.allclose(m["contacts"]["penetration"][:4], -0.2, atol=1e-6)
    if m["key"] == 4:
        assert np.allclose(pts[:, 1], 0.0, atol=1e-6)
    else:
        assert m["key"] == 1 and np.allclose(pts[:, 1], -0.2, atol=1e-6)
    assert np.allclose(np.sort(np.abs(pts[:, [0, 2]]).ravel()), 0.5, atol=1e-6)


def test_sat_separated_and_penetration_cap():
    g = np.zeros(1, scenes.BODY_DESC)
    scenes._ground(g[0])
    assert collide(DESC, g, _box((0.0, 0.51, 0.0))) is None
    m = collide(DESC, g, _box((0.0, 0.1, 0.0)))  # 0.4 deep: capped at -0.25
    assert np.allclose(m["contacts"]["penetration"][:m["contact_count"]], -0.25)


def test_sat_edge_edge_closest_points():
    s, c = math.sin(math.pi / 8), math.cos(math.pi / 8)
    a = _box((0, 0, 0), (s, 0, 0, c))          # edge along x at z = +sqrt(2)/2
    b = _box((0, 0, 1.35), (0, s, 0, c))       # edge along y at z = 1.35 - sqrt(2)/2
    m = collide(DESC, a, b)
    assert m is not None and m["key"] >= 6 and m["contact_count"] == 1
    p = m["contacts"]["position"][0]
    # the two crossed edges meet above the origin: closest points (0,0,zA), (0,0,zB)
    za, zb = math.sqrt(0.5), 1.35 - math.sqrt(0.5)
    assert abs(p[0]) < 1e-5 and abs(p[1]) < 1e-5 and abs(p[2] - 0.5 * (za + zb)) < 1e-5


def test_sphere_sphere_closed_form():
    m = collide(DESC, _sphere((0, 1, 0), 0.5), _sphere((0.6, 1.8, 0), 0.5))
    d = math.hypot(0.6, 0.8)
    assert m is not None and m["contact_count"] == 1
    assert np.allclose(m["normal"], (0.6 / d, 0.8 / d, 0), atol=1e-6)
    assert abs(m["contacts"]["penetration"][0] - (d - 1.0)) < 1e-6
    assert collide(DESC, _sphere((0, 1, 0), 0.5), _sphere((0, 2.01, 0), 0.5)) is None


def test_box_sphere_closed_form():
    box = _box((0, 0.5, 0))
    m = collide(DESC, box, _sphere((0.2, 1.3, -0.1), 0.35))  # above the top face
    assert m is not None and np.allclose(m["normal"], (0, 1, 0), atol=1e-6)
    assert abs(m["contacts"]["penetration"][0] - (1.3 - 1.0 - 0.35)) < 1e-6
    # corner region: closest point is the corner (0.5, 1, 0.5)
    cpos = np.array([0.8, 1.3, 0.8])
    m = collide(DESC, box, _sphere(tuple(cpos), 0.6))
    dvec = cpos - np.array([0.5, 1.0, 0.5])
    assert abs(m["contacts"]["penetration"][0] - (np.linalg.norm(dvec) - 0.6)) < 1e-6
    assert np.allclose(m["normal"], dvec / np.linalg.norm(dvec), atol=1e-6)


# ------------------------------------------------------------ integration
def test_pgs_free_fall_closed_form():
    body = _box((0.0, 50.0, 0.0))
    w = _world(body, iterations=8)
    dt, g = np.float32(0.01667), -9.81
    damp = math.exp(-float(dt) * 0.1)
    v, y = 0.0, 50.0
    for _ in range(40):
        w.step(1)
        v = (v + g * float(dt)) * damp
        y = y + v * float(dt)
    r = w.read_bodies()
    assert abs(r["velocity"][0][1] - v) < 2e-5 and abs(r["position"][0][1] - y) < 2e-4


def test_pgs_constant_spin_rotation_angle():
    body = _box((0.0, 50.0, 0.0))
    body["omega"] = (0.0, 3.0, 0.0)
    body["flags"] = scenes.BODY_DYNAMIC  # no gravity
    w = _world(body)
    w.step(1)
    r = w.read_bodies()
    dt = float(np.float32(0.01667))
    om = 3.0 * math.exp(-dt * 0.1)  # advect damps before integration
    ang = om * dt
    q = r["rotation"][0]
    assert abs(q[1] - math.sin(ang / 2)) < 1e-6 and abs(q[3] - math.cos(ang / 2)) < 1e-6


def test_avbd_free_fall_closed_form():
    body = _box((0.0, 50.0, 0.0))
    body["velocity"] = (1.0, 0.5, 0.0)
    w = _world(body, solver=scenes.SOLVER_AVBD)
    dt, g = float(np.float32(0.01667)), -9.81
    x = np.array([0.0, 50.0, 0.0])
    v = np.array([1.0, 0.5, 0.0])
    for _ in range(30):
        w.step(1)
        x = x + v * dt + np.array([0, g, 0]) * dt * dt
        v = v + np.array([0, g, 0]) * dt
    r = w.read_bodies()
    assert np.allclose(r["position"][0], x, atol=3e-4)
    assert np.allclose(r["velocity"][0], v, atol=3e-3)


def test_pgs_soft_coefficients_closed_form():
    g = np.zeros(1, scenes.BODY_DESC)
    scenes._ground(g[0])
    w = _world(np.concatenate([g, _box((0.0, 0.49, 0.0))]), iterations=8)
    w.step(1)
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    assert len(m) == 1 and m[0]["contact_count"] == 4
    dt = float(np.float32(0.01667))
    hertz = min(30.0, (1.0 / dt) / 3.0) * 2.0  # doubled against a static body
    omega = 2 * math.pi * hertz
    cc = dt * omega * (2 * 10.0 + dt * omega)
    mass_coef = cc / (1 + cc)
    assert np.allclose(m[0]["contacts"]["mass_coefficient"][:4], mass_coef, rtol=1e-5)


def test_resting_box_normal_impulse_balances_gravity():
    g = np.zeros(1, scenes.BODY_DESC)
    scenes._ground(g[0])
    w = _world(np.concatenate([g, _box((0.0, 0.5, 0.0))]), iterations=8,
               flags=scenes.SIM_DEFAULT_FLAGS & ~128)
    w.step(120)
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    lam = m[0]["contacts"]["normal_impulse"][:m[0]["contact_count"]].sum()
    assert abs(lam - 1.0 * 9.81 * float(np.float32(0.01667))) < 0.02 * 9.81 * 0.01667


def test_sleep_timer_cycle_for_isolated_body():
    body = _box((0.0, 50.0, 0.0))
    body["flags"] = scenes.BODY_DYNAMIC  # no gravity, no contacts, at rest
    w = _world(body)
    states = []
    for _ in range(60):
        w.step(1)
        states.append(bool(w.read_bodies()["flags"][0] & 8))
    # calm: timer +1 per step up to 30 -> sleeps on step index 29; a sleeper
    # with no manifold bleeds 10 (timer 20, wakes), then 10 calm steps re-arm
    # it on the 11th: a period of 11 steps
    asleep = [i for i, s in enumerate(states) if s]
    assert asleep[:3] == [29, 40, 51]


# -------------------------------------------------------------- invariants
def test_invariants_config1():
    w = _world(scenes.cubes_grid(), iterations=8)
    n = 1001
    for s in range(90):
        w.step(1)
        st = w.stats()
        assert st["dbg_id_sum"] == n * (n - 1) // 2
        assert st["graph_color_violations"] == 0
        assert st["manifold_node_count"] == 2 * st["collision_count"]
    # edge colouring validity: no dynamic body twice in one colour
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    rows = w.tap(beatbox.TAP_ROWS, beatbox.BODY_STATE)
    dyn = (rows["flags"] & 2) != 0
    seen = set()
    for mm in m:
        for b in (mm["obb1_index"], mm["obb2_index"]):
            if dyn[b] and mm["color"] < 32:
                assert (b, mm["color"]) not in seen
                seen.add((b, mm["color"]))


def test_islands_equal_bfs_components():
    w = _world(scenes.random_boxes(0xC020, 150, extent=3.0, y0=0.55, y1=4.0), iterations=4)
    w.step(45)
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    rows = w.tap(beatbox.TAP_ROWS, beatbox.BODY_STATE)
    isl = w.tap(TAP_ISLANDS, np.uint32)
    N = len(rows)
    dyn = (rows["flags"] & 2) != 0
    adj = {i: set() for i in range(N)}
    for mm in m:
        a, b = int(mm["obb1_index"]), int(mm["obb2_index"])
        if dyn[a] and dyn[b]:
            adj[a].add(b)
            adj[b].add(a)
    # active index = rank of the id among dynamic ids
    ids = rows["id"]
    dyn_ids = np.sort(ids[dyn])
    rank = {int(i): k for k, i in enumerate(dyn_ids)}
    seen = set()
    for i in range(N):
        if not dyn[i] or i in seen:
            continue
        comp, stack = [], [i]
        seen.add(i)
        while stack:
            x = stack.pop()
            comp.append(x)
            for y in adj[x]:
                if y not in seen:
                    seen.add(y)
                    stack.append(y)
        root = min(rank[int(ids[x])] for x in comp)
        assert all(isl[x] == root for x in comp)
    assert np.all(isl[~dyn] == 0xFFFFFFFF)
    assert w.stats()["island_count"] == len(set(isl[dyn].tolist()))


def test_avbd_colouring_valid():
    w = _world(scenes.random_boxes(0xC021, 150, extent=3.0, y0=0.55, y1=4.0), solver=scenes.SOLVER_AVBD)
    for _ in range(40):
        w.step(1)
        assert w.stats()["avbd_violations"] == 0
    bc = w.tap(TAP_BODY_COLORS, np.uint32)
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    rows = w.tap(beatbox.TAP_ROWS, beatbox.BODY_STATE)
    dyn = (rows["flags"] & 2) != 0
    for mm in m:
        a, b = int(mm["obb1_index"]), int(mm["obb2_index"])
        if dyn[a] and dyn[b]:
            assert bc[a] != bc[b]


# ------------------------------------------------- per-island path, plain PGS
ISLAND_FLAGS = scenes.SIM_DEFAULT_FLAGS & ~scenes.SIM_USE_GRAPH_COLORING


def _sphere_column(masses, speeds, overlap=0.03):
    """Ground plus unit-diameter spheres stacked on the y axis, each pair and
    the lowest sphere with the ground overlapping by `overlap`; no gravity,
    vertical initial velocities.  Every contact normal is vertical and every
    arm parallel to it, so the solve reduces to 1-D impulses."""
    g = np.zeros(1, scenes.BODY_DESC)
    scenes._ground(g[0])
    parts = [g]
    for k, (m, v) in enumerate(zip(masses, speeds)):
        s = _sphere((0.0, 0.5 - overlap + k * (1.0 - overlap), 0.0), 0.5, mass=m)
        s["flags"] = scenes.BODY_DYNAMIC
        s["velocity"] = (0.0, v, 0.0)
        parts.append(s)
    return np.concatenate(parts)


def _column_model(w, bodies, iterations, order, accum=True):
    """Sequential impulses of PGS::pre_step / solve in 1-D over the
    manifolds in the given order (first step, nothing to warm-start):
    effective mass 1/(1/ma + 1/mb), Baumgarte bias 0.2/dt * (depth - 0.01),
    accumulated impulses clamped at zero."""
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    ids = w.read_bodies(order=beatbox.ORDER_BY_ROW)["id"]
    dt = float(np.float32(0.01667))
    damp = math.exp(-dt * 0.1)
    v = {int(i): float(bodies["velocity"][i][1]) * damp for i in range(len(bodies))}
    inv = {int(i): (1.0 / float(bodies["mass"][i]) if bodies["flags"][i] & scenes.BODY_DYNAMIC else 0.0)
           for i in range(len(bodies))}
    rows = []
    for t in m:
        assert t["contact_count"] == 1
        n = t["normal"]
        assert n[0] == 0.0 and n[2] == 0.0 and abs(n[1]) == 1.0
        rows.append((int(ids[t["obb1_index"]]), int(ids[t["obb2_index"]]), float(n[1]),
                     float(t["contacts"]["penetration"][0])))
    lam = [0.0] * len(rows)
    for _ in range(iterations):
        for k in order:
            a, b, s, pen = rows[k]
            nm = 1.0 / (inv[a] + inv[b])
            bias = -0.2 / dt * min(pen + 0.01, 0.0)
            imp = nm * (-(s * (v[b] - v[a])) + bias)
            if accum:
                new = max(lam[k] + imp, 0.0)
                imp, lam[k] = new - lam[k], new
            else:
                imp = max(imp, 0.0)
            v[a] -= s * imp * inv[a]
            v[b] += s * imp * inv[b]
    return v, lam, m


@pytest.mark.parametrize("accum", [True, False])
def test_island_plain_pgs_is_serial_gauss_seidel(accum):
    """USE_GRAPH_COLORING off with PGS: one step of a sphere column equals
    1-D sequential impulses in packed (pair) order, and not in any other
    order nor with PGS_SOFT's soft constraint, so the check sees both the
    solver formula and the walk order."""
    bodies = _sphere_column([1.0, 2.0, 0.5, 3.0, 1.0], [-1.0, 0.5, -2.0, 0.7, -1.5])
    flags = ISLAND_FLAGS if accum else ISLAND_FLAGS & ~1
    w = _world(bodies, solver=scenes.SOLVER_PGS, iterations=3, flags=flags, morton_bits=63, quirks=0)
    w.step(1)
    r = w.read_bodies()
    M = len(w.tap(TAP_MANIFOLDS, MANIFOLD_TAP))
    assert M == 5
    v, lam, m = _column_model(w, bodies, 3, range(M), accum)
    got = {int(i): float(r["velocity"][i][1]) for i in range(len(bodies))}
    for i in range(1, len(bodies)):
        assert abs(got[i] - v[i]) < 1e-4, (i, got[i], v[i])
        assert r["velocity"][i][0] == 0.0 and r["velocity"][i][2] == 0.0
    if accum:
        assert np.allclose(m["contacts"]["normal_impulse"][:, 0], lam, atol=1e-4)
    # another walk order gives a different answer
    v2, _, _ = _column_model(w, bodies, 3, list(reversed(range(M))), accum)
    assert max(abs(v2[i] - v[i]) for i in range(1, len(bodies))) > 1e-2
    # and so does PGS_SOFT on the same path
    s = _world(bodies, solver=scenes.SOLVER_PGS_SOFT, iterations=3, flags=flags, morton_bits=63, quirks=0)
    s.step(1)
    assert np.abs(s.read_bodies()["velocity"] - r["velocity"]).max() > 1e-2


def test_colour_path_runs_pgs_as_pgs_soft_and_island_path_does_not():
    """With colouring on, PGS is PGS_SOFT (the reference's colour sweeps
    implement only that); with it off the two solvers differ."""
    bodies = scenes.random_boxes(0xBB21, 48, extent=2.0, y0=0.55, y1=4.0)
    out = {}
    for solver in (scenes.SOLVER_PGS, scenes.SOLVER_PGS_SOFT):
        for flags in (scenes.SIM_DEFAULT_FLAGS, ISLAND_FLAGS):
            w = _world(bodies, solver=solver, iterations=8, flags=flags)
            w.step(6)
            out[solver, flags] = w.read_bodies()
    a, b = out[scenes.SOLVER_PGS, scenes.SIM_DEFAULT_FLAGS], out[scenes.SOLVER_PGS_SOFT, scenes.SIM_DEFAULT_FLAGS]
    assert a.tobytes() == b.tobytes()
    c, d = out[scenes.SOLVER_PGS, ISLAND_FLAGS], out[scenes.SOLVER_PGS_SOFT, ISLAND_FLAGS]
    assert np.abs(c["velocity"] - d["velocity"]).max() > 1e-3
    assert np.abs(d["velocity"] - b["velocity"]).max() > 1e-3


def test_island_path_islands_do_not_interact():
    """Two sphere columns far apart in one world (two contact islands) step
    exactly as each column alone: the walk over all manifolds in packed order
    equals one serial walk per island."""
    a = _sphere_column([1.0, 2.0, 0.5], [-1.0, 0.5, -2.0])
    b = _sphere_column([3.0, 1.0], [0.7, -1.5])[1:]
    b["position"][:, 0] += 20.0
    both = np.concatenate([a, b])
    w = _world(both, solver=scenes.SOLVER_PGS, iterations=3, flags=ISLAND_FLAGS, morton_bits=63, quirks=0)
    w.step(3)
    r = w.read_bodies()
    for part, sl in ((a, slice(0, len(a))), (np.concatenate([a[:1], b]), None)):
        x = _world(part, solver=scenes.SOLVER_PGS, iterations=3, flags=ISLAND_FLAGS, morton_bits=63, quirks=0)
        x.step(3)
        rx = x.read_bodies()
        if sl is not None:
            assert rx["velocity"].tobytes() == r["velocity"][sl].tobytes()
        else:
            assert rx["velocity"][1:].tobytes() == r["velocity"][len(a):].tobytes()
