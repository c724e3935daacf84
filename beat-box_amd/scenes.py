"""Seeded synthetic scenes for the rigid-body step.

Every draw comes from SplitMix64 with an explicit seed: output k uses the state
seed + (k+1)*0x9E3779B97F4A7C15, so draws are a pure function of (seed, k) and
are vectorised with numpy.  A float in [0, 1) is (u >> 40) * 2**-24.  Draws are
made in body-id order, so CPU and GPU consume identical bytes.

Scenes return a numpy structured array with the layout of `bb_body_desc`
(include/bb_capi.h) plus the world descriptor fields to use.
"""
from __future__ import annotations

import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
MIX1 = np.uint64(0xBF58476D1CE4E5B9)
MIX2 = np.uint64(0x94D049BB133111EB)

# bb_body_desc, include/bb_capi.h
BODY_DESC = np.dtype(
    [
        ("shape", "<u4"),
        ("flags", "<u4"),
        ("position", "<f4", 3),
        ("rotation", "<f4", 4),
        ("minimum", "<f4", 3),
        ("maximum", "<f4", 3),
        ("radius", "<f4"),
        ("mass", "<f4"),
        ("velocity", "<f4", 3),
        ("omega", "<f4", 3),
        ("restitution", "<f4"),
        ("friction", "<f4"),
    ]
)
assert BODY_DESC.itemsize == 100

SHAPE_BOX, SHAPE_SPHERE = 0, 1
BODY_DYNAMIC, BODY_GRAVITY = 2, 4
SOLVER_PGS, SOLVER_PGS_SOFT, SOLVER_AVBD = 0, 1, 2
SIM_DEFAULT_FLAGS = 1 | 2 | 4 | 8 | 16 | 32 | 128
SIM_USE_GRAPH_COLORING = 32
SIM_SLEEPING_ENABLED = 128
QUIRKS_PARITY = 1 | 2 | 4
TIME_STEP = np.float32(0.01667)


def splitmix64(seed: int, counters: np.ndarray) -> np.ndarray:
    """SplitMix64 outputs for the given counters (vectorised)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (counters.astype(np.uint64) + np.uint64(1)) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * MIX1
        z = (z ^ (z >> np.uint64(27))) * MIX2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, counters: np.ndarray, lo: float, hi: float) -> np.ndarray:
    u = (splitmix64(seed, counters) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
    return lo + (hi - lo) * u


def world_desc(solver=SOLVER_PGS_SOFT, iterations=8, flags=SIM_DEFAULT_FLAGS,
               dt=float(TIME_STEP), gravity=-9.81, morton_bits=30,
               quirks=QUIRKS_PARITY, capacity=0) -> dict:
    return dict(capacity_bodies=capacity, solver=solver, iterations=iterations,
                flags=flags, dt=dt, gravity=gravity, morton_bits=morton_bits,
                quirks=quirks)


def _ground(body: np.ndarray, half: float = 50.0, friction: float = 0.5) -> None:
    """Static ground box, centre (0,-half,0): the reference floor
    (src/scene_manager.hpp:257)."""
    body["shape"] = SHAPE_BOX
    body["flags"] = 0
    body["position"] = (0.0, -half, 0.0)
    body["rotation"] = (0.0, 0.0, 0.0, 1.0)
    body["minimum"] = (-half, -half, -half)
    body["maximum"] = (half, half, half)
    body["mass"] = 0.0
    body["friction"] = friction
    body["restitution"] = 0.0


def yaw_quat(theta: np.ndarray) -> np.ndarray:
    q = np.zeros((theta.shape[0], 4), np.float64)
    q[:, 1] = np.sin(theta * 0.5)
    q[:, 3] = np.cos(theta * 0.5)
    return q.astype(np.float32)


def cubes_grid(seed: int = 0xBB01, nx: int = 10, ny: int = 10, nz: int = 10,
               spacing: float = 1.5, jitter: float = 0.1, base_y: float = 2.5):
    """Plain method small: 1,000 unit cubes (density 1, friction 0.5,
    restitution 0) on a 10x10x10 grid, 1.5 m spacing, jitter U(-0.1,0.1),
    random yaw, dropped from 2 m onto the static ground.  Body 0 = ground."""
    n = nx * ny * nz
    bodies = np.zeros(n + 1, BODY_DESC)
    _ground(bodies[0])
    idx = np.arange(n)
    i = idx % nx
    k = (idx // nx) % nz
    j = idx // (nx * nz)
    F = 4
    c = idx.astype(np.uint64) * np.uint64(F)
    jx = uniform(seed, c + np.uint64(0), -jitter, jitter)
    jy = uniform(seed, c + np.uint64(1), -jitter, jitter)
    jz = uniform(seed, c + np.uint64(2), -jitter, jitter)
    yaw = uniform(seed, c + np.uint64(3), 0.0, 2.0 * math.pi)
    b = bodies[1:]
    b["shape"] = SHAPE_BOX
    b["flags"] = BODY_DYNAMIC | BODY_GRAVITY
    b["position"][:, 0] = ((i - (nx - 1) / 2.0) * spacing + jx).astype(np.float32)
    b["position"][:, 1] = (base_y + spacing * j + jy).astype(np.float32)
    b["position"][:, 2] = ((k - (nz - 1) / 2.0) * spacing + jz).astype(np.float32)
    b["rotation"] = yaw_quat(yaw)
    b["minimum"] = -0.5
    b["maximum"] = 0.5
    b["mass"] = 1.0
    b["friction"] = 0.5
    b["restitution"] = 0.0
    return bodies


def random_boxes(seed: int, n: int, extent: float = 6.0, y0: float = 0.6,
                 y1: float = 6.0, half: float = 0.5, with_ground: bool = True,
                 tilt: bool = True):
    """Small random scenes for parity tests: n cubes in a box region with a
    random orientation (Shoemake) and the ground box."""
    bodies = np.zeros(n + (1 if with_ground else 0), BODY_DESC)
    off = 0
    if with_ground:
        _ground(bodies[0])
        off = 1
    idx = np.arange(n, dtype=np.uint64) * np.uint64(8)
    b = bodies[off:]
    b["shape"] = SHAPE_BOX
    b["flags"] = BODY_DYNAMIC | BODY_GRAVITY
    b["position"][:, 0] = uniform(seed, idx + np.uint64(0), -extent, extent)
    b["position"][:, 1] = uniform(seed, idx + np.uint64(1), y0, y1)
    b["position"][:, 2] = uniform(seed, idx + np.uint64(2), -extent, extent)
    if tilt:
        b["rotation"] = shoemake(seed, idx + np.uint64(3))
    else:
        b["rotation"] = yaw_quat(uniform(seed, idx + np.uint64(3), 0.0, 2 * math.pi))
    b["minimum"] = -half
    b["maximum"] = half
    b["mass"] = 1.0
    b["friction"] = 0.5
    return bodies


def shoemake(seed: int, c: np.ndarray) -> np.ndarray:
    """Uniform random rotation (Shoemake 1992) from three draws c, c+1, c+2."""
    u1 = uniform(seed, c, 0.0, 1.0)
    u2 = uniform(seed, c + np.uint64(1), 0.0, 2 * math.pi)
    u3 = uniform(seed, c + np.uint64(2), 0.0, 2 * math.pi)
    a = np.sqrt(1.0 - u1)
    bb = np.sqrt(u1)
    q = np.stack([a * np.sin(u2), a * np.cos(u2), bb * np.sin(u3), bb * np.cos(u3)], 1)
    return q.astype(np.float32)


def batch_world_sizes(seed: int = 0xBB04, worlds: int = 4096, lo: int = 16, hi: int = 2048):
    u = uniform(seed, np.arange(worlds, dtype=np.uint64) * np.uint64(4),
                math.log(lo), math.log(hi))
    return np.clip(np.rint(np.exp(u)), lo, hi).astype(np.int64)


def batch_world(seed: int, w: int, n_boxes: int):
    """One batched world: even w = planar box pyramid (rows L..1 truncated to
    n_boxes), odd w = 4x4-footprint tower of ceil(n/16) layers; box half-extent
    U(0.3, 0.7) per world; its own ground box (body 0)."""
    h = float(uniform(seed, np.array([w * 4 + 1], np.uint64), 0.3, 0.7)[0])
    bodies = np.zeros(n_boxes + 1, BODY_DESC)
    _ground(bodies[0])
    pos = np.zeros((n_boxes, 3), np.float64)
    if w % 2 == 0:
        L = 1
        while L * (L + 1) // 2 < n_boxes:
            L += 1
        k = 0
        gap = 2.0 * h * 1.04
        for r in range(L):
            cnt = L - r
            for c in range(cnt):
                if k >= n_boxes:
                    break
                pos[k] = ((c - (cnt - 1) / 2.0) * gap, h * 1.002 + r * 2.0 * h * 1.002, 0.0)
                k += 1
    else:
        gap = 2.0 * h * 1.03
        for k in range(n_boxes):
            layer, cell = divmod(k, 16)
            cx, cz = divmod(cell, 4)
            pos[k] = ((cx - 1.5) * gap, h * 1.002 + layer * 2.0 * h * 1.002, (cz - 1.5) * gap)
    b = bodies[1:]
    b["shape"] = SHAPE_BOX
    b["flags"] = BODY_DYNAMIC | BODY_GRAVITY
    b["position"] = pos.astype(np.float32)
    b["rotation"] = (0.0, 0.0, 0.0, 1.0)
    b["minimum"] = -h
    b["maximum"] = h
    b["mass"] = np.float32((2 * h) ** 3)
    b["friction"] = 0.5
    return bodies


def batch_worlds(seed: int = 0xBB04, worlds: int = 4096, lo: int = 16, hi: int = 2048):
    """Batched worlds: body counts log-uniform in [lo, hi]."""
    sizes = batch_world_sizes(seed, worlds, lo, hi)
    return [batch_world(seed, w, int(n)) for w, n in enumerate(sizes)]


def _mixed_pile(seed: int, n: int, ext_lo: float, ext_hi: float, x_half: float, y_lo: float, y_hi: float,
                with_ground: bool = True):
    """n bodies (even ids boxes with per-axis half-extents U(ext_lo, ext_hi),
    odd ids spheres with radius U(ext_lo, ext_hi)), density 1, uniform random
    orientation (Shoemake), centres x, z U(-x_half, x_half), y U(y_lo, y_hi),
    zero velocity; body 0 = the static ground when with_ground."""
    off = 1 if with_ground else 0
    bodies = np.zeros(n + off, BODY_DESC)
    if with_ground:
        _ground(bodies[0])
    F = np.uint64(12)
    c = np.arange(n, dtype=np.uint64) * F
    b = bodies[off:]
    sphere = (np.arange(n) % 2) == 1
    hx = uniform(seed, c + np.uint64(0), ext_lo, ext_hi)
    hy = uniform(seed, c + np.uint64(1), ext_lo, ext_hi)
    hz = uniform(seed, c + np.uint64(2), ext_lo, ext_hi)
    b["shape"] = np.where(sphere, SHAPE_SPHERE, SHAPE_BOX)
    b["flags"] = BODY_DYNAMIC | BODY_GRAVITY
    b["position"][:, 0] = uniform(seed, c + np.uint64(3), -x_half, x_half)
    b["position"][:, 1] = uniform(seed, c + np.uint64(4), y_lo, y_hi)
    b["position"][:, 2] = uniform(seed, c + np.uint64(5), -x_half, x_half)
    b["rotation"] = shoemake(seed, c + np.uint64(6))
    r = hx
    b["radius"] = np.where(sphere, r, 0.0).astype(np.float32)
    b["minimum"] = np.where(sphere[:, None], -r[:, None], -np.stack([hx, hy, hz], 1)).astype(np.float32)
    b["maximum"] = -b["minimum"]
    vol = np.where(sphere, 4.0 / 3.0 * math.pi * r ** 3, 8.0 * hx * hy * hz)
    b["mass"] = vol.astype(np.float32)
    b["friction"] = 0.5
    return bodies


def large_pile(seed: int = 0xBB02, n: int = 100_000):
    """Single large pile: 50/50 boxes and spheres, half-extents/radii
    U(0.25, 1.0) m, centres uniform in a 60 x 200 x 60 m column above the
    ground (y in [1, 201]), zero velocity.  Run with 63-bit keys and AVBD,
    10 iterations."""
    return _mixed_pile(seed, n, 0.25, 1.0, 30.0, 1.0, 201.0)


def sphere_gas(seed: int = 0xBB03, n: int = 1_000_000, half: float = 500.0):
    """Broad-phase stress: n spheres, radius U(0.1, 0.5) m, centres uniform
    in a (2*half)^3 cube, velocity N(0, 2 m/s) per axis (Box-Muller), no
    ground; run with gravity 0."""
    bodies = np.zeros(n, BODY_DESC)
    c = np.arange(n, dtype=np.uint64) * np.uint64(10)
    r = uniform(seed, c, 0.1, 0.5)
    bodies["shape"] = SHAPE_SPHERE
    bodies["flags"] = BODY_DYNAMIC
    for k in range(3):
        bodies["position"][:, k] = uniform(seed, c + np.uint64(1 + k), -half, half)
    bodies["rotation"] = (0.0, 0.0, 0.0, 1.0)
    bodies["radius"] = r
    bodies["minimum"] = -r[:, None]
    bodies["maximum"] = r[:, None]
    bodies["mass"] = (4.0 / 3.0 * math.pi * r ** 3).astype(np.float32)
    for k in range(3):
        u1 = uniform(seed, c + np.uint64(4 + 2 * k), 2.0 ** -24, 1.0)
        u2 = uniform(seed, c + np.uint64(5 + 2 * k), 0.0, 2.0 * math.pi)
        bodies["velocity"][:, k] = (2.0 * np.sqrt(-2.0 * np.log(u1)) * np.cos(u2)).astype(np.float32)
    bodies["friction"] = 0.5
    return bodies


def dense_spheres(seed: int, side: int = 5, radius: float = 0.5, spacing: float = 0.45):
    """Pair-capacity stress: side^3 spheres of one radius on a cubic grid
    whose spacing is below the radius, so each sphere's box overlaps the
    boxes of the spheres up to two grid steps away on every axis (about 60
    pairs per sphere, well over 8 per body); centres jittered by
    U(-0.02, 0.02), no ground; run with gravity 0."""
    n = side ** 3
    bodies = np.zeros(n, BODY_DESC)
    idx = np.arange(n)
    c = idx.astype(np.uint64) * np.uint64(3)
    bodies["shape"] = SHAPE_SPHERE
    bodies["flags"] = BODY_DYNAMIC
    for k, g in enumerate((idx % side, (idx // side) % side, idx // (side * side))):
        bodies["position"][:, k] = ((g - (side - 1) / 2.0) * spacing
                                    + uniform(seed, c + np.uint64(k), -0.02, 0.02)).astype(np.float32)
    bodies["rotation"] = (0.0, 0.0, 0.0, 1.0)
    bodies["radius"] = radius
    bodies["minimum"] = -radius
    bodies["maximum"] = radius
    bodies["mass"] = np.float32(4.0 / 3.0 * math.pi * radius ** 3)
    bodies["friction"] = 0.5
    return bodies


def expanding_spheres(seed: int, side: int = 5, rate: float = 100.0):
    """Overflow in the first step only: the dense_spheres block moving
    apart with velocity rate * position (a homologous expansion).  The first
    broad phase sees every sphere's box overlap its neighbours' (several
    times the 8n+64 pair slots); a step later the spacing has grown past the
    diameter and the pairs are gone.  Run with gravity 0."""
    bodies = dense_spheres(seed, side=side)
    bodies["velocity"] = (bodies["position"].astype(np.float64) * rate).astype(np.float32)
    return bodies


def crossing_cloud(seed: int, planks: int = 20, side: int = 5, speed: float = 20.0):
    """Overflow in a few middle steps only: `planks` static thin boards (half
    extents 10 x 0.05 x 0.05 m, turned 45 degrees about z, all on the
    diagonal y = x through the origin, so each board's box is a 14 x 14 m
    square 0.1 m thick in z) and a side x side layer of spheres (r 0.3 m, 2 m
    apart, at least 1.5 m off the diagonal) 3 m below them in z, moving
    through their plane at `speed` m/s along +z.  The boards' mutual pairs
    (planks * (planks - 1) / 2) fit the 8n + 64 pair slots; while the layer
    crosses the boards' plane every sphere's box overlaps every board's
    (planks more pairs per sphere), which does not; no sphere touches a
    board or another sphere, so the layer flies on unchanged and the pairs
    are gone again.  Run with gravity 0."""
    pl = np.zeros(planks, BODY_DESC)
    pl["shape"] = SHAPE_BOX
    pl["flags"] = 0
    c = math.cos(math.pi / 8.0)
    sn = math.sin(math.pi / 8.0)
    pl["rotation"] = (0.0, 0.0, sn, c)
    pl["minimum"] = (-10.0, -0.05, -0.05)
    pl["maximum"] = (10.0, 0.05, 0.05)
    pl["friction"] = 0.5
    pts = []
    k = np.arange(-(side - 1), side, 2, dtype=np.float64) * 1.0
    for x in k:
        for y in k:
            if abs(y - x) >= 1.5:
                pts.append((x, y))
    cnt = np.arange(len(pts), dtype=np.uint64) * np.uint64(3)
    jitter = uniform(seed, cnt, -0.05, 0.05)
    sp = np.zeros(len(pts), BODY_DESC)
    sp["shape"] = SHAPE_SPHERE
    sp["flags"] = BODY_DYNAMIC
    for i, (x, y) in enumerate(pts):
        sp["position"][i] = (x + jitter[i], y - jitter[i], -3.0)
    sp["rotation"] = (0.0, 0.0, 0.0, 1.0)
    sp["radius"] = 0.3
    sp["minimum"] = -0.3
    sp["maximum"] = 0.3
    sp["mass"] = (4.0 / 3.0 * math.pi * 0.3 ** 3)
    sp["velocity"] = (0.0, 0.0, speed)
    sp["friction"] = 0.5
    return np.concatenate([pl, sp])


def serving_pile(seed: int, n: int):
    """One serving request: a random pile of n bodies, 50/50 boxes and
    spheres with sizes (edge / diameter) U(0.2, 1.0) m, centres in a column
    whose footprint grows with n, on the ground."""
    x_half = max(2.0, 0.45 * n ** (1.0 / 3.0))
    height = max(4.0, n * 0.5 / (4.0 * x_half * x_half))
    return _mixed_pile(seed, n, 0.1, 0.5, x_half, 0.6, 0.6 + height)


def serving_requests(seed: int = 0xBB05, rate: float = 500.0, duration: float = 600.0,
                     lo: int = 64, hi: int = 20_000, h_lo: int = 60, h_hi: int = 600):
    """Serving stream: Poisson arrivals (exponential gaps at `rate` per s)
    within `duration` seconds; each request has a body count log-uniform in
    [lo, hi] and a horizon uniform in {h_lo..h_hi} steps.  Returns arrays
    (arrival_s, bodies, horizon, scene_seed)."""
    m = int(rate * duration * 1.2 + 64)
    c = np.arange(m, dtype=np.uint64) * np.uint64(4)
    gaps = -np.log(uniform(seed, c, 2.0 ** -24, 1.0)) / rate
    t = np.cumsum(gaps)
    keep = t < duration
    nb = np.clip(np.rint(np.exp(uniform(seed, c + np.uint64(1), math.log(lo), math.log(hi)))), lo, hi)
    hz = np.floor(uniform(seed, c + np.uint64(2), h_lo, h_hi + 1)).astype(np.int64)
    hz = np.clip(hz, h_lo, h_hi)
    sd = splitmix64(seed, c + np.uint64(3))
    return t[keep], nb[keep].astype(np.int64), hz[keep], sd[keep]


# ------------------------------------------------------------ voxel bodies
SHAPE_VOXEL = 2


def voxel_shape_word(shape_id: int) -> int:
    """Body shape word of voxel shape `shape_id` (1-based)."""
    return SHAPE_VOXEL | (int(shape_id) << 8)


def voxel_grid(dims, solid) -> np.ndarray:
    """Boolean cells g[x, y, z] of a dims grid where solid(x, y, z) holds."""
    x, y, z = np.meshgrid(*(np.arange(n) for n in dims), indexing="ij")
    return np.asarray(solid(x, y, z), bool) & np.ones(dims, bool)


# The concave shapes of the reference's voxel scene (src/scene_manager.hpp:
# 626-638): an L, a cross and a square frame with a hole, 0.5 m cells,
# density 2.
VOXEL_SIZE, VOXEL_DENSITY = 0.5, 2.0
VOXEL_SHAPES = {
    "L": voxel_grid((6, 6, 2), lambda x, y, z: (y < 2) | (x < 2)),
    "cross": voxel_grid((6, 6, 2), lambda x, y, z: ((x >= 2) & (x < 4)) | ((y >= 2) & (y < 4))),
    "frame": voxel_grid((8, 8, 2), lambda x, y, z: ~((x >= 2) & (x < 6) & (y >= 2) & (y < 6))),
}


def voxel_body(shape_id: int, pos, rot=(0.0, 0.0, 0.0, 1.0), friction: float = 0.6,
               dynamic: bool = True) -> np.ndarray:
    b = np.zeros(1, BODY_DESC)
    b["shape"] = voxel_shape_word(shape_id)
    b["flags"] = (BODY_DYNAMIC | BODY_GRAVITY) if dynamic else 0
    b["position"] = pos
    b["rotation"] = rot
    b["friction"] = friction
    return b


def _static_box(pos, half, friction=0.6) -> np.ndarray:
    b = np.zeros(1, BODY_DESC)
    b["shape"] = SHAPE_BOX
    b["position"] = pos
    b["rotation"] = (0.0, 0.0, 0.0, 1.0)
    b["minimum"] = tuple(-h for h in half)
    b["maximum"] = half
    b["friction"] = friction
    return b


def voxel_pile():
    """The reference's concave-body scene (src/scene_manager.hpp:614-655):
    floor, a static panel and a static post, three Ls, three crosses and
    three frames (the frames' holes fit the post).  Returns the shapes in id
    order (name, cells) and the bodies; shape ids follow that order."""
    shapes = [("L", VOXEL_SHAPES["L"]), ("cross", VOXEL_SHAPES["cross"]), ("frame", VOXEL_SHAPES["frame"])]
    g = np.zeros(1, BODY_DESC)
    _ground(g[0], friction=0.6)
    q_id, q_x90 = (0.0, 0.0, 0.0, 1.0), (0.7071, 0.0, 0.0, 0.7071)
    parts = [g, _static_box((0.0, 22.0, 14.0), (5.0, 0.2, 5.0)), _static_box((8.0, 4.0, 14.0), (0.35, 4.0, 0.35))]
    for sid, pos, rot in [(1, (-6.0, 2.5, 13.0), q_id), (1, (-5.4, 5.5, 14.6), q_x90), (1, (-2.0, 8.0, 13.8), q_id),
                          (2, (-3.5, 2.5, 14.8), q_id), (2, (-4.5, 11.0, 14.0), q_x90), (2, (0.5, 4.0, 14.2), q_id),
                          (3, (-1.0, 12.5, 14.4), q_id), (3, (2.5, 7.0, 13.4), q_x90), (3, (8.0, 11.0, 14.0), q_x90)]:
        parts.append(voxel_body(sid, pos, rot))
    return shapes, np.concatenate(parts)


def voxel_mixed_pile(seed: int, n: int):
    """A random pile of n bodies on the ground: voxel Ls, crosses and frames
    mixed with boxes and spheres, random yaw, dropped from a column.  Uses
    shape ids 1..3 in VOXEL_SHAPES order, like voxel_pile."""
    shapes = [("L", VOXEL_SHAPES["L"]), ("cross", VOXEL_SHAPES["cross"]), ("frame", VOXEL_SHAPES["frame"])]
    c = np.arange(n, dtype=np.uint64) * np.uint64(8)
    kind = np.floor(uniform(seed, c, 0.0, 5.0)).astype(np.int64)
    half = max(4.0, 1.6 * n ** 0.5)
    px = uniform(seed, c + np.uint64(1), -half, half)
    pz = uniform(seed, c + np.uint64(2), -half, half)
    py = 2.5 + np.arange(n) * 4.5 / max(1.0, n / 6.0)
    yaw = uniform(seed, c + np.uint64(3), 0.0, 2.0 * math.pi)
    q = yaw_quat(yaw)
    g = np.zeros(1, BODY_DESC)
    _ground(g[0], friction=0.6)
    parts = [g]
    for i in range(n):
        rot = tuple(float(v) for v in q[i])
        pos = (float(px[i]), float(py[i]), float(pz[i]))
        if kind[i] < 3:
            parts.append(voxel_body(int(kind[i]) + 1, pos, rot))
        else:
            b = np.zeros(1, BODY_DESC)
            b["flags"] = BODY_DYNAMIC | BODY_GRAVITY
            b["position"] = pos
            b["rotation"] = rot
            b["mass"] = 1.0
            b["friction"] = 0.6
            if kind[i] == 3:
                b["shape"] = SHAPE_BOX
                b["minimum"] = (-0.5, -0.5, -0.5)
                b["maximum"] = (0.5, 0.5, 0.5)
            else:
                b["shape"] = SHAPE_SPHERE
                b["radius"] = 0.5
            parts.append(b)
    return shapes, np.concatenate(parts)


# ---------------------------------------------------------------------------
# The reference's own test scenes (src/scene_manager.hpp), as body lists in
# load order (body id = position in the list).  They carry the reference's
# known answers (towers stand, ramp stick/slide outcomes and resting friction
# force, every cube of the box pool asleep) and are stepped in parity mode.
# Floats are computed as the reference computes them, in float32.

def _ref_body(pos, rot, half, mass, friction, dynamic):
    b = np.zeros(1, BODY_DESC)[0]
    b["shape"] = SHAPE_BOX
    b["flags"] = (BODY_DYNAMIC | BODY_GRAVITY) if dynamic else 0
    b["position"] = np.asarray(pos, np.float32)
    b["rotation"] = np.asarray(rot, np.float32)
    h = np.asarray(half, np.float32)
    b["minimum"] = -h
    b["maximum"] = h
    b["mass"] = mass
    b["friction"] = friction
    return b


def _ref_floor_and_panel(friction, panel_pos, panel_half):
    return [_ref_body((0.0, -50.0, 0.0), (0, 0, 0, 1), (50.0, 50.0, 50.0), 0.0, friction, False),
            _ref_body(panel_pos, (0, 0, 0, 1), panel_half, 0.0, friction, False)]


def reference_towers():
    """scene_3 (src/scene_manager.hpp:327-353): floor, a static panel and three
    towers of 5, 8 and 12 unit cubes (mass 5, friction 0.5) at x = -6, 0, 6,
    z = 14, spawned with a 1 mm gap per level.  Known answer: they stand."""
    out = _ref_floor_and_panel(0.5, (0.0, 22.0, 14.0), (5.0, 0.2, 5.0))
    for x, h in ((-6.0, 5), (0.0, 8), (6.0, 12)):
        for i in range(h):
            y = np.float32(0.5) + np.float32(i) * np.float32(1.001)
            out.append(_ref_body((x, y, 14.0), (0, 0, 0, 1), (0.5, 0.5, 0.5), 5.0, 0.5, True))
    return np.array(out, BODY_DESC)


# scene_4 ramp angles and the indices of its cubes: (ramp, effective friction,
# first body id); the 2-cube tower's top cube is listed last.
REF_RAMP_DEG = (15.0, 30.0)


def reference_ramps():
    """scene_4 (src/scene_manager.hpp:355-467): floor and panel (friction 1),
    a 15 deg ramp (downhill -x) and a 30 deg ramp (downhill +x), both static
    with friction 1, and unit cubes of mass 5 with friction 0.04 / 0.25 / 1.0
    (effective sqrt(mu * 1) = 0.2 / 0.5 / 1.0) parked on them, two of each
    per ramp, plus a second cube stacked on the gentle ramp's first mu = 1
    cube.  Returns (bodies, info): info[k] = (ramp index, mu_eff) of body
    4 + k.  Known answers (Coulomb): on 15 deg (tan 0.268) mu 0.2 slides,
    0.5 and 1.0 stick and the tower stands; on 30 deg (tan 0.577) 0.2 and
    0.5 slide, 1.0 sticks; a resting cube's friction force is m g sin(theta)."""
    f32 = np.float32
    deg = f32(3.14159265) / f32(180.0)
    th = (f32(15.0) * deg, f32(30.0) * deg)
    qa = (0.0, 0.0, float(np.float32(np.sin(np.float64(th[0] * f32(0.5))))),
          float(np.float32(np.cos(np.float64(th[0] * f32(0.5))))))
    qb = (0.0, 0.0, -float(np.float32(np.sin(np.float64(th[1] * f32(0.5))))),
          float(np.float32(np.cos(np.float64(th[1] * f32(0.5))))))
    s = lambda x: f32(np.sin(np.float64(x)))  # noqa: E731
    c = lambda x: f32(np.cos(np.float64(x)))  # noqa: E731
    xl = ((c(th[0]), s(th[0])), (c(th[1]), -s(th[1])))
    yl = ((-s(th[0]), c(th[0])), (s(th[1]), c(th[1])))
    centre = ((f32(-7.0), f32(2.3), f32(14.0)), (f32(7.0), f32(3.9), f32(14.0)))
    quat = (qa, qb)
    out = _ref_floor_and_panel(1.0, (0.0, 20.0, 14.0), (5.0, 0.2, 5.0))
    for k in range(2):
        out.append(_ref_body(centre[k], quat[k], (7.0, 0.4, 4.5), 0.0, 1.0, False))
    info = []

    def cube(k, u, z, mu, lift):
        cx, cy, _ = centre[k]
        u, lift = f32(u), f32(lift)
        p = (cx + u * xl[k][0] + lift * yl[k][0], cy + u * xl[k][1] + lift * yl[k][1], f32(z))
        out.append(_ref_body(p, quat[k], (0.5, 0.5, 0.5), 5.0, mu, True))
        info.append((k, float(np.float32(np.sqrt(np.float64(mu))))))

    lift0 = f32(0.4) + f32(0.5) + f32(0.005)
    for u, z, mu in ((3.5, 16.2, 0.04), (3.5, 17.8, 0.04), (1.75, 13.4, 0.25), (1.75, 15.0, 0.25),
                     (0.0, 10.6, 1.0), (0.0, 12.2, 1.0)):
        cube(0, u, z, mu, lift0)
    cube(0, 0.0, 10.6, 1.0, lift0 + f32(1.005))
    for u, z, mu in ((-3.5, 16.2, 0.04), (-3.5, 17.8, 0.04), (-1.75, 13.4, 0.25), (-1.75, 15.0, 0.25),
                     (0.0, 10.6, 1.0), (0.0, 12.2, 1.0)):
        cube(1, u, z, mu, lift0)
    return np.array(out, BODY_DESC), info


def reference_box_pool():
    """scene_7 (src/scene_manager.hpp:744-823): floor, a panel above, four
    static walls around an 8 x 8 pit and 432 unit cubes (mass 5, friction
    0.6) raining in from y = 12..43 in 12 waves of 6 x 6, with the scene's
    index-hashed jitter and 4-13 deg tilts.  Known answer (:739-743): every
    cube ends asleep."""
    f32 = np.float32
    out = _ref_floor_and_panel(0.6, (0.0, 50.0, 14.0), (8.0, 0.2, 8.0))
    for pos, half in (((4.25, 5.0, 14.0), (0.25, 5.0, 4.5)), ((-4.25, 5.0, 14.0), (0.25, 5.0, 4.5)),
                      ((0.0, 5.0, 18.25), (4.5, 5.0, 0.25)), ((0.0, 5.0, 9.75), (4.5, 5.0, 0.25))):
        out.append(_ref_body(pos, (0, 0, 0, 1), half, 0.0, 0.6, False))
    for iy in range(12):
        for iz in range(6):
            for ix in range(6):
                h1 = ix * 3 + iy * 5 + iz * 7
                h2 = ix * 7 + iy * 3 + iz * 5
                h3 = ix * 5 + iy * 7 + iz * 3
                jx = (f32(h1 % 11) - f32(5.0)) * f32(0.05)
                jz = (f32(h2 % 11) - f32(5.0)) * f32(0.05)
                jy = (f32(h3 % 11) - f32(5.0)) * f32(0.20)
                phi = f32(h3 % 16) * f32(0.3927)
                ang = f32(0.07) + f32(h1 % 7) * f32(0.025)
                sh = f32(np.sin(np.float64(f32(0.5) * ang)))
                q = (f32(np.cos(np.float64(phi))) * sh, 0.0, f32(np.sin(np.float64(phi))) * sh,
                     f32(np.cos(np.float64(f32(0.5) * ang))))
                p = (f32(-3.0) + f32(1.2) * f32(ix) + jx, f32(12.0) + f32(2.8) * f32(iy) + jy,
                     f32(11.0) + f32(1.2) * f32(iz) + jz)
                out.append(_ref_body(p, q, (0.5, 0.5, 0.5), 5.0, 0.6, True))
    return np.array(out, BODY_DESC)
