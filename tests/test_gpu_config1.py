"""The 1,000-cube drop (the reference's scene_2, src/scene_manager.hpp:193-268)
stepped for its whole 200-step horizon on the GPU and on the oracle.

This is the one configuration the reference itself can run, stepped the way
its fixed-step loop steps it (src/renderer_manager.cpp:248-262).  Every step's
taps (keys, order, BVH, pairs, islands, colours, manifolds with impulses),
body states and counters are compared bit for bit.  On top of that, the pose
tolerance of the north-star definition is asserted on its own at step 60:
|dx| <= 1e-4 m and rotation angle <= 1e-3 rad.  The compared window must
cover the contact phase (more than 1,000 manifolds) and, under AVBD, the
settling phase (more than 250 bodies asleep).
"""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes
from oracle_bind import OracleWorld
from parity_util import compare_step


POS_TOL_M = 1e-4
ROT_TOL_RAD = 1e-3


def pose_error(g: np.ndarray, o: np.ndarray) -> tuple[float, float]:
    """Largest position distance (m) and rotation angle (rad) between two body
    readbacks in id order.  The angle is that of the relative rotation
    conj(qa) * qb of the normalised quaternions, 2 atan2(|r.v|, |r.w|): it is
    0 for identical quaternions whatever their float32 norm (an arccos of the
    dot product is not: a norm 3e-7 below 1 alone reads as 1.5e-3 rad)."""
    dx = np.linalg.norm(g["position"].astype(np.float64) - o["position"].astype(np.float64), axis=1)
    qa = g["rotation"].astype(np.float64)
    qb = o["rotation"].astype(np.float64)
    qa /= np.linalg.norm(qa, axis=1, keepdims=True)
    qb /= np.linalg.norm(qb, axis=1, keepdims=True)
    va, wa, vb, wb = qa[:, :3], qa[:, 3:], qb[:, :3], qb[:, 3:]
    rv = wa * vb - wb * va - np.cross(va, vb)
    rw = wa[:, 0] * wb[:, 0] + np.sum(va * vb, axis=1)
    ang = 2.0 * np.arctan2(np.linalg.norm(rv, axis=1), np.abs(rw))
    return float(dx.max()), float(ang.max())


def test_pose_error_measure():
    """The tolerance measure itself: zero for identical poses, the rotation
    angle for a known relative rotation, the distance for a known shift."""
    rng = np.random.default_rng(5)
    q = rng.normal(size=(50, 4)).astype(np.float32)
    q *= np.float32(1.0 - 3e-7) / np.linalg.norm(q, axis=1, keepdims=True).astype(np.float32)
    a = np.zeros(50, beatbox.BODY_STATE)
    a["rotation"] = q
    a["position"] = rng.normal(size=(50, 3)).astype(np.float32)
    assert pose_error(a, a) == (0.0, 0.0)
    b = a.copy()
    th = 1e-3  # about +z: q * (0, 0, sin(th/2), cos(th/2))
    s, c = np.sin(th / 2), np.cos(th / 2)
    qa = q.astype(np.float64)
    qb = np.stack([qa[:, 3] * 0 + qa[:, 0] * c + qa[:, 1] * s,
                   qa[:, 1] * c - qa[:, 0] * s,
                   qa[:, 2] * c + qa[:, 3] * s,
                   qa[:, 3] * c - qa[:, 2] * s], axis=1)
    b["rotation"] = qb.astype(np.float32)
    b["position"][7, 0] += np.float32(2e-4)
    dx, ang = pose_error(a, b)
    assert abs(dx - 2e-4) < 1e-6
    assert abs(ang - th) < 2e-6


def run_config1(solver: int, steps: int, dt: float | None = None):
    kw = scenes.world_desc(solver=solver, iterations=8)
    if dt is not None:
        kw["dt"] = dt
    desc = beatbox.WorldDesc.make(**kw)
    bodies = scenes.cubes_grid()
    gpu = beatbox.RigidBodyManager(desc)
    gpu.add_bodies(bodies)
    ora = OracleWorld(desc)
    ora.add_bodies(bodies)
    e0 = gpu.energy()
    max_manifolds = 0
    max_pen_mm = 0
    first_bad: list = []
    err60 = None
    for s in range(steps):
        gpu.simulate(1)
        ora.step(1)
        bad = compare_step(gpu, ora, desc.morton_bits, s, check_taps=True)
        if bad and not first_bad:
            first_bad = bad
        st = gpu.read_back_sim_config()
        max_manifolds = max(max_manifolds, st["collision_count"])
        max_pen_mm = max(max_pen_mm, st["dbg_pen"])
        if s + 1 == 60:
            err60 = pose_error(gpu.read_bodies(), ora.read_bodies())
    final = gpu.read_back_sim_config()
    e1 = gpu.energy()
    print(f"config-1 solver {solver}: {steps} steps, max manifolds {max_manifolds}, "
          f"sleeping {final['sleeping_count']}, max penetration {max_pen_mm} mm, "
          f"energy {e0:.3f} -> {e1:.3f} J, step-60 pose error {err60}", flush=True)
    return dict(first_bad=first_bad, err60=err60, max_manifolds=max_manifolds,
                final=final, e0=e0, e1=e1)


def check(res):
    assert not res["first_bad"], "\n".join(res["first_bad"][:20])
    dx, ang = res["err60"]
    assert dx <= POS_TOL_M and ang <= ROT_TOL_RAD, f"step 60 pose error {dx} m, {ang} rad"
    assert res["max_manifolds"] > 1000, "the compared window must reach the contact phase"


@pytest.mark.gpu
def test_config1_pgs_soft_200_steps():
    res = run_config1(scenes.SOLVER_PGS_SOFT, 200)
    check(res)


@pytest.mark.gpu
def test_config1_avbd_200_steps():
    res = run_config1(scenes.SOLVER_AVBD, 200)
    check(res)
    assert res["final"]["sleeping_count"] > 250, "AVBD run must reach the settling phase"


@pytest.mark.gpu
def test_config1_pgs_soft_dt_one_sixtieth():
    """The config's dt = 1/60 (the reference constant is 0.01667f)."""
    res = run_config1(scenes.SOLVER_PGS_SOFT, 200, dt=float(np.float32(1.0 / 60.0)))
    check(res)
