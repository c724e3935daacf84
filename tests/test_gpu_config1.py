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

pytestmark = pytest.mark.gpu

POS_TOL_M = 1e-4
ROT_TOL_RAD = 1e-3


def pose_error(g: np.ndarray, o: np.ndarray) -> tuple[float, float]:
    """Largest position distance (m) and rotation angle (rad) between two body
    readbacks in id order."""
    dx = np.linalg.norm(g["position"].astype(np.float64) - o["position"].astype(np.float64), axis=1)
    dot = np.abs(np.sum(g["rotation"].astype(np.float64) * o["rotation"].astype(np.float64), axis=1))
    ang = 2.0 * np.arccos(np.clip(dot, 0.0, 1.0))
    return float(dx.max()), float(ang.max())


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
    dx, ang = res["err60"]
    assert dx <= POS_TOL_M and ang <= ROT_TOL_RAD, f"step 60 pose error {dx} m, {ang} rad"
    assert not res["first_bad"], "\n".join(res["first_bad"][:20])
    assert res["max_manifolds"] > 1000, "the compared window must reach the contact phase"


def test_config1_pgs_soft_200_steps():
    res = run_config1(scenes.SOLVER_PGS_SOFT, 200)
    check(res)


def test_config1_avbd_200_steps():
    res = run_config1(scenes.SOLVER_AVBD, 200)
    check(res)
    assert res["final"]["sleeping_count"] > 250, "AVBD run must reach the settling phase"


def test_config1_pgs_soft_dt_one_sixtieth():
    """The config's dt = 1/60 (the reference constant is 0.01667f)."""
    res = run_config1(scenes.SOLVER_PGS_SOFT, 200, dt=float(np.float32(1.0 / 60.0)))
    check(res)
