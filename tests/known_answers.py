"""Measurements behind the reference's own known answers (test helpers).

The reference's test scenes state what must happen in them
(src/scene_manager.hpp): on the friction ramps of scene_4 every cube sticks or
slides as Coulomb's law says against tan(theta) and a resting cube's friction
force is m g sin(theta) "to 4 digits" (:355-370, the regression baseline of
src/shaders/simulation/avbd.slang:174-175); the towers of scene_3 stand
(:327-333); every cube of scene_7's pit ends asleep with no fresh contact and
no residual speed (:739-743).  Under AVBD the contact lambdas are forces in
newtons (src/shaders/simulation/avbd.slang:726-808: the dual assigns the
clamped force estimate), so the statics of a resting body can be read from
the manifold tap and set against closed forms: the normal force on a contact
carries the weight above it, m g cos(theta) on a slope, and the friction
force of a sliding contact sits on the cone, mu |F_n|.

These functions take any world with `tap(kind, dtype)` and
`read_bodies(order)`: the oracle or the HIP library.
"""
from __future__ import annotations

import numpy as np

import beatbox
import scenes
from beatbox import MANIFOLD_TAP, ORDER_BY_ROW, TAP_MANIFOLDS

G = 9.81
CUBE_MASS = 5.0
RAMP_IDS = (2, 3)  # scene_4's two static ramps, body ids
RAMP_THETA = tuple(np.radians(d) for d in scenes.REF_RAMP_DEG)
TOWER_HEIGHTS = ((2, 5), (7, 8), (15, 12))  # scene_3: (first body id, height)


def reference_desc(sleeping: bool = True) -> beatbox.WorldDesc:
    """The reference's own settings: bb_world_desc_default (AVBD, 5
    iterations, dt 0.01667, parity quirks, 30-bit keys, the default flags)."""
    d = beatbox.default_desc()
    if not sleeping:
        d.flags &= ~scenes.SIM_SLEEPING_ENABLED
    return d


def contact_forces(world) -> dict:
    """Per manifold, keyed by the sorted pair of body ids: (|F_n|, F_t
    vector, every contact sticking, the manifold normal) with
    F_n = -sum(lambda_n) and F_t = sum(lambda_t0 t0 + lambda_t1 t1) in
    float64."""
    rows = world.read_bodies(ORDER_BY_ROW)
    out = {}
    for m in world.tap(TAP_MANIFOLDS, MANIFOLD_TAP):
        a, b = int(rows["id"][m["obb1_index"]]), int(rows["id"][m["obb2_index"]])
        cc = int(m["contact_count"])
        c = m["contacts"][:cc]
        lam_t = c["tangent_impulse"].astype(np.float64)
        t0 = m["tangent"][0].astype(np.float64)
        t1 = m["tangent"][1].astype(np.float64)
        ft = (lam_t[:, :1] * t0 + lam_t[:, 1:2] * t1).sum(axis=0)
        fn = -float(c["normal_impulse"].astype(np.float64).sum())
        out[(min(a, b), max(a, b))] = (fn, ft, bool(np.all(c["stick"] == 1)), m["normal"].astype(np.float64))
    return out


def ramp_cases():
    """scene_4 cubes as (body id, ramp index, mu_eff, load in cubes, sticks
    by Coulomb): body 4 + k for info[k]; body 10 stands on body 8, which
    therefore carries two cubes."""
    _, info = scenes.reference_ramps()
    cases = []
    for k, (ramp, mu) in enumerate(info):
        body = 4 + k
        if body == 10:
            continue  # the tower's top cube rests on body 8, not on the ramp
        load = 2 if body == 8 else 1
        cases.append((body, ramp, mu, load, mu > np.tan(RAMP_THETA[ramp])))
    return cases


def slope_speed(world, body: int, ramp: int) -> float:
    """Speed of a body along its ramp's downhill direction."""
    v = world.read_bodies()["velocity"][body].astype(np.float64)
    th = RAMP_THETA[ramp]
    downhill = np.array([-np.cos(th), -np.sin(th), 0.0]) if ramp == 0 else np.array([np.cos(th), -np.sin(th), 0.0])
    return float(v @ downhill)


def spinning_stack(w_bottom: float, w_top: float):
    """Two frictionless unit cubes stacked on a frictionless floor under AVBD
    (which neither damps nor assists a body with fewer than three contacts),
    each spinning about the vertical axis at the given rate.  Frictionless
    contacts apply no torque about the vertical, so each cube keeps its rate;
    one above sqrt(0.0225) = 0.15 rad/s keeps that cube's sleep timer from
    counting (RB_sim.slang:556-737).  Returns (desc, bodies): rows 1 and 2 are
    the bottom and top cubes."""
    desc = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=8))
    b = np.zeros(3, scenes.BODY_DESC)
    scenes._ground(b[0], friction=0.0)
    for k, (y, w) in ((1, (0.5, w_bottom)), (2, (1.5, w_top))):
        b[k]["shape"] = scenes.SHAPE_BOX
        b[k]["flags"] = scenes.BODY_DYNAMIC | scenes.BODY_GRAVITY
        b[k]["position"] = (0.0, y, 0.0)
        b[k]["rotation"] = (0.0, 0.0, 0.0, 1.0)
        b[k]["minimum"] = (-0.5, -0.5, -0.5)
        b[k]["maximum"] = (0.5, 0.5, 0.5)
        b[k]["mass"] = 1.0
        b[k]["friction"] = 0.0
        b[k]["omega"] = (0.0, w, 0.0)
    return desc, b
