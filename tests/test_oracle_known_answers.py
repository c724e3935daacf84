"""The oracle against the reference's own known answers and documented rules.

Each test sets up a scene whose outcome the reference states in its own
words, or whose numbers follow from a rule it documents, and checks the CPU
oracle against that outcome computed here from physics or from the rule, not
from the oracle's code:

- scene_4 friction ramps (src/scene_manager.hpp:355-467, the AVBD regression
  baseline of src/shaders/simulation/avbd.slang:174-175): stick or slide per
  Coulomb against tan(theta); contact forces equal the statics (normal m g
  cos(theta), resting friction m g sin(theta) to 4 digits, sliding friction
  on the cone mu |F_n|); sliders accelerate at g (sin(theta) - mu cos(theta)).
- scene_3 towers (src/scene_manager.hpp:327-353) stand, and each contact of a
  stack carries the weight above it.
- scene_7 box pool (src/scene_manager.hpp:739-823): every cube asleep, no
  fresh contact, no residual speed.
- Sleeping rules (src/shaders/simulation/RB_sim.slang:556-737): quiet timer
  of 30 steps under (0.12 m/s)^2 and (0.15 rad/s)^2, velocity zeroed on
  falling asleep, a contactless sleeper bled by 10 a step (so it wakes the
  next step and sleeps again 10 steps later); velocity damping exp(-0.1 dt)
  per step on the PGS path (src/shaders/extensions.slang:127-139), applied
  before the sleep test (task order, src/rigid_body_manager.cpp:1303-1314),
  none under AVBD.
- Warm-start gates (src/shaders/extensions.slang:273-473): a contact
  matched by feature ids within 0.1 m keeps its old position until it has
  drifted 0.1 m; beyond 0.1 m it is inherited by proximity within 0.35 m;
  beyond that it starts fresh.
"""
from __future__ import annotations

import numpy as np

import beatbox
import scenes
from known_answers import (CUBE_MASS, G, RAMP_IDS, RAMP_THETA, TOWER_HEIGHTS, contact_forces, ramp_cases,
                           spinning_stack,
                           reference_desc, slope_speed)
from oracle_bind import OracleWorld
from beatbox import MANIFOLD_TAP, TAP_MANIFOLDS

MG = CUBE_MASS * G


def _world(desc, bodies) -> OracleWorld:
    o = OracleWorld(desc)
    o.add_bodies(bodies)
    return o


def test_ramps_stick_or_slide_per_coulomb():
    """With the reference's settings (sleeping on), after 400 steps: mu_eff
    0.2 slides on both ramps, 0.5 slides only on the 30 deg ramp, 1.0 sticks
    on both; the stuck cubes and the 2-cube tower have not moved and sleep."""
    bodies, _ = scenes.reference_ramps()
    o = _world(reference_desc(), bodies)
    p0 = o.read_bodies()["position"].astype(np.float64)
    o.step(400)
    b = o.read_bodies()
    disp = np.linalg.norm(b["position"].astype(np.float64) - p0, axis=1)
    for body, ramp, mu, _load, sticks in ramp_cases():
        if sticks:
            assert disp[body] < 0.02, (body, ramp, mu, disp[body])
            assert b["flags"][body] & 8, f"stuck cube {body} is not asleep"
        else:
            assert disp[body] > 1.0, (body, ramp, mu, disp[body])
    assert disp[10] < 0.02 and b["position"][10][1] > b["position"][8][1] + 0.9  # the tower stands


def test_ramp_contact_forces_are_the_statics():
    """Sleeping off (the forces of a sleeper are not re-solved), step 100:
    every cube's ramp contact carries F_n = L m g cos(theta) for a load of L
    cubes; a sticking cube's friction is L m g sin(theta) and its contacts
    stick; a sliding cube's friction is on the cone, mu_eff F_n, and its
    contacts do not stick.  4 digits for single cubes; the tower's base
    cube, which also carries the top cube's moment, to 2e-3."""
    bodies, _ = scenes.reference_ramps()
    o = _world(reference_desc(sleeping=False), bodies)
    o.step(100)
    f = contact_forces(o)
    for body, ramp, mu, load, sticks in ramp_cases():
        fn, ft, stick, _n = f[(RAMP_IDS[ramp], body)]
        th = RAMP_THETA[ramp]
        tol = 5e-4 if load == 1 else 2e-3
        assert abs(fn / (load * MG * np.cos(th)) - 1) < tol, (body, fn)
        if sticks:
            assert abs(np.linalg.norm(ft) / (load * MG * np.sin(th)) - 1) < tol, (body, ft)
            assert stick, body
        else:
            assert abs(np.linalg.norm(ft) / (mu * fn) - 1) < 1e-3, (body, ft, fn)
            assert not stick, body
    # the tower's top cube rests on its base cube: the same statics
    fn, ft, stick, _n = f[(8, 10)]
    assert abs(fn / (MG * np.cos(RAMP_THETA[0])) - 1) < 2e-3
    assert abs(np.linalg.norm(ft) / (MG * np.sin(RAMP_THETA[0])) - 1) < 2e-3 and stick


def test_ramp_sliders_accelerate_at_coulomb_rate():
    """Sliding cubes gain speed along the slope at g (sin(theta) - mu_eff
    cos(theta)), measured between steps 40 and 100, to 1%."""
    bodies, _ = scenes.reference_ramps()
    desc = reference_desc(sleeping=False)
    o = _world(desc, bodies)
    o.step(40)
    sliders = [(b, r, mu) for b, r, mu, _l, s in ramp_cases() if not s]
    v0 = {b: slope_speed(o, b, r) for b, r, _ in sliders}
    o.step(60)
    t = 60 * float(np.float32(desc.dt))
    for b, r, mu in sliders:
        a = (slope_speed(o, b, r) - v0[b]) / t
        th = RAMP_THETA[r]
        expect = G * (np.sin(th) - mu * np.cos(th))
        assert abs(a / expect - 1) < 0.01, (b, r, mu, a, expect)


def test_towers_stand_and_each_contact_carries_the_weight_above():
    """Reference settings, 300 steps: the 5-, 8- and 12-cube towers stand
    (no cube moved sideways by 1 mm or tilted by 0.1 deg) and every cube
    sleeps.  Sleeping off, step 120: the contact under the cube at level i
    of an h-cube tower carries (h - i) m g, to 1e-3, with no friction."""
    bodies = scenes.reference_towers()
    o = _world(reference_desc(), bodies)
    p0 = o.read_bodies()["position"].astype(np.float64)
    o.step(300)
    b = o.read_bodies()
    dyn = slice(2, len(bodies))
    d = b["position"][dyn].astype(np.float64) - p0[dyn]
    assert np.abs(d[:, [0, 2]]).max() < 1e-3
    q = b["rotation"][dyn].astype(np.float64)
    tilt = 2 * np.degrees(np.arccos(np.clip(np.abs(q[:, 3]) / np.linalg.norm(q, axis=1), 0, 1)))
    assert tilt.max() < 0.1
    assert o.stats()["sleeping_count"] == len(bodies) - 2

    o = _world(reference_desc(sleeping=False), bodies)
    o.step(120)
    f = contact_forces(o)
    seen = 0
    for first, h in TOWER_HEIGHTS:
        for level in range(h):
            below = 0 if level == 0 else first + level - 1  # the floor is body 0
            fn, ft, _stick, _n = f[(below, first + level)]
            assert abs(fn / ((h - level) * MG) - 1) < 1e-3, (first, level, fn)
            assert np.linalg.norm(ft) < 1e-3 * MG
            seen += 1
    assert seen == 25


def test_box_pool_ends_fully_asleep():
    """scene_7's success criterion (src/scene_manager.hpp:739-743): 432
    cubes rained into the pit all sleep, no contact is fresh and no body
    moves (sleeping = 432, fresh = 0, maxv = 0).  The oracle gets there
    after 4,196 steps (70 s of simulated time) and stays."""
    o = _world(reference_desc(), scenes.reference_box_pool())
    o.step(4200)
    st = o.stats()
    assert (st["sleeping_count"], st["dbg_fresh"], st["dbg_maxv"]) == (432, 0, 0), st


def _lone_body(solver: int, v=(0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0)) -> OracleWorld:
    desc = beatbox.WorldDesc.make(**scenes.world_desc(solver=solver, iterations=8))
    b = scenes.random_boxes(1, 1, with_ground=False)
    b["flags"] = scenes.BODY_DYNAMIC  # no gravity: nothing but damping acts
    b["position"] = (0.0, 5.0, 0.0)
    b["rotation"] = (0.0, 0.0, 0.0, 1.0)
    b["velocity"] = v
    b["omega"] = w
    return _world(desc, b)


def _asleep_steps(o: OracleWorld, steps: int) -> list:
    out = []
    for s in range(1, steps + 1):
        o.step(1)
        if o.read_bodies()["flags"][0] & 8:
            out.append(s)
    return out


def _first_calm_step(speed: float, limit: float, damped: bool) -> int:
    """First step whose speed, damped by exp(-0.1 dt) per step before the
    test (float32, as the reference multiplies), squares below limit^2."""
    dt = np.float32(scenes.TIME_STEP)
    damp = np.float32(np.exp(-np.float64(dt) * 0.1)) if damped else np.float32(1.0)
    v = np.float32(speed)
    for s in range(1, 10_000):
        v = np.float32(v * damp)
        if np.float32(v * v) < np.float32(limit * limit):
            return s
    raise AssertionError("never calm")


def test_sleep_timer_rules_on_a_lone_body():
    """A body with no contact and no gravity: it sleeps after 30 calm steps,
    the next step the support-loss bleed (-10) wakes it, and 10 calm steps
    later it sleeps again: asleep at c + 29, c + 40, c + 51, ... for the
    first calm step c.  Under PGS_SOFT the speed is damped every step, so a
    body just above the threshold becomes calm after a computable number of
    steps; AVBD does not damp, so it never does."""
    for solver, damped in ((scenes.SOLVER_PGS_SOFT, True), (scenes.SOLVER_AVBD, False)):
        for v in (0.05, 0.119, 0.125, 0.2):
            c = _first_calm_step(v, 0.12, damped) if (damped or v < 0.12) else None
            got = _asleep_steps(_lone_body(solver, v=(v, 0.0, 0.0)), 90)
            want = [s for s in range(c + 29, 91, 11)] if c is not None else []
            assert got == want, (solver, v, got, want)
        for w in (0.149, 0.155):
            c = _first_calm_step(w, 0.15, damped) if (damped or w < 0.15) else None
            got = _asleep_steps(_lone_body(solver, w=(0.0, w, 0.0)), 90)
            want = [s for s in range(c + 29, 91, 11)] if c is not None else []
            assert got == want, (solver, w, got, want)


def test_sleep_veto_by_a_loud_neighbour():
    """The neighbourhood veto (RB_sim.slang:556-737): a dynamic body in
    contact with a dynamic neighbour whose sleep timer has not reached 30 may
    not sleep, whatever its own timer; a static neighbour never vetoes.  Two
    cubes at rest on the floor are calm from step 1, so both sleep at step 30
    (= 1 + 29).  With either cube spinning at 0.2 rad/s (never calm, never
    asleep), the other stays calm but never sleeps: the veto works in both
    directions of a contact."""
    for w_bottom, w_top in ((0.0, 0.0), (0.0, 0.2), (0.2, 0.0)):
        desc, bodies = spinning_stack(w_bottom, w_top)
        o = _world(desc, bodies)
        asleep = {1: [], 2: []}
        for s in range(1, 121):
            o.step(1)
            r = o.read_bodies()
            for k in (1, 2):
                if r["flags"][k] & 8:
                    asleep[k].append(s)
                spin = (w_bottom, w_top)[k - 1]
                v2 = float(np.dot(r["velocity"][k], r["velocity"][k]))
                w2 = float(np.dot(r["omega"][k], r["omega"][k]))
                if spin == 0.0:
                    assert v2 < 0.0144 and w2 < 0.0225, (w_bottom, w_top, s, k)  # calm every step
                else:
                    assert abs(np.sqrt(w2) - spin) < 1e-3, (w_bottom, w_top, s, k)  # keeps its rate
        if w_bottom == w_top == 0.0:
            assert asleep == {1: list(range(30, 121)), 2: list(range(30, 121))}, asleep
        else:
            assert asleep == {1: [], 2: []}, (w_bottom, w_top, asleep)


def _slide(speed: float, steps: int):
    """A unit box resting on the floor, both frictionless, pushed along x:
    per step the box position (start of step), fresh count, and the stored
    contact x positions and warm-start counters of its manifold."""
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    desc.flags &= ~scenes.SIM_SLEEPING_ENABLED
    b = scenes.random_boxes(1, 1)
    b["friction"] = 0.0
    b["position"][1] = (-20.0, 0.5, 0.0)
    b["rotation"][1] = (0.0, 0.0, 0.0, 1.0)
    b["velocity"][1] = (speed, 0.0, 0.0)
    b["omega"][1] = 0.0
    o = _world(desc, b)
    x_prev = -20.0
    out = []
    for _ in range(steps):
        o.step(1)
        m = o.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
        assert len(m) == 1 and m["contact_count"][0] == 4
        c = m["contacts"][0][:4]
        out.append((x_prev, o.stats()["dbg_fresh"], np.sort(c["position"][:, 0].astype(np.float64)),
                    c["warm_start"].copy()))
        x_prev = float(o.read_bodies()["position"][1][0])
    return out


def test_warm_start_gates_on_a_sliding_box():
    """The bottom face's four contacts move with the box (x -/+ 0.5 from
    its centre at the start of the step).  At 1.8 m/s (3 cm a step) they
    match by feature ids and the stored positions stay put until the drift
    from the stored point reaches 0.1 m, then jump to the new corners.  At
    9 and 19.5 m/s (15 and 32.5 cm) the id match fails the 0.1 m gate and
    proximity within 0.35 m carries them, with the new positions.  At 22.5
    m/s (37.5 cm) nothing carries them: all four are fresh every step."""
    rows = _slide(1.8, 30)
    stored = rows[0][0]
    jumps = 0
    for s, (x, fresh, cx, ws) in enumerate(rows):
        if s > 0:
            assert fresh == 0 and np.all(ws == s)
            if abs(x - stored) >= 0.1 + 1e-4:
                stored, jumps = x, jumps + 1
            else:
                assert abs(x - stored) < 0.1 - 1e-4, "step too close to the gate to decide"
        assert np.allclose(cx, [stored - 0.5, stored - 0.5, stored + 0.5, stored + 0.5], atol=1e-5), (s, cx, stored)
    assert jumps >= 6
    for speed in (9.0, 19.5):
        for s, (x, fresh, cx, ws) in enumerate(_slide(speed, 20)):
            assert np.allclose(cx, [x - 0.5, x - 0.5, x + 0.5, x + 0.5], atol=1e-5)
            if s > 0:
                assert fresh == 0 and np.all(ws == s), (speed, s, fresh, ws)
    for s, (x, fresh, cx, ws) in enumerate(_slide(22.5, 20)):
        assert fresh == 4 and np.all(ws == 0), (s, fresh, ws)


# ------------------------------------------------------------ SAT rules
# Box-box SAT (src/shaders/simulation/collision_detection.slang:575-760): face
# axes of A and B and, unless two box axes are within 0.9998 of parallel,
# edge axes; the face of B wins over the face of A only when 0.95 s_B >
# s_A + 0.01 (separations s <= 0), which keeps the reference face stable from
# frame to frame.  Clipping (:336-420) keeps an incident vertex that lies
# within 0.005 of a side plane of the reference face as it is.
AXIS_A_Y, AXIS_B_Y = 1, 4


def _box(pos, quat=(0.0, 0.0, 0.0, 1.0), half=(0.5, 0.5, 0.5)):
    b = np.zeros(1, scenes.BODY_DESC)
    b["shape"] = scenes.SHAPE_BOX
    b["flags"] = scenes.BODY_DYNAMIC | scenes.BODY_GRAVITY
    b["position"] = pos
    b["rotation"] = quat
    b["minimum"] = -np.asarray(half, np.float32)
    b["maximum"] = np.asarray(half, np.float32)
    b["mass"] = 1.0
    b["friction"] = 0.5
    return b


def test_sat_face_bias_threshold():
    """A: a unit cube tilted by phi about z; B: a 2 x 1 x 2 slab centred
    above it at height y.  Only the two y faces are close (the z axes are
    parallel, so no edge axis is tried):
      s_A = y cos(phi) - (0.5 + sin(phi) + 0.5 cos(phi))
      s_B = y - (0.5 + 0.5 sin(phi) + 0.5 cos(phi)).
    The key is B's y face exactly when 0.95 s_B > s_A + 0.01; sweeping y
    crosses that line for each phi."""
    from oracle_bind import collide
    desc = beatbox.WorldDesc.make(**scenes.world_desc())
    flips = 0
    for phi in (0.004, 0.01, 0.015):
        q = (0.0, 0.0, float(np.sin(phi / 2)), float(np.cos(phi / 2)))
        a = _box((0.0, 0.0, 0.0), q)
        phi32 = 2.0 * np.arctan2(np.float64(np.float32(q[2])), np.float64(np.float32(q[3])))
        s, c = np.sin(phi32), np.cos(phi32)
        last = None
        for y in np.linspace(0.80, 1.0, 81):
            y = float(np.float32(y))
            s_a = y * c - (0.5 + s + 0.5 * c)
            s_b = y - (0.5 + 0.5 * s + 0.5 * c)
            margin = 0.95 * s_b - (s_a + 0.01)
            if abs(margin) < 1e-5 or s_b > -1e-4:
                continue
            m = collide(desc, a, _box((0.0, y, 0.0), half=(1.0, 0.5, 1.0)))
            assert m is not None and m["contact_count"] > 0, (phi, y)
            want = AXIS_B_Y if margin > 0 else AXIS_A_Y
            assert m["key"] == want, (phi, y, s_a, s_b, m["key"])
            flips += last is not None and want != last
            last = want
    assert flips == 3  # the threshold is crossed once per tilt


def test_clip_keeps_vertices_within_five_mm_of_a_side_plane():
    """A unit cube (incident, B) 1 cm into a unit cube below it (reference,
    A: equal separations leave A's face), shifted along x by dx: its bottom
    vertices at x = 0.5 + dx stick out past A's side plane x = 0.5.  Within
    0.005 they are kept where they are; farther out they are clipped to the
    plane."""
    from oracle_bind import collide
    desc = beatbox.WorldDesc.make(**scenes.world_desc())
    base = _box((0.0, 0.0, 0.0))
    for dx, x_max in ((0.002, 0.502), (0.004, 0.504), (0.006, 0.5), (0.02, 0.5)):
        m = collide(desc, base, _box((dx, 0.99, 0.0)))
        assert m is not None and m["key"] == AXIS_A_Y and m["contact_count"] == 4, dx
        px = m["contacts"]["position"][:4, 0]
        assert abs(float(px.max()) - x_max) < 1e-6, (dx, px)
        assert abs(float(px.min()) - (dx - 0.5)) < 1e-6, (dx, px)
