"""The reference's own known answers on the GPU.

The HIP step runs the reference's test scenes with the reference's settings
and must (1) equal the oracle bit for bit at the checked steps, every tap
included, and (2) show the outcomes the reference states for those scenes,
measured on the GPU's own results: the scene_4 ramps stick or slide per
Coulomb and their contact forces are the statics (src/scene_manager.hpp:
355-370), the scene_3 towers stand and their stacks carry their weight
(:327-353), and every cube of the scene_7 pit ends asleep (:739-743).
"""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes
from known_answers import (CUBE_MASS, G, RAMP_IDS, RAMP_THETA, TOWER_HEIGHTS, contact_forces, ramp_cases,
                           reference_desc, spinning_stack)
from oracle_bind import OracleWorld
from parity_util import compare_step

pytestmark = pytest.mark.gpu

MG = CUBE_MASS * G


def _pair(desc, bodies):
    gpu = beatbox.RigidBodyManager(desc)
    gpu.add_bodies(bodies)
    ora = OracleWorld(desc)
    ora.add_bodies(bodies)
    return gpu, ora


def _advance(gpu, ora, steps, check_every=0, bits=30):
    """Steps both; compares every tap and all state at the last step (and
    every check_every steps)."""
    done = 0
    while done < steps:
        k = min(check_every or steps, steps - done)
        gpu.simulate(k)
        ora.step(k)
        done += k
        bad = compare_step(gpu, ora, bits, done)
        assert not bad, "\n".join(bad[:20])


def test_ramps_on_gpu():
    bodies, _ = scenes.reference_ramps()
    print("ramps: reference settings, 400 steps", flush=True)
    gpu, ora = _pair(reference_desc(), bodies)
    p0 = gpu.read_bodies()["position"].astype(np.float64)
    _advance(gpu, ora, 400, check_every=50)
    b = gpu.read_bodies()
    disp = np.linalg.norm(b["position"].astype(np.float64) - p0, axis=1)
    for body, _ramp, _mu, _load, sticks in ramp_cases():
        assert (disp[body] < 0.02 and b["flags"][body] & 8) if sticks else disp[body] > 1.0, body
    assert disp[10] < 0.02

    print("ramps: sleeping off, 100 steps", flush=True)
    gpu, ora = _pair(reference_desc(sleeping=False), bodies)
    _advance(gpu, ora, 100, check_every=10)
    f = contact_forces(gpu)
    for body, ramp, mu, load, sticks in ramp_cases():
        fn, ft, stick, _n = f[(RAMP_IDS[ramp], body)]
        th = RAMP_THETA[ramp]
        tol = 5e-4 if load == 1 else 2e-3
        assert abs(fn / (load * MG * np.cos(th)) - 1) < tol, (body, fn)
        if sticks:
            assert abs(np.linalg.norm(ft) / (load * MG * np.sin(th)) - 1) < tol and stick, body
        else:
            assert abs(np.linalg.norm(ft) / (mu * fn) - 1) < 1e-3 and not stick, body


def test_towers_on_gpu():
    bodies = scenes.reference_towers()
    print("towers: reference settings, 300 steps", flush=True)
    gpu, ora = _pair(reference_desc(), bodies)
    p0 = gpu.read_bodies()["position"].astype(np.float64)
    _advance(gpu, ora, 300, check_every=100)
    d = gpu.read_bodies()["position"][2:].astype(np.float64) - p0[2:]
    assert np.abs(d[:, [0, 2]]).max() < 1e-3
    assert gpu.read_back_sim_config()["sleeping_count"] == len(bodies) - 2

    print("towers: sleeping off, 120 steps", flush=True)
    gpu, ora = _pair(reference_desc(sleeping=False), bodies)
    _advance(gpu, ora, 120, check_every=40)
    f = contact_forces(gpu)
    for first, h in TOWER_HEIGHTS:
        for level in range(h):
            below = 0 if level == 0 else first + level - 1
            fn, _ft, _stick, _n = f[(below, first + level)]
            assert abs(fn / ((h - level) * MG) - 1) < 1e-3, (first, level, fn)


def test_box_pool_on_gpu():
    """4,200 steps of the 432-cube pit: equal to the oracle at every
    1,050th step and at the end, and every cube asleep with no fresh
    contact and no residual speed."""
    print("box pool: reference settings, 4200 steps", flush=True)
    gpu, ora = _pair(reference_desc(), scenes.reference_box_pool())
    _advance(gpu, ora, 4200, check_every=1050)
    st = gpu.read_back_sim_config()
    assert (st["sleeping_count"], st["dbg_fresh"], st["dbg_maxv"]) == (432, 0, 0), st


def test_sleep_and_warm_start_rules_on_gpu():
    """The lone-body sleep cycle and the sliding-box warm-start cases of the
    oracle's rule tests, compared with the oracle at every step."""
    cases = []
    for solver in (scenes.SOLVER_PGS_SOFT, scenes.SOLVER_AVBD):
        for v in (0.119, 0.125):
            b = scenes.random_boxes(1, 1, with_ground=False)
            b["flags"] = scenes.BODY_DYNAMIC
            b["position"] = (0.0, 5.0, 0.0)
            b["rotation"] = (0.0, 0.0, 0.0, 1.0)
            b["velocity"] = (v, 0.0, 0.0)
            b["omega"] = 0.0
            cases.append((scenes.world_desc(solver=solver, iterations=8), b, 70))
    for speed in (1.8, 9.0, 22.5):
        b = scenes.random_boxes(1, 1)
        b["friction"] = 0.0
        b["position"][1] = (-20.0, 0.5, 0.0)
        b["rotation"][1] = (0.0, 0.0, 0.0, 1.0)
        b["velocity"][1] = (speed, 0.0, 0.0)
        b["omega"][1] = 0.0
        kw = scenes.world_desc(iterations=8)
        kw["flags"] &= ~scenes.SIM_SLEEPING_ENABLED
        cases.append((kw, b, 20))
    for kw, bodies, steps in cases:
        print(f"rule case: solver {kw['solver']} v {bodies['velocity'][-1][0]}", flush=True)
        gpu, ora = _pair(beatbox.WorldDesc.make(**kw), bodies)
        _advance(gpu, ora, steps, check_every=1)


def test_sleep_veto_by_a_loud_neighbour_on_gpu():
    """The oracle's neighbourhood-veto rule case on the GPU, equal to the
    oracle every 10 steps: the resting stack sleeps at step 30, and a cube
    spinning at 0.2 rad/s (above or below) keeps the other awake."""
    for w_bottom, w_top in ((0.0, 0.0), (0.0, 0.2), (0.2, 0.0)):
        print(f"veto case: spin {w_bottom} below, {w_top} above", flush=True)
        gpu, ora = _pair(*spinning_stack(w_bottom, w_top))
        _advance(gpu, ora, 29, check_every=29)
        assert not np.any(gpu.read_bodies()["flags"][1:] & 8)
        _advance(gpu, ora, 91, check_every=10)
        asleep = gpu.read_bodies()["flags"][1:] & 8
        assert np.all(asleep) if w_bottom == w_top == 0.0 else not np.any(asleep), (w_bottom, w_top)
