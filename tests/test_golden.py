"""The committed fixtures of tests/golden/ reproduced by the CPU oracle (CPU
suite) and by the HIP step through the C-ABI (GPU suite), bit-exactly."""
from __future__ import annotations

import os

import numpy as np
import pytest

import beatbox
import scenes
from beatbox import (LBVH_NODE, MANIFOLD_TAP, TAP_EDGE_COLORS, TAP_ISLANDS, TAP_LBVH,
                     TAP_MANIFOLDS, TAP_MORTON, TAP_PAIRS, TAP_SORTED, WorldDesc)
from oracle_bind import OracleWorld, collide

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _same(a, b):
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def _check_taps(w, f, prefix):
    assert _same(w.tap(TAP_MORTON, np.uint32), f[prefix + "morton"])
    assert _same(w.tap(TAP_SORTED, np.uint32), f[prefix + "sorted"])
    assert _same(w.tap(TAP_LBVH, LBVH_NODE), f[prefix + "lbvh"])
    assert _same(w.tap(TAP_PAIRS, np.uint32), f[prefix + "pairs"])
    assert _same(w.tap(TAP_ISLANDS, np.uint32), f[prefix + "islands"])
    assert _same(w.tap(TAP_EDGE_COLORS, np.uint32), f[prefix + "colors"])
    assert _same(w.tap(TAP_MANIFOLDS, MANIFOLD_TAP), f[prefix + "manifolds"])
    got = w.read_bodies()
    want = f[prefix + "bodies"]
    for k in ("id", "flags", "position", "rotation", "velocity", "omega"):
        assert _same(got[k], want[k]), k


def _pile(world_cls):
    f = load("pile64.npz")
    w = world_cls(WorldDesc.make(**scenes.world_desc(iterations=8)))
    w.add_bodies(f["scene"])
    return w, f


def _step(w, n):
    if isinstance(w, OracleWorld):
        w.step(n)
    else:
        w.simulate(n)


def test_oracle_sat_cases():
    f = load("sat_cases.npz")
    desc = WorldDesc.make(**scenes.world_desc())
    for i in range(len(f["have"])):
        m = collide(desc, f["a"][i:i + 1], f["b"][i:i + 1])
        assert (m is not None) == bool(f["have"][i])
        if m is not None:
            assert m.tobytes() == f["manifolds"][i].tobytes()


@pytest.mark.parametrize("impl", ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def test_pile64_fixture(impl):
    cls = OracleWorld if impl == "oracle" else beatbox.RigidBodyManager
    w, f = _pile(cls)
    _step(w, 1)
    _check_taps(w, f, "s1_")
    _step(w, 5)
    _check_taps(w, f, "s6_")


@pytest.mark.parametrize("impl", ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def test_config1_fixture(impl):
    f = load("config1.npz")
    desc = WorldDesc.make(**scenes.world_desc(iterations=8))
    w = OracleWorld(desc) if impl == "oracle" else beatbox.RigidBodyManager(desc)
    w.add_bodies(scenes.cubes_grid())
    for s in range(4):
        _step(w, 1)
        assert _same(w.tap(TAP_MORTON, np.uint32), f[f"s{s}_morton"])
        assert _same(w.tap(TAP_SORTED, np.uint32).reshape(-1, 4)[:, 2].copy(), f[f"s{s}_order"])
        assert _same(w.tap(TAP_PAIRS, np.uint32), f[f"s{s}_pairs"])
        assert _same(w.tap(TAP_ISLANDS, np.uint32), f[f"s{s}_islands"])
        assert _same(w.tap(TAP_EDGE_COLORS, np.uint32), f[f"s{s}_colors"])
        m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
        keys = np.stack([m["obb1_index"], m["obb2_index"], m["key"].astype(np.uint32),
                         m["contact_count"].astype(np.uint32)], 1)
        assert _same(keys, f[f"s{s}_mkeys"])
    got, want = w.read_bodies(), f["s3_bodies"]
    for k in ("position", "rotation", "velocity", "omega", "flags"):
        assert _same(got[k], want[k]), k


@pytest.mark.parametrize("impl", ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
@pytest.mark.parametrize("solver", ["pgs", "avbd"])
def test_freefall_fixture(impl, solver):
    f = load("freefall.npz")
    sid = scenes.SOLVER_PGS_SOFT if solver == "pgs" else scenes.SOLVER_AVBD
    desc = WorldDesc.make(**scenes.world_desc(solver=sid))
    w = OracleWorld(desc) if impl == "oracle" else beatbox.RigidBodyManager(desc)
    w.add_bodies(f[solver + "_scene"])
    traj = f[solver]
    for s in range(len(traj)):
        _step(w, 1)
        r = w.read_bodies()[0]
        for k in ("position", "rotation", "velocity", "omega"):
            assert r[k].tobytes() == traj[s][k].tobytes(), (s, k)
