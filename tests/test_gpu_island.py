"""The per-island solver path (USE_GRAPH_COLORING off: the reference's CPS /
CS / IP / CSR passes, one serial walk per contact island, rigid_body_manager
.cpp:1390-1399) with plain PGS and PGS_SOFT, against the oracle bit for bit:
poses, velocities, manifold taps (impulses, tangent masses, colours) and
counters.  The GPU solves it with parallel sweeps over dependency levels plus
a serial overflow bucket; these scenes reach levels past the 32 buckets, so
both parts run."""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes
from oracle_bind import OracleWorld
from parity_util import compare_step

pytestmark = pytest.mark.gpu

ISLAND = scenes.SIM_DEFAULT_FLAGS & ~scenes.SIM_USE_GRAPH_COLORING


def _parity(bodies, desc_kw, steps, toggles=None):
    desc = beatbox.WorldDesc.make(**desc_kw)
    gpu = beatbox.RigidBodyManager(desc)
    gpu.add_bodies(bodies)
    ora = OracleWorld(desc)
    ora.add_bodies(bodies)
    for s in range(steps):
        if toggles and s in toggles:
            on = toggles[s]
            if on:
                gpu.set_sim_flags(scenes.SIM_USE_GRAPH_COLORING)
            else:
                gpu.clear_sim_flags(scenes.SIM_USE_GRAPH_COLORING)
            ora.set_flags(gpu.get_sim_flags())
        gpu.simulate(1)
        ora.step(1)
        bad = compare_step(gpu, ora, desc.morton_bits, s)
        assert not bad, "\n".join(bad[:20])
    gpu.destroy()


def test_island_plain_pgs_parity_quirks():
    bodies = scenes.random_boxes(0xBB31, 64, extent=3.0, y0=0.55, y1=5.0)
    _parity(bodies, scenes.world_desc(solver=scenes.SOLVER_PGS, iterations=8, flags=ISLAND), steps=30)


def test_island_pgs_soft_63bit():
    bodies = scenes.random_boxes(0xBB32, 96, extent=4.0, y0=0.55, y1=8.0)
    kw = scenes.world_desc(iterations=8, flags=ISLAND, morton_bits=63, quirks=0)
    _parity(bodies, kw, steps=25)


def test_island_plain_pgs_pile_spheres_and_boxes():
    bodies = scenes.serving_pile(41, 300)
    kw = scenes.world_desc(solver=scenes.SOLVER_PGS, iterations=8, flags=ISLAND, morton_bits=63, quirks=0)
    _parity(bodies, kw, steps=20)


def test_island_plain_pgs_no_accumulation_no_friction():
    bodies = scenes.random_boxes(0xBB33, 48, extent=2.0, y0=0.55, y1=4.0)
    flags = ISLAND & ~(1 | 2)
    _parity(bodies, scenes.world_desc(solver=scenes.SOLVER_PGS, iterations=4, flags=flags,
                                      morton_bits=63, quirks=0), steps=15)


def test_colouring_switched_at_run_time():
    """PGS switches between the colour sweeps (run as PGS_SOFT) and the
    per-island plain PGS between steps, warm start carried across."""
    bodies = scenes.random_boxes(0xBB34, 64, extent=3.0, y0=0.55, y1=5.0)
    kw = scenes.world_desc(solver=scenes.SOLVER_PGS, iterations=8, morton_bits=63, quirks=0)
    _parity(bodies, kw, steps=24, toggles={6: False, 12: True, 18: False})


def test_island_worlds_in_a_batch_and_teams_equal_alone():
    """Island-mode worlds stepped in one batch (graph replay, a team for the
    largest) equal each world stepped alone."""
    plain = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_PGS, iterations=8, flags=ISLAND,
                                                       morton_bits=63, quirks=0))
    soft = beatbox.WorldDesc.make(**scenes.world_desc(iterations=6, flags=ISLAND, morton_bits=63, quirks=0))
    colour = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    worlds = [(plain, scenes.serving_pile(43, 700)), (soft, scenes.batch_world(0xBB04, 6, 400)),
              (colour, scenes.serving_pile(44, 200)), (plain, scenes.random_boxes(0xBB35, 64, extent=3.0,
                                                                                   y0=0.55, y1=5.0))]
    b = beatbox.Batch()
    for d, bodies in worlds:
        b.add_world(d, bodies)
    b.set_param(beatbox.PARAM_TEAM_BODIES, 600)
    b.step(10, use_graph=True)
    b.sync()
    for k, (d, bodies) in enumerate(worlds):
        w = beatbox.RigidBodyManager(d)
        w.add_bodies(bodies)
        w.simulate(10)
        x, y = b.read_world(k, len(bodies)), w.read_bodies()
        for f in ("position", "rotation", "velocity", "omega", "flags"):
            assert x[f].tobytes() == y[f].tobytes(), (k, f)
        assert b.stats(k)["collision_count"] == w.read_back_sim_config()["collision_count"] > 0
        w.destroy()
    b.destroy()


def test_large_island_world_team_equals_workgroup():
    """A 5,000-body plain-PGS world on the per-island path: the team path
    (level fixpoint and sweeps over team barriers) equals one workgroup, bit
    for bit."""
    bodies = scenes.serving_pile(45, 5000)
    desc = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_PGS, iterations=8, flags=ISLAND,
                                                      morton_bits=63, quirks=0))
    runs = []
    for team in (2048, 0):
        b = beatbox.Batch()
        b.add_world(desc, bodies)
        b.set_param(beatbox.PARAM_TEAM_BODIES, team)
        b.step(12)
        b.sync()
        runs.append((b.read_world(0, len(bodies)), b.stats(0)))
        b.destroy()
    (x, sx), (y, sy) = runs
    for f in ("position", "rotation", "velocity", "omega", "flags"):
        assert x[f].tobytes() == y[f].tobytes(), f
    for k in ("collision_count", "contact_count", "island_count", "overflow"):
        assert sx[k] == sy[k], k
    assert sx["collision_count"] > 1000 and sx["overflow"] == 0
