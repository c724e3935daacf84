"""bb_world_step_noalloc against the entry points it stands for.

The allocation-free call steps a fresh world in the library's static device
scratch.  Its bodies (by id, with island slots and rows) and its last step's
counters must equal, bit for bit, what bb_world_create + bb_world_add_bodies +
bb_world_step + bb_world_read_bodies + bb_world_read_stats give for the same
world.  Two calls in a row, the second with fewer bodies: the second call
reuses the scratch the first one left, so anything it read before writing
would show up here.
"""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes

pytestmark = pytest.mark.gpu

FIELDS = ("id", "flags", "island_index", "row", "position", "rotation", "velocity", "omega")


def allocating(desc, bodies, nsteps):
    w = beatbox.RigidBodyManager(desc)
    w.add_bodies(bodies)
    w.simulate(nsteps)
    return w.read_bodies(), w.read_back_sim_config()


def check_equal(rc, got, stats, ref, ref_stats, what):
    assert rc == 0, f"{what}: status {rc}"
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f"{what}: field {f} differs"
    for name, _ in beatbox.StepStats._fields_:
        assert getattr(stats, name) == ref_stats[name], f"{what}: counter {name}"


def test_noalloc_equals_world_step_twice_second_smaller():
    # 1: the 1,000-cube drop under PGS_SOFT, into its contact phase
    desc1 = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    cubes = scenes.cubes_grid()
    print(f"noalloc call 1: cubes_grid n={len(cubes)} steps=70", flush=True)
    rc, got, st = beatbox.step_noalloc(desc1, cubes, 70)
    ref, ref_st = allocating(desc1, cubes, 70)
    assert ref_st["collision_count"] > 100
    check_equal(rc, got, st, ref, ref_st, "cubes_grid PGS_SOFT")
    # 2: a smaller mixed pile under AVBD (63-bit keys), in the same scratch
    desc2 = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                       morton_bits=63, quirks=0))
    pile = scenes.serving_pile(0xA11C, 300)
    print(f"noalloc call 2: serving_pile n={len(pile)} steps=60", flush=True)
    rc, got, st = beatbox.step_noalloc(desc2, pile, 60)
    ref, ref_st = allocating(desc2, pile, 60)
    assert ref_st["collision_count"] > 50
    check_equal(rc, got, st, ref, ref_st, "serving_pile AVBD")
    # 3: smaller again, the per-island plain-PGS path, zero steps then some
    desc3 = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_PGS, iterations=8))
    desc3.flags &= ~scenes.SIM_USE_GRAPH_COLORING
    boxes = scenes.random_boxes(0xA11D, 64, extent=2.0, y0=0.55, y1=4.0)
    print(f"noalloc call 3: random_boxes n={len(boxes)} steps=0 and 40", flush=True)
    for nsteps in (0, 40):
        rc, got, st = beatbox.step_noalloc(desc3, boxes, nsteps)
        ref, ref_st = allocating(desc3, boxes, nsteps)
        check_equal(rc, got, st, ref, ref_st, f"random_boxes plain PGS, {nsteps} steps")
