"""Checkpoints (bb_world_snapshot / bb_world_restore): a world restored from
a blob continues bit for bit like the world that wrote it; restored without
warm start it equals the original after skip_warm_starting_once (the
reference's resume after a pause); the read-back right after a restore
equals the one at the snapshot; blobs that are not checkpoints are refused."""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes

pytestmark = pytest.mark.gpu

FIELDS = ("position", "rotation", "velocity", "omega", "flags", "island_index")
STATS = ("broad_phase_collision_count", "collision_count", "contact_count", "island_count",
         "sleeping_count", "dbg_fresh", "dbg_id_sum")


def _worlds():
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    parity = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    return [(pgs, scenes.serving_pile(61, 300)), (avbd, scenes.serving_pile(62, 250)),
            (parity, scenes.cubes_grid())]


def _state(w):
    return w.read_bodies(), w.read_back_sim_config()


def _same(a, b):
    (xa, sa), (xb, sb) = a, b
    for f in FIELDS:
        assert xa[f].tobytes() == xb[f].tobytes(), f
    for k in STATS:
        assert sa[k] == sb[k], k


@pytest.mark.parametrize("k", [0, 1, 2])
def test_restore_continues_bit_identically(k):
    desc, bodies = _worlds()[k]
    a = beatbox.RigidBodyManager(desc)
    a.add_bodies(bodies)
    a.simulate(30)
    at_snap = _state(a)
    blob = a.snapshot()
    a.simulate(20)
    after = _state(a)
    b = beatbox.RigidBodyManager.restore(blob)
    _same(_state(b), at_snap)
    b.simulate(20)
    _same(_state(b), after)
    assert after[1]["contact_count"] > 0
    # cold restore == the original resumed with skip_warm_starting_once
    c = beatbox.RigidBodyManager(desc)
    c.add_bodies(bodies)
    c.simulate(30)
    c.skip_warm_starting_once()
    c.simulate(20)
    d = beatbox.RigidBodyManager.restore(blob, warm_start=False)
    d.simulate(20)
    _same(_state(d), _state(c))
    for w in (a, b, c, d):
        w.destroy()


def test_snapshot_of_fresh_world_and_bad_blobs():
    desc, bodies = _worlds()[0]
    a = beatbox.RigidBodyManager(desc)
    a.add_bodies(bodies)
    blob = a.snapshot()
    a.simulate(10)
    b = beatbox.RigidBodyManager.restore(blob)
    b.simulate(10)
    _same(_state(b), _state(a))
    for bad in (blob[:-1], b"x" * len(blob), blob[:16]):
        with pytest.raises(beatbox.BBError):
            beatbox.RigidBodyManager.restore(bad)
    a.destroy()
    b.destroy()


@pytest.mark.parametrize("k", [0, 1, 2])
def test_restored_world_equals_the_oracle(k):
    """A world checkpointed at step 30, restored (warm) and stepped 20 more
    equals the oracle's uninterrupted world at step 50, every tap included;
    restored cold it equals the oracle with skip_warm_starting_once at 30."""
    from oracle_bind import OracleWorld
    from parity_util import compare_step
    desc, bodies = _worlds()[k]
    a = beatbox.RigidBodyManager(desc)
    a.add_bodies(bodies)
    a.simulate(30)
    blob = a.snapshot()
    a.destroy()
    for warm in (True, False):
        ora = OracleWorld(desc)
        ora.add_bodies(bodies)
        ora.step(30)
        if not warm:
            ora.skip_warm_start_once()
        ora.step(20)
        b = beatbox.RigidBodyManager.restore(blob, warm_start=warm)
        b.simulate(20)
        bad = compare_step(b, ora, desc.morton_bits, 50)
        assert not bad, "\n".join(bad[:20])
        b.destroy()
