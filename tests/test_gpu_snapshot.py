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


def _batch(team_bodies):
    """The three worlds above and the reference's voxel scene under AVBD in
    one batch; worlds of team_bodies bodies or more run as workgroup teams."""
    shapes, vox = scenes.voxel_pile()
    vdesc = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                       morton_bits=63, quirks=0))
    worlds = _worlds() + [(vdesc, vox)]
    b = beatbox.Batch()
    b.set_param(beatbox.PARAM_TEAM_BODIES, team_bodies)
    for _, cells in shapes:
        b.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
    for desc, bodies in worlds:
        b.add_world(desc, bodies)
    return b, worlds, shapes


def _batch_state(b, worlds):
    b.sync()
    return [(b.read_world(w, len(bodies)), b.stats(w)) for w, (_, bodies) in enumerate(worlds)]


def _batch_same(x, y):
    for w, ((xa, sa), (xb, sb)) in enumerate(zip(x, y)):
        for f in FIELDS:
            assert xa[f].tobytes() == xb[f].tobytes(), (w, f)
        assert sa == sb, w


def test_batch_restore_continues_bit_identically():
    """A batch checkpointed at step 30 (bb_batch_snapshot) and restored
    (bb_batch_restore) reads back as it was and continues, graph replay
    included, bit for bit like the batch that wrote it: per world rows,
    counters and step index, for PGS_SOFT and AVBD, a team world and voxel
    shapes.  Restored cold, each world equals the world stepped alone that
    skipped warm starting once at step 30; bad blobs are refused."""
    a, worlds, shapes = _batch(team_bodies=900)
    a.step(30, use_graph=True)
    at_snap = _batch_state(a, worlds)
    blob = a.snapshot()
    a.step(20, use_graph=True)
    after = _batch_state(a, worlds)
    b = beatbox.Batch.restore(blob)
    b.set_param(beatbox.PARAM_TEAM_BODIES, 900)
    _batch_same(_batch_state(b, worlds), at_snap)
    b.step(20, use_graph=True)
    _batch_same(_batch_state(b, worlds), after)
    assert all(s["contact_count"] > 0 for _, s in after)
    assert at_snap[2][1]["step_index"] == 29  # the index of the last step taken
    c = beatbox.Batch.restore(blob, warm_start=False)
    c.step(20)
    cold = _batch_state(c, worlds)
    for w, (desc, bodies) in enumerate(worlds):
        r = beatbox.RigidBodyManager(desc)
        if w == 3:
            for _, cells in shapes:
                r.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
        r.add_bodies(bodies)
        r.simulate(30)
        r.skip_warm_starting_once()
        r.simulate(20)
        x = r.read_bodies()
        for f in ("position", "rotation", "velocity", "omega", "flags"):
            assert x[f].tobytes() == cold[w][0][f].tobytes(), (w, f)
        r.destroy()
    for bad in (blob[:-1], blob + b"\0", b"x" * len(blob), blob[:16]):
        with pytest.raises(beatbox.BBError):
            beatbox.Batch.restore(bad)
    for x in (a, b, c):
        x.destroy()
