"""Pair capacity past the reference's fixed buffer caps (RigidBodyManager,
src/rigid_body_manager.cpp:132-282): a world
starts with 8n+64 pair slots; a step whose broad phase finds more pairs sets
the overflow counter, and the single-world step then doubles the capacity and
runs the step again from the state it read.  Spheres packed closer than their
radius give 27 to 42 pairs per body, so the first step overflows; the grown
world follows the oracle (which has no capacity) bit for bit, taps included,
a batch world of the same scene reports the overflow instead, and a grown
world's checkpoint restores with room for its pairs."""
from __future__ import annotations

import pytest

import beatbox
import scenes
from oracle_bind import OracleWorld
from parity_util import compare_step

pytestmark = pytest.mark.gpu

DESC = scenes.world_desc(iterations=8, morton_bits=63, quirks=0, gravity=0.0)


@pytest.mark.parametrize("side", [5, 10])
def test_pair_overflow_grows_and_matches_oracle(side):
    bodies = scenes.dense_spheres(0xBB71 + side, side=side)
    n = len(bodies)
    desc = beatbox.WorldDesc.make(**DESC)
    gpu = beatbox.RigidBodyManager(desc)
    gpu.add_bodies(bodies)
    ora = OracleWorld(desc)
    ora.add_bodies(bodies)
    for s in range(6):
        gpu.simulate(1)
        ora.step(1)
        bad = compare_step(gpu, ora, desc.morton_bits, s)
        assert not bad, "\n".join(bad[:20])
        st = gpu.read_back_sim_config()
        assert st["overflow"] == 0, s
        if s == 0:
            assert st["broad_phase_collision_count"] > 3 * (8 * n + 64)
    gpu.destroy()


def test_batch_world_reports_the_overflow():
    bodies = scenes.dense_spheres(0xBB76)
    b = beatbox.Batch()
    b.add_world(beatbox.WorldDesc.make(**DESC), bodies)
    b.step(1)
    with pytest.raises(beatbox.BBError) as e:
        b.sync()
    assert e.value.code == 5  # BB_ERR_OVERFLOW
    st = b.stats(0)
    assert st["overflow"] & 1 and st["broad_phase_collision_count"] > 8 * len(bodies) + 64
    assert b.overflow() == [0]
    b.sync()  # the mark is collected
    b.destroy()


def test_overflow_in_an_early_step_is_kept_until_collected():
    """Worlds 0 and 2 overflow in the first of three steps only (a block of
    spheres flying apart); after the three steps their own counters say
    nothing, but the marks stay until bb_batch_overflow collects them, and
    the sync reports BB_ERR_OVERFLOW until then.  The single-world step of
    the same world grows its capacity and replays, and equals the oracle."""
    desc = beatbox.WorldDesc.make(**DESC)
    boom = scenes.expanding_spheres(0xBB78)
    calm = scenes.random_boxes(0xBB79, 40, extent=2.0, y0=0.55, y1=3.0)
    b = beatbox.Batch()
    for bodies in (boom, calm, scenes.expanding_spheres(0xBB7A)):
        b.add_world(desc, bodies)
    b.step(3, use_graph=True)
    with pytest.raises(beatbox.BBError) as e:
        b.sync()
    assert e.value.code == 5
    assert all(b.stats(w)["overflow"] == 0 for w in range(3))
    assert b.overflow() == [0, 2]
    b.sync()
    assert b.overflow() == []
    b.step(2)
    b.sync()
    b.destroy()
    gpu = beatbox.RigidBodyManager(desc)
    gpu.add_bodies(boom)
    ora = OracleWorld(desc)
    ora.add_bodies(boom)
    for s in range(3):
        gpu.simulate(1)
        ora.step(1)
        bad = compare_step(gpu, ora, desc.morton_bits, s)
        assert not bad, "\n".join(bad[:20])
    gpu.sync()
    gpu.destroy()


def test_grown_world_checkpoint_continues_bit_identically():
    bodies = scenes.dense_spheres(0xBB77)
    desc = beatbox.WorldDesc.make(**DESC)
    a = beatbox.RigidBodyManager(desc)
    a.add_bodies(bodies)
    a.simulate(2)
    assert a.read_back_sim_config()["broad_phase_collision_count"] > 8 * len(bodies) + 64
    blob = a.snapshot()
    b = beatbox.RigidBodyManager.restore(blob, warm_start=True)
    a.simulate(5)
    b.simulate(5)
    x, y = a.read_bodies(), b.read_bodies()
    for f in ("position", "rotation", "velocity", "omega", "flags"):
        assert x[f].tobytes() == y[f].tobytes(), f
    a.destroy()
    b.destroy()


def test_batch_pair_scale_gives_room_and_equals_growth():
    """BB_PARAM_PAIR_SCALE: the same dense world that overflows a batch at
    the default capacity fits at four times it, reports no overflow, and
    equals the single world that grows its capacity and replays."""
    desc = beatbox.WorldDesc.make(**DESC)
    bodies = scenes.dense_spheres(0xBB76)
    b = beatbox.Batch()
    b.set_param(beatbox.PARAM_PAIR_SCALE, 4)
    assert b.get_param(beatbox.PARAM_PAIR_SCALE) == 4
    b.add_world(desc, bodies)
    b.add_world(desc, scenes.random_boxes(0xBB79, 40, extent=2.0, y0=0.55, y1=3.0))
    b.step(1)
    b.sync()  # no overflow mark although the first step has more pairs than 8n + 64
    assert b.stats(0)["broad_phase_collision_count"] > 8 * len(bodies) + 64
    b.step(3, use_graph=True)
    b.sync()
    assert b.overflow() == []
    w = beatbox.RigidBodyManager(desc)
    w.add_bodies(bodies)
    w.simulate(4)
    got, ref = b.read_world(0, len(bodies)), w.read_bodies()
    for f in ("position", "rotation", "velocity", "omega", "flags"):
        assert got[f].tobytes() == ref[f].tobytes(), f
    with pytest.raises(beatbox.BBError):
        b.set_param(beatbox.PARAM_PAIR_SCALE, 0)
    b.destroy()
    w.destroy()


def test_overflow_in_a_middle_step_only():
    """A layer of spheres flying through the plane of twenty static boards
    (scenes.crossing_cloud): 190 board pairs against 384 pair slots, 590
    pairs in steps 9 and 10 while every sphere's box overlaps every board's,
    190 again after.  A batch world flags steps 9 and 10 only; its mark
    stays until collected, through the later calm steps.  The single world
    grows its capacity in step 9 and replays that step from the state it
    read, and every step, taps included, equals the oracle bit for bit."""
    desc = beatbox.WorldDesc.make(**DESC)
    bodies = scenes.crossing_cloud(0xBB7D)
    cap = 8 * len(bodies) + 64
    b = beatbox.Batch()
    b.add_world(desc, bodies)
    flagged = []
    for s in range(14):
        b.step(1)
        try:
            b.sync()
        except beatbox.BBError as e:
            assert e.code == 5 and s in (9, 10), s  # BB_ERR_OVERFLOW until the mark is collected
        st = b.stats(0)
        if st["overflow"]:
            flagged.append(s)
        assert st["broad_phase_collision_count"] == (590 if s in (9, 10) else 190), s
        if s == 10:
            # the mark outlives the counters: collected only now
            assert b.overflow() == [0]
            assert b.overflow() == []
    assert flagged == [9, 10]
    b.destroy()
    gpu = beatbox.RigidBodyManager(desc)
    gpu.add_bodies(bodies)
    ora = OracleWorld(desc)
    ora.add_bodies(bodies)
    for s in range(14):
        gpu.simulate(1)
        ora.step(1)
        bad = compare_step(gpu, ora, desc.morton_bits, s)
        assert not bad, (s, "\n".join(bad[:20]))
        assert gpu.read_back_sim_config()["overflow"] == 0, s
    assert 590 > cap
    gpu.destroy()
