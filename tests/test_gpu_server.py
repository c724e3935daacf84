"""Serving loop: continuous batching through the C-ABI.  Each world's final
state equals the same world stepped alone (bit-exact: the step is a pure
function of the world's own state, wherever it sits in the device arrays and
whatever else is resident); admission respects the HBM budget and the
resident-body cap; counters and latency percentiles are reported."""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes

pytestmark = pytest.mark.gpu

FIELDS = ("position", "rotation", "velocity", "omega", "flags")


def _alone(desc, bodies, steps):
    w = beatbox.RigidBodyManager(desc)
    w.add_bodies(bodies)
    w.simulate(steps)
    out = w.read_bodies()
    st = w.read_back_sim_config()
    w.destroy()
    return out, st


def _requests():
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    parity = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    # plain PGS on the per-island path (USE_GRAPH_COLORING off)
    island = beatbox.WorldDesc.make(**scenes.world_desc(
        solver=scenes.SOLVER_PGS, iterations=8, flags=scenes.SIM_DEFAULT_FLAGS & ~scenes.SIM_USE_GRAPH_COLORING,
        morton_bits=63, quirks=0))
    reqs = [
        (pgs, scenes.batch_world(0xBB04, 3, 120), 9),
        (island, scenes.serving_pile(14, 160), 8),
        (avbd, scenes.serving_pile(11, 150), 7),
        (parity, scenes.random_boxes(0xBB11, 64, extent=3.0, y0=0.55, y1=5.0), 12),
        (pgs, scenes.batch_world(0xBB04, 8, 300), 1),
        (avbd, scenes.serving_pile(12, 90), 15),
        (pgs, scenes.serving_pile(13, 200), 5),
    ]
    return reqs


def test_server_results_equal_alone():
    reqs = _requests()
    srv = beatbox.Server(budget_bytes=1 << 30, max_worlds=4)
    tickets = []
    # staggered arrivals: two now, the rest while others are running
    for d, bodies, steps in reqs[:2]:
        tickets.append(srv.submit(d, bodies, steps))
    srv.step(3)
    for d, bodies, steps in reqs[2:]:
        tickets.append(srv.submit(d, bodies, steps))
        srv.step(2)
    srv.drain()
    rep = srv.report()
    assert rep["submitted"] == rep["admitted"] == rep["retired"] == len(reqs)
    assert rep["queued"] == 0 and rep["resident_worlds"] == 0 and rep["bytes_in_use"] == 0
    assert rep["world_steps"] == sum(r[2] for r in reqs)
    assert rep["body_steps"] == sum(len(r[1]) * r[2] for r in reqs)
    assert rep["step_ms_p50"] > 0 and rep["step_ms_p99"] >= rep["step_ms_p50"]
    for t, (d, bodies, steps) in zip(tickets, reqs):
        state, got, stats = srv.poll(t, len(bodies))
        assert state == beatbox.REQ_DONE, t
        ref, ref_stats = _alone(d, bodies, steps)
        for f in FIELDS:
            assert got[f].tobytes() == ref[f].tobytes(), (t, f)
        assert stats["collision_count"] == ref_stats["collision_count"]
        assert stats["contact_count"] == ref_stats["contact_count"]
        # released after the first poll
        assert srv.poll(t, len(bodies))[0] == beatbox.REQ_UNKNOWN
    srv.destroy()


def test_server_admission_limits():
    """A resident-body cap admits one world at a time; a budget too small for
    a world refuses it at submit; every accepted request completes."""
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    worlds = [scenes.batch_world(0xBB04, w, 100 + 30 * w) for w in range(4)]
    srv = beatbox.Server(budget_bytes=1 << 30, max_worlds=8, max_resident_bodies=150)
    ts = [srv.submit(pgs, b, 4) for b in worlds]
    srv.step(1)
    rep = srv.report()
    assert rep["resident_worlds"] == 1 and rep["queued"] == 3
    assert srv.poll(ts[1], len(worlds[1]))[0] == beatbox.REQ_QUEUED
    srv.drain()
    rep = srv.report()
    assert rep["retired"] == 4 and rep["server_steps"] == 16
    for t, b in zip(ts, worlds):
        state, got, _ = srv.poll(t, len(b))
        assert state == beatbox.REQ_DONE
        ref, _ = _alone(pgs, b, 4)
        assert got["position"].tobytes() == ref["position"].tobytes()
    srv.destroy()
    tiny = beatbox.Server(budget_bytes=4 << 20, max_worlds=2)
    with pytest.raises(beatbox.BBError) as e:
        tiny.submit(pgs, scenes.batch_world(0xBB04, 5, 2000), 3)
    assert e.value.code == 7  # BB_ERR_BUDGET
    tiny.destroy()


def test_server_reruns_a_world_that_overflowed():
    """A served world whose step drops pairs past its capacity is run again
    from its first step with twice the pair capacity (as many times as it
    takes) before its result is reported, so the result equals the world
    stepped alone by bb_world_step, which grows and replays the same way.
    One world overflows in its first step only (spheres flying apart), one in
    every step (a dense block); a calm world beside them is unaffected."""
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0, gravity=0.0))
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    reqs = [(desc, scenes.expanding_spheres(0xBB7B), 4), (pgs, scenes.batch_world(0xBB04, 9, 150), 6),
            (desc, scenes.dense_spheres(0xBB7C), 5)]
    srv = beatbox.Server(budget_bytes=1 << 30, max_worlds=4)
    tickets = [srv.submit(d, b, s) for d, b, s in reqs]
    srv.drain()
    rep = srv.report()
    assert rep["reruns"] >= 2
    for t, (d, bodies, steps) in zip(tickets, reqs):
        state, got, stats = srv.poll(t, len(bodies))
        assert state == beatbox.REQ_DONE, t
        ref, ref_stats = _alone(d, bodies, steps)
        for f in FIELDS:
            assert got[f].tobytes() == ref[f].tobytes(), (t, f)
        assert stats["overflow"] == 0
        assert stats["broad_phase_collision_count"] == ref_stats["broad_phase_collision_count"]
    srv.destroy()


def test_server_checkpoint_restores_bit_exact():
    """A checkpoint taken mid-run (worlds resident in every pool, a team
    world among them, requests still queued, results not yet polled) and
    restored into a new server, with the same pools or with one pool: every
    request's result equals the original server's and the world stepped
    alone, bit for bit; new tickets continue after the checkpoint's."""
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    reqs = _requests() + [(avbd, scenes.serving_pile(15, 1300), 9), (avbd, scenes.serving_pile(16, 600), 8)]
    kw = dict(budget_bytes=1 << 31, max_worlds=3, team_bodies=1000, split_bodies=256)
    a = beatbox.Server(**kw)
    tickets = [a.submit(d, b, s) for d, b, s in reqs]
    a.step(4)
    blob = a.checkpoint()
    rep = a.report()
    print(f"checkpoint: {len(blob)} bytes, resident {rep['resident_worlds']}, queued {rep['queued']}, "
          f"retired {rep['retired']}", flush=True)
    assert rep["resident_worlds"] >= 3 and rep["queued"] >= 1 and rep["retired"] >= 1
    restored = [beatbox.Server.restore(blob, **kw),
                beatbox.Server.restore(blob, budget_bytes=1 << 31, max_worlds=16, team_bodies=1000)]
    a.drain()
    for srv in restored:
        srv.drain()
    for t, (d, bodies, steps) in zip(tickets, reqs):
        state, want, want_stats = a.poll(t, len(bodies))
        assert state == beatbox.REQ_DONE, t
        ref, _ = _alone(d, bodies, steps)
        for f in FIELDS:
            assert want[f].tobytes() == ref[f].tobytes(), (t, f)
        for k, srv in enumerate(restored):
            state, got, stats = srv.poll(t, len(bodies))
            assert state == beatbox.REQ_DONE, (k, t)
            for f in FIELDS:
                assert got[f].tobytes() == want[f].tobytes(), (k, t, f)
            for key in ("collision_count", "contact_count", "broad_phase_collision_count", "overflow"):
                assert stats[key] == want_stats[key], (k, t, key)
    for srv in restored:
        assert srv.submit(reqs[0][0], reqs[0][1], 1) > max(tickets)
        srv.destroy()
    a.destroy()


def test_server_voxel_requests_equal_alone_and_checkpoint():
    """Voxel shapes registered with the server: requests with concave voxel
    bodies (PGS_SOFT and AVBD, six pair slots per default one) are served
    next to a box pile, each equal to the world stepped alone; a checkpoint
    taken with them resident carries the shapes, and the restored server
    finishes them the same way."""
    shapes, mixed = scenes.voxel_mixed_pile(0xBB65, 150)
    _, ref = scenes.voxel_pile()
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    reqs = [(pgs, mixed, 14), (avbd, ref, 12), (pgs, scenes.serving_pile(16, 120), 10)]

    def alone(d, bodies, steps):
        w = beatbox.RigidBodyManager(d)
        for _, cells in shapes:
            w.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
        w.add_bodies(bodies)
        w.simulate(steps)
        out = w.read_bodies()
        w.destroy()
        return out

    want = [alone(*r) for r in reqs]
    srv = beatbox.Server(budget_bytes=1 << 30, max_worlds=4)
    with pytest.raises(beatbox.BBError):  # no shapes registered yet
        srv.submit(pgs, mixed, 1)
    for _, cells in shapes:
        srv.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
    tickets = [srv.submit(*r) for r in reqs]
    srv.step(5)
    blob = srv.checkpoint()
    srv.drain()
    restored = beatbox.Server.restore(blob, budget_bytes=1 << 30, max_worlds=4)
    restored.drain()
    for server in (srv, restored):
        for t, (d, bodies, steps), w in zip(tickets, reqs, want):
            state, got, stats = server.poll(t, len(bodies))
            assert state == beatbox.REQ_DONE, t
            for f in FIELDS:
                assert got[f].tobytes() == w[f].tobytes(), (t, f)
        assert server.report()["reruns"] == 0
    srv.destroy()
    restored.destroy()
