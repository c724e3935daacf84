"""The batched and serving workloads at their real size and shape.

- The 4,096-world batch (seed 0xBB04, 16-2,048 bodies per world, box
  pyramids and towers, PGS_SOFT 8 iterations, 63-bit keys) is stepped the way
  the benchmark steps it: two launch lanes, the autotuner, hipGraph replay.
  32 worlds spread over the size range, the largest included, are compared
  bit for bit with the oracle and with the same world stepped alone.
- A serving stream shaped like the benchmark's (AVBD 10 iterations, random
  piles of 64-20,000 bodies, staggered arrivals, workgroup teams from 1,024
  bodies, more requests than slots): every retired world equals the same
  world stepped alone, bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes
from oracle_bind import OracleWorld

pytestmark = pytest.mark.gpu

FIELDS = ("position", "rotation", "velocity", "omega", "flags")
STATS = ("broad_phase_collision_count", "collision_count", "contact_count", "island_count",
         "sleeping_count", "graph_color_count", "avbd_color_count", "dbg_id_sum")


def _sample(sizes: np.ndarray, k: int) -> list:
    """k world indices spread over the size range (by rank), the largest world
    included."""
    order = np.argsort(sizes, kind="stable")
    picks = {int(order[int(round(q))]) for q in np.linspace(0, len(sizes) - 1, k)}
    picks.add(int(order[-1]))
    return sorted(picks)


def test_batch_4096_worlds_sampled_equal_oracle_and_alone():
    seed, steps, warm = 0xBB04, 12, 2
    sizes = scenes.batch_world_sizes(seed, 4096, 16, 2048)
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    b = beatbox.Batch()
    worlds = []
    for w, n in enumerate(sizes):
        bodies = scenes.batch_world(seed, w, int(n))
        worlds.append(bodies)
        b.add_world(desc, bodies)
    assert b.world_count == 4096 and b.body_count > 1_500_000
    b.step(warm, use_graph=True)
    b.sync()
    b.autotune(2)  # replays the next step per candidate, state unchanged
    b.step(steps - warm, use_graph=True)
    b.sync()
    assert b.overflow() == []
    picks = _sample(np.array([len(x) for x in worlds]), 32)
    assert len(picks) >= 30 and max(len(worlds[w]) for w in picks) == max(len(x) for x in worlds)
    for w in picks:
        bodies = worlds[w]
        got = b.read_world(w, len(bodies))
        sg = b.stats(w)
        ora = OracleWorld(desc)
        ora.add_bodies(bodies)
        ora.step(steps)
        ref = ora.read_bodies()
        so = ora.stats()
        for f in FIELDS:
            assert got[f].tobytes() == ref[f].tobytes(), ("oracle", w, len(bodies), f)
        for k in STATS:
            assert sg[k] == so[k], ("oracle", w, k)
        alone = beatbox.RigidBodyManager(desc)
        alone.add_bodies(bodies)
        alone.simulate(steps)
        ga = alone.read_bodies()
        for f in FIELDS:
            assert got[f].tobytes() == ga[f].tobytes(), ("alone", w, f)
        alone.destroy()
    b.destroy()


@pytest.mark.parametrize("team_bodies,split_bodies,npools", [(1024, 0, 1), (1024, 1024, 2), (8192, 1024, 4),
                                                              (4096, 256, 4)])
def test_serving_stream_shape_equals_alone(team_bodies, split_bodies, npools):
    """One lockstep pool, two pools (one workgroup per world below 1,024
    bodies, teams above) and four pools (one workgroup per world below the
    split s, 512-thread workgroups over [s, 4s) and [4s, team), teams above),
    each stepping at its own cadence: every world equals itself stepped
    alone."""
    desc = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    sizes = [20000, 64, 3000, 150, 1500, 700, 9000, 90, 2500, 400, 1100, 5000]
    horizons = [7, 12, 9, 6, 11, 8, 6, 10, 7, 9, 8, 6]
    reqs = [(scenes.serving_pile(0xBB50 + k, n), h) for k, (n, h) in enumerate(zip(sizes, horizons))]
    print(f"serving stream: team_bodies {team_bodies}, split_bodies {split_bodies}", flush=True)
    srv = beatbox.Server(budget_bytes=24 << 30, max_worlds=6, team_bodies=team_bodies, split_bodies=split_bodies)
    tickets = []
    for k, (bodies, h) in enumerate(reqs):
        tickets.append(srv.submit(desc, bodies, h))
        srv.step(1 + k % 3)  # arrivals while others run
    srv.drain()
    rep = srv.report()
    assert rep["retired"] == len(reqs) and rep["reruns"] == 0
    assert rep["step_ms_p99"] >= rep["step_ms_p50"] > 0
    pools = rep["pools"]
    assert len(pools) == npools
    for k, p in enumerate(pools):  # every world's steps counted once, in the pool of its size
        hi = pools[k + 1]["min_bodies"] if k + 1 < len(pools) else 1 << 32
        want = sum(h for n, h in zip(sizes, horizons) if p["min_bodies"] <= n < hi)
        assert p["world_steps"] == want, (k, p)
        assert p["step_ms_p99"] >= p["step_ms_p50"] > 0
    for t, (bodies, h) in zip(tickets, reqs):
        state, got, stats = srv.poll(t, len(bodies))
        assert state == beatbox.REQ_DONE
        alone = beatbox.RigidBodyManager(desc)
        alone.add_bodies(bodies)
        alone.simulate(h)
        ref = alone.read_bodies()
        sr = alone.read_back_sim_config()
        alone.destroy()
        for f in FIELDS:
            assert got[f].tobytes() == ref[f].tobytes(), (len(bodies), f)
        for k in ("collision_count", "contact_count", "avbd_color_count", "sleeping_count"):
            assert stats[k] == sr[k], (len(bodies), k)
    srv.destroy()
