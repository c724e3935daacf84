"""Full-size workloads checked through properties that do not need the
oracle to run at their size:
- the sphere gas (1,000,000 spheres): the broad phase's pair set equals the
  exact set of overlapping sphere AABBs (centre +- radius, inclusive), found
  on the host with a uniform grid;
- the large pile (100,000 boxes and spheres, AVBD): the team path (a group of
  workgroups with team barriers) gives the same trajectory and counters as
  the one-workgroup path, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes


def exact_aabb_pairs(mn: np.ndarray, mx: np.ndarray, cell: float) -> set:
    """All (i, j), i < j, whose boxes overlap (inclusive), for boxes no wider
    than `cell`: overlapping boxes have min corners in the same or adjacent
    grid cells, so each box is tested against the 27 cells around its own.
    The cell key is linear in the cell coordinates, so a neighbour cell's key
    is the box's key plus a constant and the queries stay sorted."""
    c = np.floor(mn / cell).astype(np.int64)
    c -= c.min(axis=0)
    dims = c.max(axis=0) + 3
    k = ((c[:, 0] + 1) * dims[1] + (c[:, 1] + 1)) * dims[2] + (c[:, 2] + 1)
    order = np.argsort(k, kind="stable")
    sk = k[order]
    out = set()
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                q = sk + (dx * dims[1] + dy) * dims[2] + dz
                lo = np.searchsorted(sk, q, "left")
                cnt = np.searchsorted(sk, q, "right") - lo
                tot = int(cnt.sum())
                if tot == 0:
                    continue
                first = np.repeat(np.cumsum(cnt) - cnt, cnt)
                ii = order[np.repeat(np.arange(len(sk)), cnt)]
                jj = order[np.repeat(lo, cnt) + (np.arange(tot) - first)]
                keep = ii < jj
                ii, jj = ii[keep], jj[keep]
                hit = np.all((mn[ii] <= mx[jj]) & (mx[ii] >= mn[jj]), axis=1)
                out.update(zip(ii[hit].tolist(), jj[hit].tolist()))
    return out


def test_exact_aabb_pairs_matches_brute_force():
    """The host pair finder itself, against O(n^2) on a small dense set."""
    rng = np.random.default_rng(7)
    p = rng.uniform(-3.0, 3.0, (300, 3)).astype(np.float32)
    r = rng.uniform(0.1, 0.5, 300).astype(np.float32)[:, None]
    mn, mx = p - r, p + r
    brute = {(i, j) for i in range(300) for j in range(i + 1, 300)
             if np.all(mn[i] <= mx[j]) and np.all(mx[i] >= mn[j])}
    assert exact_aabb_pairs(mn, mx, 1.0) == brute and len(brute) > 100


@pytest.mark.gpu
def test_sphere_gas_full_size_pairs():
    bodies = scenes.sphere_gas()
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0, gravity=0.0))
    b = beatbox.Batch()
    b.add_world(desc, bodies)
    b.step(1)
    b.sync()
    st = b.stats(0)
    assert st["overflow"] == 0
    rows = b.tap(0, beatbox.TAP_PAIRS, np.uint32).reshape(-1, 2)
    ids = b.read_world(0, len(bodies), order=beatbox.ORDER_BY_ROW)["id"]
    got = {(min(a, c), max(a, c)) for a, c in zip(ids[rows[:, 0]].tolist(), ids[rows[:, 1]].tolist())}
    b.destroy()
    p = bodies["position"].astype(np.float32)
    r = bodies["radius"].astype(np.float32)[:, None]
    want = exact_aabb_pairs(p - r, p + r, 1.0)
    assert len(rows) == st["broad_phase_collision_count"] == len(got)
    assert got == want and len(want) > 500


@pytest.mark.gpu
def test_large_pile_full_size_team_equals_workgroup():
    bodies = scenes.large_pile()
    desc = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    runs = []
    # the team at all of the resident workgroups, rows handed to workgroups
    # past the default walk budget; the one workgroup walking every row alone
    for team, budget, team_blocks in ((8192, 0, 2048), (0, 0xFFFFFFFF, 0)):
        b = beatbox.Batch()
        b.add_world(desc, bodies)
        b.set_param(beatbox.PARAM_TEAM_BODIES, team)
        b.set_param(beatbox.PARAM_WALK_BUDGET, budget)
        b.set_param(beatbox.PARAM_TEAM_BLOCKS, team_blocks)
        b.step(2)
        b.sync()
        runs.append((b.read_world(0, len(bodies)), b.stats(0)))
        b.destroy()
    (x, sx), (y, sy) = runs
    for f in ("position", "rotation", "velocity", "omega", "flags"):
        assert x[f].tobytes() == y[f].tobytes(), f
    for k in ("broad_phase_collision_count", "collision_count", "contact_count", "avbd_color_count",
              "dbg_id_sum", "overflow"):
        assert sx[k] == sy[k], k
    assert sx["collision_count"] > 0
