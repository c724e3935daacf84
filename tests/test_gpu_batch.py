"""Batched worlds: every world of a batch gives bit-identical results to the
same world stepped alone (segmented kernels are order independent), and a
hipGraph replay equals eager launches."""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes
from oracle_bind import OracleWorld
from parity_util import compare_step

pytestmark = pytest.mark.gpu


def _worlds(seed, sizes):
    return [scenes.batch_world(seed, w, n) for w, n in enumerate(sizes)]


def test_batch_equals_alone_and_oracle():
    sizes = [16, 40, 120, 33, 300, 64]
    worlds = _worlds(0xBB04, sizes)
    kw = scenes.world_desc(iterations=8, morton_bits=63, quirks=0)
    desc = beatbox.WorldDesc.make(**kw)
    b = beatbox.Batch()
    for bodies in worlds:
        b.add_world(desc, bodies)
    b.step(12)
    b.sync()
    for w, bodies in enumerate(worlds):
        alone = beatbox.RigidBodyManager(desc)
        alone.add_bodies(bodies)
        alone.simulate(12)
        ga = alone.read_bodies()
        gb = b.read_world(w, len(bodies))
        for f in ("position", "rotation", "velocity", "omega", "flags"):
            assert ga[f].tobytes() == gb[f].tobytes(), (w, f)
        ora = OracleWorld(desc)
        ora.add_bodies(bodies)
        ora.step(12)
        o = ora.read_bodies()
        for f in ("position", "rotation", "velocity", "omega", "flags"):
            assert o[f].tobytes() == gb[f].tobytes(), ("oracle", w, f)


def test_batch_mixed_solvers_and_graph_replay():
    sizes = [50, 20, 90, 200]
    worlds = _worlds(0xBB44, sizes)
    descs = [beatbox.WorldDesc.make(**scenes.world_desc(
        solver=scenes.SOLVER_AVBD if w % 2 else scenes.SOLVER_PGS_SOFT,
        iterations=8, morton_bits=63, quirks=0)) for w in range(len(sizes))]
    eager = beatbox.Batch()
    graph = beatbox.Batch()
    for d, bodies in zip(descs, worlds):
        eager.add_world(d, bodies)
        graph.add_world(d, bodies)
    eager.step(10, use_graph=False)
    graph.step(10, use_graph=True)
    eager.sync()
    graph.sync()
    for w, bodies in enumerate(worlds):
        a = eager.read_world(w, len(bodies))
        g = graph.read_world(w, len(bodies))
        for f in ("position", "rotation", "velocity", "omega"):
            assert a[f].tobytes() == g[f].tobytes(), (w, f)
        assert eager.stats(w)["collision_count"] == graph.stats(w)["collision_count"]


def _quat_matrix(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_energy_matches_closed_form():
    """Energy report before any step: potential m*g*y of a pile, and the
    kinetic energy of spinning, moving boxes and spheres evaluated as
    m|v|^2/2 + w.(R I_b R^T).w/2 in matrix form."""
    pile = scenes.batch_world(0xBB04, 2, 40)
    rng = np.random.default_rng(7)
    spin = np.zeros(6, scenes.BODY_DESC)
    for i in range(6):
        b = spin[i]
        b["shape"] = scenes.SHAPE_SPHERE if i % 3 == 2 else scenes.SHAPE_BOX
        b["flags"] = scenes.BODY_DYNAMIC | scenes.BODY_GRAVITY
        b["position"] = (4.0 * i, 3.0 + i, -2.0)
        q = rng.normal(size=4)
        b["rotation"] = q / np.linalg.norm(q)
        half = np.array([0.3, 0.5, 0.9]) * (1 + 0.1 * i)
        b["minimum"] = -half
        b["maximum"] = half
        b["radius"] = 0.4 + 0.05 * i
        b["mass"] = 1.5 + i
        b["velocity"] = rng.normal(size=3)
        b["omega"] = rng.normal(size=3) * 2.0
        b["friction"] = 0.5
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    bt = beatbox.Batch()
    bt.add_world(desc, pile)
    bt.add_world(desc, spin)
    e = bt.energy()
    g = -float(desc.gravity)
    dyn = (pile["flags"] & scenes.BODY_DYNAMIC) != 0
    e_pile = float(np.sum(pile["mass"][dyn].astype(np.float64) * g * pile["position"][dyn, 1]))
    assert abs(e[0] - e_pile) <= 1e-9 * abs(e_pile)
    e_spin = 0.0
    for b in spin:
        m = float(b["mass"])
        if b["shape"] == scenes.SHAPE_SPHERE:
            ib = np.eye(3) * 0.4 * m * float(b["radius"]) ** 2
        else:
            ext = (b["maximum"] - b["minimum"]).astype(np.float64)
            ib = np.diag([m / 12 * (ext[1] ** 2 + ext[2] ** 2), m / 12 * (ext[0] ** 2 + ext[2] ** 2),
                          m / 12 * (ext[0] ** 2 + ext[1] ** 2)])
        r = _quat_matrix(b["rotation"])
        w = b["omega"].astype(np.float64)
        v = b["velocity"].astype(np.float64)
        e_spin += 0.5 * m * v @ v + 0.5 * w @ (r @ ib @ r.T) @ w + m * g * float(b["position"][1])
    assert abs(e[1] - e_spin) <= 1e-6 * abs(e_spin)


def test_profile_counts_equal_world_stats():
    sizes = [30, 200, 64, 500]
    worlds = _worlds(0xBB04, sizes)
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    b = beatbox.Batch()
    for bodies in worlds:
        b.add_world(desc, bodies)
    b.step(5)
    step_ms, kern_ms, counts = b.profile(4, counts=True)
    assert np.all(step_ms > 0) and np.all(kern_ms[[0, 1, 2, 3, 4, 6]] > 0)
    st = [b.stats(w) for w in range(len(sizes))]
    last = counts[-1]
    assert last[0] == sum(s["rigid_body_count"] for s in st)
    assert last[1] == sum(s["active_rigid_body_count"] for s in st)
    assert last[2] == sum(s["broad_phase_collision_count"] for s in st)
    assert last[3] == sum(s["collision_count"] for s in st)
    assert last[4] == sum(s["contact_count"] for s in st)
    assert last[5] == sum(s["graph_color_count"] * s["collision_count"] for s in st)
    assert last[6] == max(s["dbg_pen"] for s in st)
    assert last[7] == sum(s["sleeping_count"] for s in st)


def test_profile_splits_team_time_by_stage():
    """A world stepped by a team runs as one kernel; the profile splits its
    time by stage from the team's own clock marks: every stage of an AVBD
    team world gets time, the first stage keeps only its share, and the
    stages still add up to the steps.  Profiling does not change the
    trajectory."""
    bodies = scenes.serving_pile(46, 3000)
    desc = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    runs = []
    for prof in (True, False):
        b = beatbox.Batch()
        b.add_world(desc, bodies)
        b.set_param(beatbox.PARAM_TEAM_BODIES, 1024)
        b.step(3)
        if prof:
            step_ms, kern_ms = b.profile(4)
            assert np.all(kern_ms[[0, 1, 2, 3, 5, 6]] > 0), kern_ms
            total = float(step_ms.sum())
            assert abs(float(kern_ms.sum()) - total) < 0.05 * total, (kern_ms, total)
            assert kern_ms[0] < 0.6 * total and kern_ms[5] > kern_ms[0], kern_ms
            # the build stage's parts by the team's marks inside it: every part
            # gets time (the lead workgroup's climb too) and they add up to
            # the build stage, which also holds the team kernel's launch and
            # exit (outside the marks; tens of microseconds)
            parts = b.build_split()
            assert parts.shape == (beatbox.BUILD_COUNT,) and np.all(parts > 0), parts
            assert abs(float(parts.sum()) - float(kern_ms[0])) < 0.05 * float(kern_ms[0]) + 0.05, (parts, kern_ms)
            # and the AVBD stage's parts
            av = b.avbd_split()
            assert av.shape == (beatbox.AVBD_COUNT,) and np.all(av > 0), av
            assert abs(float(av.sum()) - float(kern_ms[5])) < 0.05 * float(kern_ms[5]), (av, kern_ms)
        else:
            b.step(4)
        b.sync()
        runs.append(b.read_world(0, len(bodies)))
        b.destroy()
    for f in ("position", "rotation", "velocity", "omega", "flags"):
        assert runs[0][f].tobytes() == runs[1][f].tobytes(), f


def _bodies_equal(a, b, n_worlds, sizes):
    for w in range(n_worlds):
        x = a.read_world(w, sizes[w])
        y = b.read_world(w, sizes[w])
        for f in ("position", "rotation", "velocity", "omega", "flags"):
            assert x[f].tobytes() == y[f].tobytes(), (w, f)


def test_launch_params_do_not_change_results():
    """Block size and the LDS velocity cache are pure launch choices: every
    setting gives bit-identical trajectories (worlds larger and smaller than
    the LDS row count, PGS and AVBD)."""
    sizes = [700, 40, 300, 1200]
    worlds = _worlds(0xBB04, sizes)
    descs = [beatbox.WorldDesc.make(**scenes.world_desc(
        solver=scenes.SOLVER_AVBD if w == 1 else scenes.SOLVER_PGS_SOFT,
        iterations=8, morton_bits=63, quirks=0)) for w in range(len(sizes))]
    runs = []
    # (block, LDS rows, one launch per step, AVBD lanes per body, PGS prefetch,
    # colour-sweep workgroup of the first lane)
    for blk, lds, fused, lanes, pre, pb in ((256, 0, 0, 4, 0, 0), (512, 256, 0, 4, 0, 0), (1024, 768, 0, 4, 1, 0),
                                            (256, 512, 0, 4, 1, 0), (512, 2048, 0, 4, 0, 0),
                                            (256, 512, 1, 4, 1, 0), (512, 256, 1, 2, 0, 0), (256, 0, 0, 1, 1, 0),
                                            (512, 512, 0, 8, 1, 0), (1024, 256, 1, 8, 0, 0),
                                            (256, 128, 0, 4, 0, 64), (256, 256, 0, 4, 1, 128),
                                            (512, 64, 0, 4, 0, 64), (128, 128, 0, 4, 0, 64),
                                            (128, 256, 1, 2, 1, 0), (128, 0, 0, 8, 0, 128)):
        print(f"launch params: block {blk} lds {lds} fused {fused} lanes {lanes} prefetch {pre} pgs block {pb}",
              flush=True)
        b = beatbox.Batch()
        for d, bodies in zip(descs, worlds):
            b.add_world(d, bodies)
        b.set_param(beatbox.PARAM_BLOCK, blk)
        b.set_param(beatbox.PARAM_PGS_LDS_ROWS, lds)
        b.set_param(beatbox.PARAM_FUSED, fused)
        b.set_param(beatbox.PARAM_AVBD_LANES, lanes)
        b.set_param(beatbox.PARAM_PGS_PREFETCH, pre)
        b.set_param(beatbox.PARAM_PGS_BLOCK, pb)
        b.step(15, use_graph=True)
        b.sync()
        assert b.get_param(beatbox.PARAM_BLOCK) == blk
        assert b.get_param(beatbox.PARAM_PGS_LDS_ROWS) == min(lds, max(sizes) + 1)
        assert (b.get_param(beatbox.PARAM_FUSED), b.get_param(beatbox.PARAM_AVBD_LANES),
                b.get_param(beatbox.PARAM_PGS_PREFETCH)) == (fused, lanes, pre)
        runs.append(b)
    for r in runs[1:]:
        _bodies_equal(runs[0], r, len(sizes), [s + 1 for s in sizes])


def test_autotune_leaves_state_unchanged():
    """The autotuner replays the next step from the same state; afterwards
    the batch continues exactly as an untuned batch, and the counters of the
    last real step are restored."""
    sizes = [400, 64, 900, 150, 1200]
    worlds = _worlds(0xBB04, sizes)
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    tuned, plain = beatbox.Batch(), beatbox.Batch()
    # world 1 runs AVBD (the lanes-per-body tile is tuned), world 4 runs as a
    # workgroup team (the sort's sub-team is tuned)
    for b in (tuned, plain):
        b.set_param(beatbox.PARAM_TEAM_BODIES, 1000)
    for w, bodies in enumerate(worlds):
        tuned.add_world(avbd if w == 1 else desc, bodies)
        plain.add_world(avbd if w == 1 else desc, bodies)
    tuned.step(6, use_graph=True)
    plain.step(6, use_graph=True)
    before = [tuned.stats(w) for w in range(len(sizes))]
    ms = tuned.autotune(2)
    assert ms > 0.0
    assert [tuned.stats(w) for w in range(len(sizes))] == before
    tuned.step(7, use_graph=True)
    plain.step(7, use_graph=True)
    tuned.sync()
    plain.sync()
    _bodies_equal(tuned, plain, len(sizes), [s + 1 for s in sizes])
    for w in range(len(sizes)):
        a, b = tuned.stats(w), plain.stats(w)
        a.pop("step_index"), b.pop("step_index")
        assert a == b


def test_team_path_equals_workgroup_path():
    """Worlds stepped by a team of workgroups (k_team: team barriers, team
    scans, the multi-workgroup radix sort) give bit-identical trajectories to
    the one-workgroup path, for PGS_SOFT and AVBD, 30- and 63-bit keys."""
    pgs63 = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    parity = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    worlds = [(pgs63, scenes.batch_world(0xBB04, 7, 3000)),
              (avbd, scenes.serving_pile(21, 2600)),
              (parity, scenes.cubes_grid()),
              (pgs63, scenes.batch_world(0xBB04, 2, 64)),
              (avbd, scenes.serving_pile(22, 700))]
    runs = []
    # 0: every world by one workgroup (the two largest in the 512-thread lane),
    # no broad-phase row handed to a workgroup; 600: four worlds as teams, the
    # radix sort on 1 or (default) 32 of their workgroups, walk budgets of 1
    # node (64 rows per world walked by a workgroup) or the default, the team
    # launch at half or all of the resident workgroups; 2000: two teams
    # running beside both lanes
    for team, sort_blocks, budget, team_blocks in ((0, 0, 0xFFFFFFFF, 0), (600, 1, 1, 0), (600, 0, 0, 2048),
                                                   (2000, 2, 4, 0)):
        print(f"team {team} sort {sort_blocks} walk budget {budget} team blocks {team_blocks}", flush=True)
        b = beatbox.Batch()
        for d, bodies in worlds:
            b.add_world(d, bodies)
        b.set_param(beatbox.PARAM_TEAM_BODIES, team)
        b.set_param(beatbox.PARAM_SORT_BLOCKS, sort_blocks)
        b.set_param(beatbox.PARAM_WALK_BUDGET, budget)
        b.set_param(beatbox.PARAM_TEAM_BLOCKS, team_blocks)
        b.set_param(beatbox.PARAM_PGS_PREFETCH, 1 if sort_blocks == 1 else 0)
        b.step(8, use_graph=True)
        b.sync()
        runs.append(b)
    for w, (d, bodies) in enumerate(worlds):
        x = runs[0].read_world(w, len(bodies))
        sx = runs[0].stats(w)
        for r in runs[1:]:
            y = r.read_world(w, len(bodies))
            for f in ("position", "rotation", "velocity", "omega", "flags", "island_index"):
                assert x[f].tobytes() == y[f].tobytes(), (w, f)
            sy = r.stats(w)
            assert sy["overflow"] == 0, w
            for k in ("broad_phase_collision_count", "collision_count", "contact_count", "island_count",
                      "graph_color_count", "avbd_color_count", "dbg_id_sum"):
                assert sx[k] == sy[k], (w, k)


def test_two_lanes_equal_one_lane():
    """Worlds at or above BB_PARAM_LANE_BODIES run the per-world kernels with
    BB_PARAM_LANE_BLOCK threads on a second stream, concurrently with the
    rest (graph and eager steps); every world's trajectory and counters are
    bit-identical to the one-lane step's."""
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    worlds = [(pgs, scenes.batch_world(0xBB04, w, n)) for w, n in ((1, 1500), (2, 90), (3, 1100), (4, 700))]
    worlds += [(avbd, scenes.serving_pile(41, 1200)), (avbd, scenes.serving_pile(42, 300))]
    runs = []
    for lane, lblock in ((0, 512), (1024, 512), (600, 1024)):
        b = beatbox.Batch()
        for d, bodies in worlds:
            b.add_world(d, bodies)
        b.set_param(beatbox.PARAM_LANE_BODIES, lane)
        b.set_param(beatbox.PARAM_LANE_BLOCK, lblock)
        b.step(6, use_graph=True)
        b.step(3, use_graph=False)
        b.sync()
        runs.append(b)
    for w, (d, bodies) in enumerate(worlds):
        ref = runs[0].read_world(w, len(bodies))
        for r in runs[1:]:
            got = r.read_world(w, len(bodies))
            for f in ("position", "rotation", "velocity", "omega", "flags"):
                assert got[f].tobytes() == ref[f].tobytes(), (w, f)
            assert r.stats(w)["contact_count"] == runs[0].stats(w)["contact_count"]
    for r in runs:
        r.destroy()
