"""The step's HIP kernels, compiled for the host (tests/cpu_emu) with
AddressSanitizer, run on small scenes: every access stays inside its
allocation, and the team path (k_team, several workgroups per world with the
global barrier, team scans and the multi-workgroup radix sort) gives the same
trajectories as the one-workgroup path, as on the GPU.  The emulation runs
every workgroup of a launch on its own host thread and every work-item as a
fiber; it checks indexing and team logic, not the GPU's memory model."""
from __future__ import annotations

import importlib.util
import os
import struct
import subprocess

import numpy as np
import pytest

import beatbox
import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("cpu_emu_build", os.path.join(HERE, "cpu_emu", "build.py"))
emu_build = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(emu_build)

ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1")


@pytest.fixture(scope="module")
def emu_main():
    return emu_build.build()


def _scene(path, worlds, shapes=()):
    """Scene file of the emulation driver; `shapes` are voxel shapes
    (cells, voxel size, density) added to the batch before the worlds."""
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(worlds) | (0x80000000 if shapes else 0)))
        if shapes:
            f.write(struct.pack("<I", len(shapes)))
            for cells, vs, rho in shapes:
                d, words = beatbox.voxel_desc(cells, vs, rho)
                f.write(struct.pack("<3IffI", *cells.shape, vs, rho, len(words)))
                f.write(words.tobytes())
        for d, bodies in worlds:
            raw = bytes(d)
            f.write(struct.pack("<I", len(raw)))
            f.write(raw)
            arr = np.ascontiguousarray(bodies, dtype=scenes.BODY_DESC)
            f.write(struct.pack("<I", len(arr)))
            f.write(arr.tobytes())


def _run(exe, scene, out, steps, team_bodies, block, graph=1, lane_bodies=0, single=0, team_spin=0,
         max_blocks=0, sort_blocks=0, walk_budget=0, pgs_block=0, pgs_lds=512):
    env = dict(ENV, EMU_MAX_BLOCKS=str(max_blocks)) if max_blocks else ENV
    r = subprocess.run([exe, scene, str(steps), str(team_bodies), str(block), str(graph), out, str(lane_bodies),
                        str(single), str(team_spin), str(sort_blocks), str(walk_budget), str(pgs_block),
                        str(pgs_lds)],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    return r.stdout, open(out, "rb").read()


def _world_lines(log: str) -> list:
    """The per-world counter lines of the driver's log (without the lines
    that list collected overflow marks)."""
    return [l for l in log.splitlines() if " world " in l and " overflow world " not in l]


def test_team_and_workgroup_paths_in_bounds_and_identical(emu_main, tmp_path):
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    parity = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    scene = str(tmp_path / "scene.bin")
    _scene(scene, [(avbd, scenes.serving_pile(22, 700)),
                   (pgs, scenes.batch_world(0xBB04, 7, 900)),
                   (parity, scenes.random_boxes(0xBB11, 64, extent=3.0, y0=0.55, y1=5.0))])
    log_w, w = _run(emu_main, scene, str(tmp_path / "w.bin"), 6, 0, 0)
    log_t, t = _run(emu_main, scene, str(tmp_path / "t.bin"), 6, 600, 0)
    assert log_w == log_t
    assert w == t
    # the radix sort on a sub-team of one workgroup (the others wait at the
    # barrier after it): the same order, so the same steps
    log_s, x = _run(emu_main, scene, str(tmp_path / "s.bin"), 6, 600, 0, sort_blocks=1)
    assert log_s == log_w and x == w
    # the step did work: pairs and contacts on the last step of the AVBD pile
    last = [l for l in log_w.splitlines() if l.startswith("step 5 world 0")][0].split()
    assert int(last[5]) > 0 and int(last[9]) > 0


def test_large_workgroups_in_bounds(emu_main, tmp_path):
    """1,024-thread per-world kernels (heavy register pressure on the GPU) on
    an AVBD and a PGS world, without graphs; 128-thread ones the same."""
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    scene = str(tmp_path / "scene.bin")
    _scene(scene, [(avbd, scenes.serving_pile(23, 400)), (pgs, scenes.serving_pile(24, 500))])
    a, x = _run(emu_main, scene, str(tmp_path / "a.bin"), 4, 0, 1024, graph=0)
    b, y = _run(emu_main, scene, str(tmp_path / "b.bin"), 4, 0, 256, graph=0)
    assert a == b and x == y
    # 128-thread workgroups: the radix sort's digit scan takes two digits per thread
    c, z = _run(emu_main, scene, str(tmp_path / "c.bin"), 4, 0, 128, graph=0)
    assert c == a and z == x


def test_emulated_step_equals_oracle(emu_main, tmp_path):
    """The emulated kernels reproduce the oracle bit for bit (poses,
    velocities, flags) on a PGS_SOFT pile, an AVBD pile and a parity-mode
    world stepped together in one batch, so changes to the kernels can be
    checked against the oracle on the host."""
    from oracle_bind import OracleWorld

    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    parity = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    worlds = [(pgs, scenes.serving_pile(31, 300)), (avbd, scenes.serving_pile(32, 250)),
              (parity, scenes.random_boxes(0xBB12, 96, extent=3.0, y0=0.55, y1=5.0))]
    scene = str(tmp_path / "scene.bin")
    _scene(scene, worlds)
    steps = 8
    _, raw = _run(emu_main, scene, str(tmp_path / "o.bin"), steps, 0, 0)
    got = np.frombuffer(raw, dtype=beatbox.BODY_STATE)
    at = 0
    for d, bodies in worlds:
        o = OracleWorld(d)
        o.add_bodies(bodies)
        o.step(steps)
        ref = o.read_bodies()
        mine = got[at:at + len(bodies)]
        at += len(bodies)
        for f in ("position", "rotation", "velocity", "omega", "flags", "id"):
            assert mine[f].tobytes() == ref[f].tobytes(), f
    assert at == len(got)


def test_two_lanes_equal_one_lane(emu_main, tmp_path):
    """Worlds above the lane threshold run the per-world kernels with
    512-thread workgroups on a second stream (captured into the same graph);
    the trajectories equal the one-lane step's."""
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    scene = str(tmp_path / "scene.bin")
    _scene(scene, [(pgs, scenes.batch_world(0xBB04, 3, 60)), (pgs, scenes.batch_world(0xBB04, 8, 400)),
                   (avbd, scenes.serving_pile(33, 350)), (pgs, scenes.serving_pile(34, 120))])
    for graph in (1, 0):
        a, x = _run(emu_main, scene, str(tmp_path / "a.bin"), 5, 0, 0, graph=graph, lane_bodies=0)
        b, y = _run(emu_main, scene, str(tmp_path / "b.bin"), 5, 0, 0, graph=graph, lane_bodies=300)
        assert a == b and x == y, graph
    # the first lane's colour sweeps in one-wave and two-wave workgroups,
    # with fewer LDS rows than the worlds have bodies
    for pgs_block, lds in ((64, 128), (128, 64)):
        c, z = _run(emu_main, scene, str(tmp_path / "c.bin"), 5, 0, 0, lane_bodies=300, pgs_block=pgs_block,
                    pgs_lds=lds)
        assert c == a and z == x, pgs_block


def test_emulated_island_path_equals_oracle(emu_main, tmp_path):
    """USE_GRAPH_COLORING off: the per-island path (plain PGS and PGS_SOFT),
    its dependency-level sweeps and serial overflow reproduce the oracle's
    serial walk bit for bit, next to a colour-sweep world in the same batch."""
    from oracle_bind import OracleWorld

    island = scenes.SIM_DEFAULT_FLAGS & ~scenes.SIM_USE_GRAPH_COLORING
    plain = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_PGS, iterations=8, flags=island,
                                                       morton_bits=63, quirks=0))
    soft = beatbox.WorldDesc.make(**scenes.world_desc(iterations=6, flags=island, morton_bits=63, quirks=0))
    parity = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_PGS, iterations=8, flags=island))
    colour = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    worlds = [(plain, scenes.serving_pile(41, 300)), (soft, scenes.batch_world(0xBB04, 6, 400)),
              (parity, scenes.random_boxes(0xBB13, 96, extent=3.0, y0=0.55, y1=5.0)),
              (colour, scenes.serving_pile(42, 200))]
    scene = str(tmp_path / "scene.bin")
    _scene(scene, worlds)
    steps = 8
    for team in (0, 250):
        _, raw = _run(emu_main, scene, str(tmp_path / "o.bin"), steps, team, 0)
        got = np.frombuffer(raw, dtype=beatbox.BODY_STATE)
        at = 0
        for d, bodies in worlds:
            o = OracleWorld(d)
            o.add_bodies(bodies)
            o.step(steps)
            ref = o.read_bodies()
            mine = got[at:at + len(bodies)]
            at += len(bodies)
            for f in ("position", "rotation", "velocity", "omega", "flags", "id"):
                assert mine[f].tobytes() == ref[f].tobytes(), (team, f)


def test_emulated_voxel_worlds_equal_oracle(emu_main, tmp_path):
    """Voxel bodies (concave L, cross and frame shapes, six pair slots per
    voxel pair, key-exact warm start) with boxes and spheres, PGS_SOFT and
    AVBD, next to a plain world: the emulated kernels reproduce the oracle
    bit for bit, one workgroup per world and as workgroup teams."""
    from oracle_bind import OracleWorld

    shapes, ref = scenes.voxel_pile()
    _, mixed = scenes.voxel_mixed_pile(0xBB61, 24)
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    worlds = [(avbd, ref), (pgs, mixed), (pgs, scenes.serving_pile(35, 150))]
    cells = [(c, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY) for _, c in shapes]
    scene = str(tmp_path / "scene.bin")
    _scene(scene, worlds, cells)
    steps = 40
    log, raw = _run(emu_main, scene, str(tmp_path / "o.bin"), steps, 0, 0)
    got = np.frombuffer(raw, dtype=beatbox.BODY_STATE)
    at = 0
    for d, bodies in worlds:
        o = OracleWorld(d)
        for c, vs, rho in cells:
            o.add_voxel_shape(c, vs, rho)
        o.add_bodies(bodies)
        o.step(steps)
        ref_state = o.read_bodies()
        mine = got[at:at + len(bodies)]
        at += len(bodies)
        for f in ("position", "rotation", "velocity", "omega", "flags", "id"):
            assert mine[f].tobytes() == ref_state[f].tobytes(), f
    # the voxel worlds made contact
    last = [l for l in log.splitlines() if l.startswith(f"step {steps - 1} world 1")][0].split()
    assert int(last[7]) > 0
    # every world as a workgroup team (the k_team instance with the voxel
    # narrow phase): in bounds and the same steps
    log_t, raw_t = _run(emu_main, scene, str(tmp_path / "t.bin"), steps, 12, 0)
    assert _world_lines(log_t) == _world_lines(log) and raw_t == raw
    # they did run as teams: with one workgroup resident at a time no team
    # barrier is met, and each world's step is flagged (overflow bit 1)
    log_x, _ = _run(emu_main, scene, str(tmp_path / "x.bin"), 1, 12, 0, graph=0, team_spin=200, max_blocks=1)
    for w in range(len(worlds)):
        assert f"step 0 overflow world {w}" in log_x, log_x


def test_emulated_pair_overflow_grows_and_replays_exactly(emu_main, tmp_path):
    """125 spheres packed closer than their radius: about 3,400 box pairs
    against the 8n+64 = 1,064 pair slots a world starts with.  Stepped as a
    batch world the overflow is flagged and the surplus pairs dropped; the
    single-world step doubles the capacity and runs the step again from the
    state it read (twice here), so its trajectory equals the oracle's, which
    has no pair capacity, bit for bit, and no step reports an overflow."""
    from oracle_bind import OracleWorld

    bodies = scenes.dense_spheres(0xBB71)
    d = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0, gravity=0.0))
    scene = str(tmp_path / "scene.bin")
    _scene(scene, [(d, bodies)])
    steps = 6
    log_b, raw_b = _run(emu_main, scene, str(tmp_path / "b.bin"), steps, 0, 0, graph=0)
    first = _world_lines(log_b)[0].split()
    assert "step 0 overflow world 0" in log_b
    assert int(first[11]) & 1 and int(first[5]) > 3 * 1064
    log_s, raw_s = _run(emu_main, scene, str(tmp_path / "s.bin"), steps, 0, 0, graph=0, single=1)
    lines = [l.split() for l in log_s.splitlines()]
    assert all(int(l[11]) == 0 for l in lines)
    assert int(lines[0][5]) == int(first[5])
    o = OracleWorld(d)
    o.add_bodies(bodies)
    o.step(steps)
    ref = o.read_bodies()
    got = np.frombuffer(raw_s, dtype=beatbox.BODY_STATE)
    for f in ("position", "rotation", "velocity", "omega", "flags", "id"):
        assert got[f].tobytes() == ref[f].tobytes(), f
    assert raw_b != raw_s


def _team_scene(path):
    """A 700-body AVBD pile (a team of two 512-thread workgroups at a
    600-body threshold) and a small PGS world stepped beside it."""
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    worlds = [(avbd, scenes.serving_pile(22, 700)), (pgs, scenes.serving_pile(25, 150))]
    _scene(path, worlds)
    return worlds


def test_team_that_cannot_meet_leaves_every_step_in_bounds(emu_main, tmp_path):
    """One workgroup resident at a time (EMU_MAX_BLOCKS=1): a team's barrier
    can never be met.  Each waiting workgroup gives up after the spin limit,
    and every member leaves the step at its first barrier after that
    (GroupTeam::abandon) without touching data the team did not finish:
    AddressSanitizer sees every access in bounds, each step flags the team
    world (overflow bit 1, collected as a mark) and the next step's team
    starts from a reset control block.  The other world is stepped as usual
    and equals its run with the team path off."""
    scene = str(tmp_path / "scene.bin")
    worlds = _team_scene(scene)
    log, raw = _run(emu_main, scene, str(tmp_path / "t.bin"), 3, 600, 0, graph=0, team_spin=2000, max_blocks=1)
    for s in range(3):
        assert f"step {s} overflow world 0" in log, log
        assert f"step {s} overflow world 1" not in log, log
        w0 = [l for l in _world_lines(log) if l.startswith(f"step {s} world 0 ")][0].split()
        assert int(w0[11]) & 2
    ref_log, ref = _run(emu_main, scene, str(tmp_path / "r.bin"), 3, 0, 0, graph=0)
    n0 = len(worlds[0][1])
    got = np.frombuffer(raw, dtype=beatbox.BODY_STATE)[n0:]
    want = np.frombuffer(ref, dtype=beatbox.BODY_STATE)[n0:]
    assert got.tobytes() == want.tobytes()
    assert [l for l in _world_lines(log) if " world 1 " in l] == [l for l in _world_lines(ref_log) if " world 1 " in l]


def test_team_barrier_timeouts_at_any_stage_stay_in_bounds(emu_main, tmp_path):
    """Every workgroup resident, but a spin limit so short that waits run
    out now and then at whatever barrier a slow workgroup holds up: the
    team leaves the step at that point, in any stage, and stays in bounds."""
    scene = str(tmp_path / "scene.bin")
    _team_scene(scene)
    _run(emu_main, scene, str(tmp_path / "t.bin"), 4, 600, 0, graph=1, team_spin=40)


def test_single_world_team_failure_replays_by_one_workgroup(emu_main, tmp_path):
    """bb_world_step on a team world whose team cannot meet: the step is
    undone and run again by one workgroup, so the trajectory equals the
    oracle's bit for bit and no step reports an overflow."""
    from oracle_bind import OracleWorld

    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    bodies = scenes.serving_pile(22, 700)
    scene = str(tmp_path / "scene.bin")
    _scene(scene, [(avbd, bodies)])
    steps = 3
    log, raw = _run(emu_main, scene, str(tmp_path / "s.bin"), steps, 600, 0, graph=0, single=1, team_spin=2000,
                    max_blocks=1)
    assert all(int(l.split()[11]) == 0 for l in _world_lines(log))
    o = OracleWorld(avbd)
    o.add_bodies(bodies)
    o.step(steps)
    ref = o.read_bodies()
    got = np.frombuffer(raw, dtype=beatbox.BODY_STATE)
    for f in ("position", "rotation", "velocity", "omega", "flags", "id"):
        assert got[f].tobytes() == ref[f].tobytes(), f


def test_broad_phase_rows_handed_to_the_workgroup_equal_one_thread_walks(emu_main, tmp_path):
    """Walk budgets of 1 and 4 nodes hand most rows (64 per world at most)
    to their workgroup's shared walk (frontier widened in leaf order, one
    walk per frontier node, scanned into place); budget 0xFFFFFFFF hands
    none over.  Per-world kernels and teams, 256- and 1,024-thread
    workgroups: the same pair lists, so the same trajectories, and every
    access in bounds."""
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    parity = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    scene = str(tmp_path / "scene.bin")
    _scene(scene, [(avbd, scenes.serving_pile(22, 700)), (pgs, scenes.serving_pile(26, 300)),
                   (parity, scenes.random_boxes(0xBB14, 64, extent=3.0, y0=0.55, y1=5.0)),
                   (pgs, scenes.dense_spheres(0xBB72, side=4))])
    ref_log, ref = _run(emu_main, scene, str(tmp_path / "r.bin"), 4, 0, 256, walk_budget=0xFFFFFFFF)
    for team, block, budget in ((0, 256, 1), (0, 1024, 4), (600, 256, 1), (600, 0, 4)):
        log, x = _run(emu_main, scene, str(tmp_path / "x.bin"), 4, team, block, walk_budget=budget)
        assert log == ref_log and x == ref, (team, block, budget)
