"""Voxel (concave) bodies on the GPU against the oracle, bit for bit: the
reference's voxel scene (Ls, crosses and frames, src/scene_manager.hpp:
614-655) with AVBD, and voxel shapes mixed with boxes and spheres under
PGS_SOFT, every step's taps (expanded pair slots, manifolds with keys 90..95,
islands, colours) and counters included; a voxel world in a batch equals the
world alone; bad voxel bodies are refused; a voxel world's checkpoint
continues bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes
from oracle_bind import OracleWorld
from parity_util import compare_step

pytestmark = pytest.mark.gpu


def _pair(desc_kw, shapes, bodies):
    desc = beatbox.WorldDesc.make(**desc_kw)
    gpu = beatbox.RigidBodyManager(desc)
    ora = OracleWorld(desc)
    for _, cells in shapes:
        a = gpu.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
        b = ora.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
        for k in ("shape_id", "solid_count", "surface_count", "mass"):
            assert a[k] == b[k], k
        for k in ("inv_inertia", "minimum", "maximum"):
            assert a[k].tobytes() == b[k].tobytes(), k
    gpu.add_bodies(bodies)
    ora.add_bodies(bodies)
    return desc, gpu, ora


def _parity(desc_kw, shapes, bodies, steps):
    desc, gpu, ora = _pair(desc_kw, shapes, bodies)
    for s in range(steps):
        gpu.simulate(1)
        ora.step(1)
        bad = compare_step(gpu, ora, desc.morton_bits, s)
        assert not bad, "\n".join(bad[:20])
    st = gpu.read_back_sim_config()
    gpu.destroy()
    return st


def test_reference_voxel_scene_avbd_parity():
    shapes, bodies = scenes.voxel_pile()
    kw = scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10, morton_bits=63, quirks=0)
    st = _parity(kw, shapes, bodies, steps=120)
    assert st["collision_count"] > 5


def test_voxel_boxes_spheres_pgs_soft_parity():
    shapes, bodies = scenes.voxel_mixed_pile(0xBB62, 40)
    kw = scenes.world_desc(iterations=8, morton_bits=63, quirks=0)
    st = _parity(kw, shapes, bodies, steps=80)
    assert st["collision_count"] >= 5


def test_voxel_world_in_a_batch_equals_alone():
    shapes, vox = scenes.voxel_mixed_pile(0xBB63, 30)
    desc_v = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                        morton_bits=63, quirks=0))
    desc_p = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    worlds = [(desc_p, scenes.serving_pile(53, 300)), (desc_v, vox), (desc_p, scenes.batch_world(0xBB04, 2, 200))]
    b = beatbox.Batch()
    for _, cells in shapes:
        b.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
    for d, bodies in worlds:
        b.add_world(d, bodies)
    b.step(30, use_graph=True)
    b.sync()
    for k, (d, bodies) in enumerate(worlds):
        w = beatbox.RigidBodyManager(d)
        for _, cells in shapes:
            w.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
        w.add_bodies(bodies)
        w.simulate(30)
        x, y = b.read_world(k, len(bodies)), w.read_bodies()
        for f in ("position", "rotation", "velocity", "omega", "flags"):
            assert x[f].tobytes() == y[f].tobytes(), (k, f)
        w.destroy()
    b.destroy()


def test_voxel_errors():
    w = beatbox.RigidBodyManager(beatbox.WorldDesc.make(**scenes.world_desc(morton_bits=63, quirks=0)))
    with pytest.raises(beatbox.BBError):
        w.add_bodies(scenes.voxel_body(1, (0.0, 1.0, 0.0)))  # no shape yet
    with pytest.raises(beatbox.BBError):
        w.add_voxel_shape(np.zeros((2, 2, 2), bool), 0.5, 2.0)  # empty grid
    info = w.add_voxel_shape(np.ones((2, 2, 2), bool), 0.5, 2.0)
    w.add_bodies(scenes.voxel_body(info["shape_id"], (0.0, 1.0, 0.0)))
    w.simulate(2)
    w.destroy()


def test_voxel_world_checkpoint_continues_bit_identically():
    """The checkpoint of a voxel world carries its shapes: the restored
    world continues exactly (warm start) like the one that wrote it."""
    shapes, bodies = scenes.voxel_pile()
    kw = scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10, morton_bits=63, quirks=0)
    desc = beatbox.WorldDesc.make(**kw)
    a = beatbox.RigidBodyManager(desc)
    for _, cells in shapes:
        a.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
    a.add_bodies(bodies)
    a.simulate(40)
    blob = a.snapshot()
    b = beatbox.RigidBodyManager.restore(blob, warm_start=True)
    a.simulate(30)
    b.simulate(30)
    x, y = a.read_bodies(), b.read_bodies()
    for f in ("position", "rotation", "velocity", "omega", "flags"):
        assert x[f].tobytes() == y[f].tobytes(), f
    assert a.read_back_sim_config()["collision_count"] == b.read_back_sim_config()["collision_count"] > 0
    a.destroy()
    b.destroy()


def test_large_voxel_pile_batch_equals_alone():
    """600 bodies, two thirds of them concave voxel shapes: graph replay in a
    batch next to a box world equals the world stepped alone, with no pair
    overflow."""
    shapes, vox = scenes.voxel_mixed_pile(0xBB64, 600)
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    b = beatbox.Batch()
    for _, cells in shapes:
        b.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
    b.add_world(desc, scenes.batch_world(0xBB04, 5, 400))
    b.add_world(desc, vox)
    b.step(60, use_graph=True)
    b.sync()
    w = beatbox.RigidBodyManager(desc)
    for _, cells in shapes:
        w.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
    w.add_bodies(vox)
    w.simulate(60)
    x, y = b.read_world(1, len(vox)), w.read_bodies()
    for f in ("position", "rotation", "velocity", "omega", "flags"):
        assert x[f].tobytes() == y[f].tobytes(), f
    st = b.stats(1)
    assert st["overflow"] == 0 and st["collision_count"] > 100
    w.destroy()
    b.destroy()


def test_voxel_worlds_as_teams_equal_one_workgroup():
    """The 600-body voxel pile (PGS_SOFT) and the reference's voxel scene
    (AVBD) stepped by workgroup teams (the k_team instance with the voxel
    narrow phase) give bit-identical trajectories and counters to one
    workgroup per world, graph replay included."""
    shapes, vox = scenes.voxel_mixed_pile(0xBB64, 600)
    ref_shapes, ref = scenes.voxel_pile()
    # the reference scene's shape ids name the same shapes in the mixed pile's list
    assert all(np.array_equal(a, c) for (_, a), (_, c) in zip(ref_shapes, shapes))
    pgs = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    worlds = [(pgs, vox), (avbd, ref)]
    runs = []
    for team in (len(ref), 0):
        print(f"voxel teams: team threshold {team}", flush=True)
        b = beatbox.Batch()
        b.set_param(beatbox.PARAM_TEAM_BODIES, team)
        for _, cells in shapes:
            b.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
        for d, bodies in worlds:
            b.add_world(d, bodies)
        b.step(40, use_graph=True)
        b.sync()
        runs.append([(b.read_world(w, len(bodies)), b.stats(w)) for w, (_, bodies) in enumerate(worlds)])
        b.destroy()
    for w, ((x, sx), (y, sy)) in enumerate(zip(*runs)):
        for f in ("position", "rotation", "velocity", "omega", "flags"):
            assert x[f].tobytes() == y[f].tobytes(), (w, f)
        assert sx == sy and sx["overflow"] == 0, w
    assert runs[0][0][1]["collision_count"] > 100 and runs[0][1][1]["collision_count"] > 0
