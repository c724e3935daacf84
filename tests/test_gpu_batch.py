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
