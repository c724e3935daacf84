"""GPU step vs the CPU oracle, bit-exact, through the C-ABI."""
from __future__ import annotations

import numpy as np
import pytest

import beatbox
import scenes
from oracle_bind import OracleWorld
from parity_util import compare_step

pytestmark = pytest.mark.gpu


def run_parity(bodies, desc_kw, steps, check_taps=True):
    desc = beatbox.WorldDesc.make(**desc_kw)
    gpu = beatbox.RigidBodyManager(desc)
    gpu.add_bodies(bodies)
    ora = OracleWorld(desc)
    ora.add_bodies(bodies)
    for s in range(steps):
        gpu.simulate(1)
        ora.step(1)
        bad = compare_step(gpu, ora, desc.morton_bits, s, check_taps)
        assert not bad, "\n".join(bad[:20])
    return gpu, ora


def test_small_pile_pgs_soft_30bit():
    bodies = scenes.random_boxes(0xBB11, 64, extent=3.0, y0=0.55, y1=5.0)
    run_parity(bodies, scenes.world_desc(iterations=8), steps=30)


def test_config1_cubes_grid_first_steps():
    bodies = scenes.cubes_grid()
    run_parity(bodies, scenes.world_desc(iterations=8), steps=8)
