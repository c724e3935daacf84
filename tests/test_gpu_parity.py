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


def test_small_pile_avbd_30bit():
    bodies = scenes.random_boxes(0xBB12, 64, extent=3.0, y0=0.55, y1=5.0)
    run_parity(bodies, scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=5), steps=30)


def test_config1_avbd_first_steps():
    bodies = scenes.cubes_grid()
    run_parity(bodies, scenes.world_desc(solver=scenes.SOLVER_AVBD), steps=6)


def test_extended_mode_63bit_no_quirks():
    bodies = scenes.random_boxes(0xBB13, 96, extent=4.0, y0=0.55, y1=8.0)
    kw = scenes.world_desc(iterations=8, morton_bits=63, quirks=0)
    run_parity(bodies, kw, steps=25)


def test_spheres_and_boxes_avbd_63bit():
    bodies = scenes.random_boxes(0xBB14, 80, extent=3.0, y0=0.6, y1=6.0)
    sph = bodies[1::2]
    sph["shape"] = scenes.SHAPE_SPHERE
    sph["radius"] = 0.5
    bodies[1::2] = sph
    kw = scenes.world_desc(solver=scenes.SOLVER_AVBD, morton_bits=63, quirks=0)
    run_parity(bodies, kw, steps=25)


def test_sleep_disabled_and_warm_start_off():
    bodies = scenes.random_boxes(0xBB15, 40, extent=2.0, y0=0.55, y1=3.0)
    flags = scenes.SIM_DEFAULT_FLAGS & ~(4 | 128)
    run_parity(bodies, scenes.world_desc(iterations=4, flags=flags), steps=20)
