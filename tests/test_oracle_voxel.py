"""Pins the oracle's voxel shapes and voxel narrow phase (SceneManager::
build_voxel_shape, src/scene_manager.hpp:478-566; collision_detection.slang:
1001-1512) against closed forms and brute force, independent of its code:
mass properties of solid blocks equal the box formulas and of any shape the
parallel-axis sum in float64; the surface list is the set of cells with an
empty face neighbour; a block resting on the ground or on another block
gives one manifold along the contact axis with contacts at the touching cell
centres and depth equal to the overlap; a single cell gives its four face
corners 0.318*vs*sqrt(2) from its centre; a cell buried in a solid block
gives the centre-to-centre rescue contact.  Runs on the CPU."""
from __future__ import annotations

import math

import numpy as np
import pytest

import beatbox
import scenes
from beatbox import MANIFOLD_TAP, TAP_MANIFOLDS, TAP_PAIRS, WorldDesc
from oracle_bind import OracleWorld

VS, RHO = 0.5, 2.0


def _world(solver=scenes.SOLVER_PGS_SOFT, **kw):
    return OracleWorld(WorldDesc.make(**scenes.world_desc(solver=solver, morton_bits=63, quirks=0, **kw)))


def _inertia_f64(cells: np.ndarray, vs: float, rho: float):
    """Mass, centre of mass and inertia about it of unit-density-rho cubes
    (parallel-axis point terms plus each cube's own vs^2/6 term)."""
    idx = np.argwhere(cells).astype(np.float64)
    centres = (idx + 0.5) * vs
    com = centres.mean(axis=0)
    r = centres - com
    mv = rho * vs ** 3
    I = mv * ((r * r).sum() * np.eye(3) - r.T @ r) + len(idx) * mv * vs * vs / 6.0 * np.eye(3)
    return mv * len(idx), com, I


def _surface(cells: np.ndarray) -> int:
    p = np.pad(cells, 1)
    nb = np.zeros_like(cells)
    for ax in range(3):
        for s in (-1, 1):
            nb |= ~np.roll(p, s, axis=ax)[1:-1, 1:-1, 1:-1]
    return int((cells & nb).sum())


@pytest.mark.parametrize("dims", [(3, 2, 4), (1, 1, 1), (5, 5, 5)])
def test_solid_block_mass_properties_equal_box_formula(dims):
    cells = np.ones(dims, bool)
    w = _world()
    info = w.add_voxel_shape(cells, VS, RHO)
    W, H, D = (n * VS for n in dims)
    m = RHO * W * H * D
    assert abs(info["mass"] - m) < 1e-5 * m
    box = m / 12.0 * np.diag([H * H + D * D, W * W + D * D, W * W + H * H])
    assert np.allclose(info["inv_inertia"], np.linalg.inv(box), rtol=2e-5, atol=1e-7)
    assert np.allclose(info["minimum"], [-W / 2, -H / 2, -D / 2], atol=1e-6)
    assert np.allclose(info["maximum"], [W / 2, H / 2, D / 2], atol=1e-6)
    assert info["solid_count"] == cells.sum() and info["surface_count"] == _surface(cells)


@pytest.mark.parametrize("name", ["L", "cross", "frame"])
def test_shape_mass_properties_equal_float64_parallel_axis(name):
    cells = scenes.VOXEL_SHAPES[name]
    w = _world()
    info = w.add_voxel_shape(cells, VS, RHO)
    m, com, I = _inertia_f64(cells, VS, RHO)
    assert abs(info["mass"] - m) < 1e-5 * m
    assert np.allclose(info["minimum"], -com, atol=1e-6)
    assert np.allclose(info["inv_inertia"], np.linalg.inv(I), rtol=1e-4, atol=1e-6)
    assert info["surface_count"] == _surface(cells)
    # the frame has a hole: its surface includes the hole's walls
    if name == "frame":
        assert info["surface_count"] == cells.sum()


def _block_on_ground(dims, y, rot=(0.0, 0.0, 0.0, 1.0)):
    w = _world(iterations=8)
    info = w.add_voxel_shape(np.ones(dims, bool), VS, RHO)
    g = np.zeros(1, scenes.BODY_DESC)
    scenes._ground(g[0])
    w.add_bodies(np.concatenate([g, scenes.voxel_body(info["shape_id"], (0.0, y, 0.0), rot)]))
    w.step(1)
    return w, info


def test_block_resting_on_ground_contacts_at_cell_centres():
    """A 1 m cube of 2x2x2 cells overlapping the ground by d: one manifold on
    the ground's +y face (key 90 + 3), the four bottom cells' centres, depth
    -d, normal along y from obb1 to obb2 (positions before the solve)."""
    d = 0.01
    w, _ = _block_on_ground((2, 2, 2), 0.5 - d)
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    assert len(m) == 1 and m[0]["key"] == 93 and m[0]["contact_count"] == 4
    c = m[0]["contacts"][:4]
    assert np.allclose(c["penetration"], -d, atol=2e-6)
    # one voxel pair takes six pair slots
    assert len(w.tap(TAP_PAIRS, np.uint32).reshape(-1, 2)) == 6
    ids = w.read_bodies(order=beatbox.ORDER_BY_ROW)["id"]
    body_row = int(np.nonzero(ids == 1)[0][0])
    sign = 1.0 if m[0]["obb2_index"] == body_row else -1.0
    assert np.allclose(m[0]["normal"], [0.0, sign, 0.0], atol=1e-6)
    xz = np.sort(np.round(c["position"][:, [0, 2]], 5), axis=0)
    assert np.allclose(np.abs(xz), 0.25, atol=1e-5)
    assert len({(round(a, 4), round(b, 4)) for a, b in c["position"][:, [0, 2]]}) == 4


def test_single_cell_emits_its_four_face_corners():
    d = 0.02
    w, _ = _block_on_ground((1, 1, 1), 0.25 - d)
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    assert len(m) == 1 and m[0]["contact_count"] == 4
    c = m[0]["contacts"][:4]
    r = np.hypot(c["position"][:, 0], c["position"][:, 2])
    assert np.allclose(r, 0.318 * VS * math.sqrt(2.0), rtol=1e-5)
    assert np.allclose(c["penetration"], -d, atol=2e-6)
    # corner index in bits 24-25 of the incident feature ids
    assert sorted((c["fp"][:, 2] >> 24).tolist()) == [0, 1, 2, 3]


def test_stacked_blocks_depth_equals_overlap():
    """Two 1 m voxel cubes, the upper one overlapping the lower by d: the
    voxel-voxel branch (hover band of the cell just outside the grid) gives
    depth -d at the four touching cells."""
    d = 0.01
    w = _world(iterations=8, gravity=0.0)
    info = w.add_voxel_shape(np.ones((2, 2, 2), bool), VS, RHO)
    s = info["shape_id"]
    w.add_bodies(np.concatenate([scenes.voxel_body(s, (0.0, 10.0, 0.0)), scenes.voxel_body(s, (0.0, 11.0 - d, 0.0))]))
    w.step(1)
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    assert len(m) == 1 and m[0]["contact_count"] == 4
    assert m[0]["key"] in (92, 93)
    assert np.allclose(m[0]["contacts"]["penetration"][:4], -d, atol=2e-6)
    assert np.allclose(np.abs(m[0]["normal"]), [0.0, 1.0, 0.0], atol=1e-6)


def test_cell_inside_a_solid_block_gets_the_rescue_contact():
    """One cell buried in the middle of a solid 5x5x5 block: the sampled
    cell's neighbours are all solid, so no pass finds an exit and code 0
    emits the centre-to-centre rescue contact (depth -vs/2, feature ids
    0xFFFFFF, normal +-y for coincident centres)."""
    w = _world(iterations=8, gravity=0.0)
    big = w.add_voxel_shape(np.ones((5, 5, 5), bool), VS, RHO)["shape_id"]
    small = w.add_voxel_shape(np.ones((1, 1, 1), bool), VS, RHO)["shape_id"]
    w.add_bodies(np.concatenate([scenes.voxel_body(big, (0.0, 5.0, 0.0)), scenes.voxel_body(small, (0.0, 5.0, 0.0))]))
    w.step(1)
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    assert len(m) == 1 and m[0]["key"] == 90 and m[0]["contact_count"] == 1
    c = m[0]["contacts"][0]
    assert abs(c["penetration"] + 0.5 * VS) < 1e-7 and (c["fp"] == 0xFFFFFF).all()
    assert np.allclose(c["position"], [0.0, 5.0, 0.0], atol=1e-6)
    assert np.allclose(np.abs(m[0]["normal"]), [0.0, 1.0, 0.0])


def test_resting_block_warm_starts_and_balances_gravity():
    """Key-exact warm start: a resting voxel cube keeps its manifold's
    contacts matched step after step, and the accumulated normal impulse
    carries the weight.  RigidBody::advect (extensions.slang:127-139) adds
    the gravity impulse g*dt times inv_mass to the velocity, so the impulse
    that holds a body of any mass is g*dt."""
    w = _world(iterations=8, flags=scenes.SIM_DEFAULT_FLAGS & ~128)
    info = w.add_voxel_shape(np.ones((2, 2, 2), bool), VS, RHO)
    g = np.zeros(1, scenes.BODY_DESC)
    scenes._ground(g[0])
    w.add_bodies(np.concatenate([g, scenes.voxel_body(info["shape_id"], (0.0, 0.5, 0.0))]))
    w.step(120)
    m = w.tap(TAP_MANIFOLDS, MANIFOLD_TAP)
    assert len(m) == 1
    c = m[0]["contacts"][:m[0]["contact_count"]]
    assert (c["warm_start"] > 100).all()
    dt = float(np.float32(0.01667))
    assert info["mass"] == 2.0
    assert abs(c["normal_impulse"].sum() - 9.81 * dt) < 0.02 * 9.81 * dt


def test_reference_voxel_scene_settles_with_avbd():
    shapes, bodies = scenes.voxel_pile()
    w = _world(solver=scenes.SOLVER_AVBD, iterations=10)
    infos = [w.add_voxel_shape(c, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY) for _, c in shapes]
    w.add_bodies(bodies)
    w.step(300)
    r = w.read_bodies()
    vox = r[3:]
    assert (vox["flags"] & 8).all()  # all asleep
    # every piece rests at least half a cell above the floor and none sank
    assert (vox["position"][:, 1] > 0.45).all()
    assert infos[2]["solid_count"] == 48 * 2


def test_voxel_body_validation():
    w = _world()
    with pytest.raises(ValueError):
        w.add_bodies(scenes.voxel_body(1, (0.0, 1.0, 0.0)))  # no shape 1 yet
    with pytest.raises(ValueError):
        w.add_voxel_shape(np.zeros((2, 2, 2), bool), VS, RHO)  # empty grid
