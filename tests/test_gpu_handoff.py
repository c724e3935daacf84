"""Render hand-off: per body id, the 3x4 affine transform [R | t] of its
pose (the TLAS instance layout the reference's AS_mngr.slang writes).  The
translation column is the body's position bit for bit; R is the rotation
matrix of the body's quaternion (closed form in float64, within float32
rounding); the device-pointer path writes the same bytes as the host path,
and a world of a batch gives the same transforms as the world alone."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import beatbox
import scenes

pytestmark = pytest.mark.gpu


def hip_runtime() -> C.CDLL:
    """The HIP runtime the library is linked against, for a device buffer."""
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def rotation_f64(q: np.ndarray) -> np.ndarray:
    """Rotation matrices of unit quaternions (x, y, z, w), float64."""
    x, y, z, w = (q[:, k].astype(np.float64) for k in range(4))
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1),
    ], 1)


def test_instance_transforms_match_poses():
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    bodies = scenes.serving_pile(51, 400)
    w = beatbox.RigidBodyManager(desc)
    w.add_bodies(bodies)
    w.simulate(25)
    n = len(bodies)
    t = w.instance_transforms(n)
    st = w.read_bodies()
    assert t[:, :, 3].tobytes() == np.ascontiguousarray(st["position"]).tobytes()
    rot = rotation_f64(st["rotation"])
    assert np.abs(t[:, :, :3] - rot).max() < 4e-6
    # the bodies moved and turned: the check is not of the initial poses
    assert np.abs(rot - np.eye(3)).max() > 0.1
    # the device-pointer path writes the same bytes into a device buffer
    hip = hip_runtime()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), n * 48) == 0
    try:
        assert w.instance_transforms(n, device_ptr=dev.value) is None
        w.sync()
        back = np.zeros((n, 3, 4), np.float32)
        assert hip.hipMemcpy(back.ctypes.data_as(C.c_void_p), dev, n * 48, 2) == 0  # device to host
    finally:
        hip.hipFree(dev)
    assert back.tobytes() == t.tobytes()
    with pytest.raises(beatbox.BBError):
        w.instance_transforms(n - 1)
    w.destroy()


def test_batch_instance_transforms_equal_alone():
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    worlds = [scenes.batch_world(0xBB04, 5, 300), scenes.serving_pile(52, 250)]
    b = beatbox.Batch()
    for bodies in worlds:
        b.add_world(desc, bodies)
    b.step(12, use_graph=True)
    b.sync()
    for k, bodies in enumerate(worlds):
        w = beatbox.RigidBodyManager(desc)
        w.add_bodies(bodies)
        w.simulate(12)
        assert b.instance_transforms(k, len(bodies)).tobytes() == w.instance_transforms(len(bodies)).tobytes()
        w.destroy()
    b.destroy()


def test_lbvh_export_matches_tap_and_oracle():
    """The BVH hand-off (get_lbvh_node_buffer): the 2N-1 records in the
    reference's LBVHNode layout equal the debug tap and the oracle's BVH bit
    for bit, on the host path and, byte for byte, on the device path; a
    world of a batch exports the same nodes as the world alone."""
    from oracle_bind import OracleWorld

    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    bodies = scenes.random_boxes(0xBB53, 200, extent=4.0, y0=0.55, y1=6.0)
    w = beatbox.RigidBodyManager(desc)
    w.add_bodies(bodies)
    assert len(w.lbvh_nodes()) == 0  # nothing built before the first step
    ora = OracleWorld(desc)
    ora.add_bodies(bodies)
    w.simulate(15)
    ora.step(15)
    nodes = w.lbvh_nodes()
    n = len(bodies)
    assert len(nodes) == 2 * n - 1
    assert nodes.tobytes() == w.tap(beatbox.TAP_LBVH, beatbox.LBVH_NODE).tobytes()
    assert nodes.tobytes() == ora.tap(beatbox.TAP_LBVH, beatbox.LBVH_NODE).tobytes()
    leaves = nodes["left"] < 0
    assert leaves.sum() == n and np.array_equal(np.sort(nodes["primitive_idx"][leaves]), np.arange(n))
    assert np.all(nodes["primitive_idx"][~leaves] == 0xFFFFFFFF)
    hip = hip_runtime()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), len(nodes) * 40) == 0
    try:
        with pytest.raises(beatbox.BBError):
            w.lbvh_nodes(device_ptr=dev.value, cap=len(nodes) - 1)
        assert w.lbvh_nodes(device_ptr=dev.value, cap=len(nodes)) == len(nodes)
        w.sync()
        back = np.zeros(len(nodes), beatbox.LBVH_NODE)
        assert hip.hipMemcpy(back.ctypes.data_as(C.c_void_p), dev, len(nodes) * 40, 2) == 0
    finally:
        hip.hipFree(dev)
    assert back.tobytes() == nodes.tobytes()
    b = beatbox.Batch()
    b.add_world(desc, scenes.random_boxes(0xBB54, 30, extent=2.0, y0=0.55, y1=3.0))
    b.add_world(desc, bodies)
    b.step(15)
    b.sync()
    cnt = C.c_uint32(0)
    got = np.zeros(len(nodes), beatbox.LBVH_NODE)
    assert beatbox.lib().bb_batch_lbvh_nodes(b._h, 1, got.ctypes.data_as(C.c_void_p), len(got), C.byref(cnt), 0) == 0
    assert cnt.value == len(nodes) and got.tobytes() == nodes.tobytes()
    b.destroy()
    w.destroy()
