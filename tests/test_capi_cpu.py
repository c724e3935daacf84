"""C-ABI checks that need no GPU: every symbol declared in include/bb_capi.h
is exported, descriptors are validated, capacity is enforced, and stepping
without a device fails loudly (there is no CPU fallback)."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

import beatbox
import scenes

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(REPO, "include", "bb_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bb_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = beatbox.lib()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(beatbox.EXPORTS) <= set(names)


def test_struct_layouts_match_c_compiler(tmp_path):
    """The ctypes/numpy mirrors of the header structs have the sizes and field
    offsets gcc gives them."""
    src = tmp_path / "sizes.c"
    src.write_text('''#include <stdio.h>
#include <stddef.h>
#include "bb_capi.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(bb_world_desc), sizeof(bb_body_desc),
         sizeof(bb_body_state), sizeof(bb_step_stats), sizeof(bb_lbvh_node),
         sizeof(bb_manifold_tap), sizeof(bb_contact_tap));
  printf("%zu %zu %zu\\n", offsetof(bb_body_desc, friction), offsetof(bb_manifold_tap, contacts),
         offsetof(bb_step_stats, step_index));
  printf("%zu %zu %zu\\n", sizeof(bb_server_desc), sizeof(bb_server_report),
         offsetof(bb_server_report, device_ms));
  return 0;
}
''')
    exe = tmp_path / "sizes"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    sizes = [int(x) for x in out]
    assert sizes[0] == C.sizeof(beatbox.WorldDesc)
    assert sizes[1] == scenes.BODY_DESC.itemsize
    assert sizes[2] == beatbox.BODY_STATE.itemsize == C.sizeof(beatbox.BodyState)
    assert sizes[3] == C.sizeof(beatbox.StepStats)
    assert sizes[4] == beatbox.LBVH_NODE.itemsize
    assert sizes[5] == beatbox.MANIFOLD_TAP.itemsize
    assert sizes[6] == beatbox.CONTACT_TAP.itemsize
    assert sizes[7] == scenes.BODY_DESC.fields["friction"][1]
    assert sizes[8] == beatbox.MANIFOLD_TAP.fields["contacts"][1]
    assert sizes[9] == beatbox.StepStats.step_index.offset
    assert sizes[10] == C.sizeof(beatbox.ServerDesc)
    assert sizes[11] == C.sizeof(beatbox.ServerReport)
    assert sizes[12] == beatbox.ServerReport.device_ms.offset


def test_defaults_follow_reference():
    d = beatbox.default_desc()
    assert abs(d.dt - 0.01667) < 1e-9 and abs(d.gravity + 9.81) < 1e-6
    assert d.iterations == 5 and d.solver == 2 and d.morton_bits == 30
    assert d.flags == (8 | 16 | 2 | 1 | 4 | 32 | 128)


def test_invalid_descriptors_rejected():
    w = beatbox.RigidBodyManager(beatbox.WorldDesc.make(**scenes.world_desc()))
    bad = scenes.random_boxes(1, 2)
    bad["maximum"][1] = bad["minimum"][1]  # degenerate box
    with pytest.raises(beatbox.BBError) as e:
        w.add_bodies(bad)
    assert e.value.code == 1
    nan = scenes.random_boxes(1, 2)
    nan["position"][1][0] = np.nan
    with pytest.raises(beatbox.BBError):
        w.add_bodies(nan)
    with pytest.raises(beatbox.BBError):
        beatbox.RigidBodyManager(beatbox.WorldDesc.make(**scenes.world_desc(morton_bits=31)))
    with pytest.raises(beatbox.BBError):
        beatbox.RigidBodyManager(beatbox.WorldDesc.make(**scenes.world_desc(dt=0.0)))


def test_parity_mode_keeps_reference_cap():
    # MAX_RIGID_BODY_COUNT = 1024 (include/defines.hpp:357): build_accel_structs
    # refuses larger scenes; nothing is added
    w = beatbox.RigidBodyManager(beatbox.WorldDesc.make(**scenes.world_desc()))
    ids = w.add_bodies(scenes.random_boxes(2, 1000))
    assert len(ids) == 1001 and ids[0] == 0 and ids[-1] == 1000
    with pytest.raises(beatbox.BBError) as e:
        w.add_bodies(scenes.random_boxes(3, 30, with_ground=False))
    assert e.value.code == 2
    assert beatbox.lib().bb_world_body_count(w._h) == 1001
    # extended mode (63-bit keys, quirks off) has no such cap
    x = beatbox.RigidBodyManager(beatbox.WorldDesc.make(**scenes.world_desc(morton_bits=63, quirks=0)))
    assert len(x.add_bodies(scenes.random_boxes(4, 3000))) == 3001


def test_explicit_capacity():
    w = beatbox.RigidBodyManager(beatbox.WorldDesc.make(**scenes.world_desc(capacity=10)))
    with pytest.raises(beatbox.BBError) as e:
        w.add_bodies(scenes.random_boxes(5, 20))
    assert e.value.code == 2


def test_flags_mirror_reference_setters():
    w = beatbox.RigidBodyManager(beatbox.WorldDesc.make(**scenes.world_desc(flags=0)))
    assert w.set_sim_flags(4 | 8) == 12
    assert w.set_sim_flags(1) == 13
    assert w.clear_sim_flags(4) == 9


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_step_without_gpu_fails_loudly():
    w = beatbox.RigidBodyManager(beatbox.WorldDesc.make(**scenes.world_desc()))
    w.add_bodies(scenes.random_boxes(6, 5))
    with pytest.raises(beatbox.BBError) as e:
        w.simulate(1)
    assert e.value.code == 4  # BB_ERR_NO_DEVICE


def test_server_without_gpu_fails_loudly():
    with pytest.raises(beatbox.BBError) as e:
        beatbox.Server(budget_bytes=1 << 30, max_worlds=8)
    assert e.value.code == 4  # BB_ERR_NO_DEVICE
    with pytest.raises(beatbox.BBError) as e:
        beatbox.Server(budget_bytes=0, max_worlds=8)
    assert e.value.code == 1  # BB_ERR_INVALID


def _world_parts(n=3, p=1, mutate=None):
    """The world descriptor, body descriptors and state arrays of a
    checkpoint of n boxes and p manifold slots, in blob order, with
    consistent indices unless `mutate` changes them."""
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    bodies = scenes.random_boxes(0xBB90, n - 1, extent=2.0, y0=0.55, y1=3.0)
    assert len(bodies) == n
    ids = np.arange(n, dtype=np.uint32)[::-1].copy()
    entries = np.zeros(n, np.uint32)
    entries[ids] = np.arange(n, dtype=np.uint32)
    arrays = dict(id=ids, flags=np.zeros(n, np.uint32), sleep=np.zeros(n, np.uint32), entries=entries,
                  pstart=np.zeros(n, np.uint32), islot=np.zeros(n, np.uint32),
                  rows=np.zeros((5, n, 4), np.float32), pairs=np.array([[0, 1]] * p, np.uint32),
                  geo=np.zeros((5, p, 4), np.float32), hdr=np.array([[0, 1, 5, 4]] * p, np.int32),
                  contacts=np.zeros(p * 8 * (6 * 16 + 4), np.uint8), stats=np.zeros(22, np.uint32))
    if mutate:
        mutate(arrays)
    body = np.ascontiguousarray(bodies, dtype=scenes.BODY_DESC).tobytes()
    order = ("id", "flags", "sleep", "entries", "pstart", "islot", "rows", "pairs", "geo", "hdr", "contacts",
             "stats")
    return bytes(desc), body, b"".join(arrays[k].tobytes() for k in order)


def _blob(n=3, p=1, mutate=None):
    """A checkpoint blob in the layout of bb_world_snapshot (header, body
    descriptors, then the arrays)."""
    import struct
    desc, body, state = _world_parts(n, p, mutate)
    return b"BBSNAP02" + struct.pack("<4IQ", 2, n, p, 1, 0) + desc + struct.pack("<2I", 0, 0) + body + state


def _batch_blob(worlds, pair_scale=1, world_pair_scale=1):
    """A blob in the layout of bb_batch_snapshot: header, no voxel shapes,
    then per world its record (counts, warm flags, pair scale, descriptor),
    body descriptors and arrays; worlds = [(n, p, mutate), ...]."""
    import struct
    out = b"BBBATCH1" + struct.pack("<6IQ", 1, len(worlds), 0, 0, pair_scale, 0, 7)
    for n, p, mutate in worlds:
        desc, body, state = _world_parts(n, p, mutate)
        out += struct.pack("<6I", n, p, 1, world_pair_scale, 0, 0) + desc + body + state
    return out


def _restore_status(blob: bytes) -> int:
    out = C.c_void_p()
    buf = C.create_string_buffer(blob, len(blob))
    rc = beatbox.lib().bb_world_restore(buf, len(blob), 1, C.byref(out))
    if rc == 0:
        beatbox.lib().bb_world_destroy(out)
    return rc


@pytest.mark.parametrize("what", ["ok", "id_range", "id_repeat", "entries", "pair_row", "hdr_row", "hdr_count"])
def test_restore_refuses_blobs_with_bad_indices(what):
    """bb_world_restore checks every index the next step would use to
    address device memory before anything is uploaded: a consistent blob
    gets as far as needing the device (BB_ERR_NO_DEVICE here, no GPU), a
    blob with an id out of range or repeated, an id -> row map that
    disagrees, a pair row or manifold row past the world, or more than 8
    contacts in a slot is refused with BB_ERR_INVALID."""
    def mutate(a):
        if what == "id_range":
            a["id"][1] = 7
        elif what == "id_repeat":
            a["id"][1] = a["id"][0]
        elif what == "entries":
            a["entries"][0] = 1
        elif what == "pair_row":
            a["pairs"][0, 1] = 3
        elif what == "hdr_row":
            a["hdr"][0, 0] = 9
        elif what == "hdr_count":
            a["hdr"][0, 3] = 9
    blob = _blob(mutate=mutate)
    layout_size = 72 + 3 * 100 + 3 * 104 + 1 * 88 + 16 + 8 * (6 * 16 + 4) + 22 * 4
    assert len(blob) == layout_size
    rc = _restore_status(blob)
    if what == "ok":
        assert rc in (0, 4)  # BB_OK with a GPU, BB_ERR_NO_DEVICE without
    else:
        assert rc == 1, what


@pytest.mark.parametrize("what", ["ok", "id_repeat", "hdr_row", "short", "long", "magic", "pair_scale",
                                  "world_pair_scale", "count"])
def test_batch_restore_refuses_bad_blobs_before_the_device(what):
    """bb_batch_restore reads and checks every world record of a batch
    checkpoint on the host: a consistent two-world blob gets as far as needing
    the device (BB_ERR_NO_DEVICE here, no GPU); a bad index in the second
    world, a blob one byte short or long, a wrong magic, a pair scale out of
    range or a world count the blob cannot hold is refused with
    BB_ERR_INVALID."""
    import struct

    def bad_ids(a):
        a["id"][1] = a["id"][0]

    def bad_hdr(a):
        a["hdr"][0, 0] = 9
    second = {"id_repeat": bad_ids, "hdr_row": bad_hdr}.get(what)
    blob = _batch_blob([(3, 1, None), (4, 1, second)], pair_scale=17 if what == "pair_scale" else 1,
                       world_pair_scale=3 if what == "world_pair_scale" else 1)
    if what == "short":
        blob = blob[:-1]
    elif what == "long":
        blob = blob + b"\0"
    elif what == "magic":
        blob = b"BBBATCH0" + blob[8:]
    elif what == "count":
        blob = blob[:12] + struct.pack("<I", 1 << 30) + blob[16:]
    out = C.c_void_p()
    buf = C.create_string_buffer(blob, len(blob))
    rc = beatbox.lib().bb_batch_restore(buf, len(blob), 1, C.byref(out))
    if rc == 0:
        beatbox.lib().bb_batch_destroy(out)
    if what == "ok":
        assert rc in (0, 4)  # BB_OK with a GPU, BB_ERR_NO_DEVICE without
    else:
        assert rc == 1, what


def test_noalloc_limits_checked_before_the_device():
    """bb_world_step_noalloc refuses what its static scratch cannot hold
    (more than NOALLOC_MAX_BODIES bodies: BB_ERR_CAPACITY), voxel bodies and
    empty or short outputs (BB_ERR_INVALID), before it touches a device; a
    valid call gets as far as needing one (BB_ERR_NO_DEVICE here)."""
    desc = beatbox.WorldDesc.make(**scenes.world_desc(morton_bits=63, quirks=0))

    def call(bodies, out_n=None):
        arr, ptr = beatbox._bodies_ptr(bodies)
        out = np.zeros(max(len(arr), 1), beatbox.BODY_STATE)
        n_out = len(arr) if out_n is None else out_n
        return beatbox.lib().bb_world_step_noalloc(C.byref(desc), ptr, len(arr), 3,
                                                   out.ctypes.data_as(C.c_void_p), n_out, None, None)

    big = scenes.random_boxes(7, beatbox.NOALLOC_MAX_BODIES, extent=20.0, with_ground=False)
    assert len(big) == beatbox.NOALLOC_MAX_BODIES
    over = np.concatenate([big, scenes.random_boxes(8, 1, with_ground=False)])
    assert call(over) == 2  # BB_ERR_CAPACITY
    small = scenes.random_boxes(9, 10)
    assert call(small, out_n=len(small) - 1) == 1
    assert call(small[:0]) == 1
    vox = small.copy()
    vox["shape"][3] = scenes.voxel_shape_word(1)
    assert call(vox) == 1
    parity = beatbox.WorldDesc.make(**scenes.world_desc())  # the reference cap of 1,024 bodies
    arr, ptr = beatbox._bodies_ptr(big)
    out = np.zeros(len(arr), beatbox.BODY_STATE)
    assert beatbox.lib().bb_world_step_noalloc(C.byref(parity), ptr, len(arr), 1,
                                               out.ctypes.data_as(C.c_void_p), len(out), None, None) == 2
    if not os.path.exists("/dev/kfd"):
        assert call(big) == 4  # BB_ERR_NO_DEVICE: the maximum itself is accepted


def _server_blob(mutate_state=None, **over):
    """A serving-loop checkpoint in the layout of bb_server_checkpoint: the
    header, one resident world (its record, body descriptors and the state
    arrays of _blob), nothing queued, no results."""
    import struct

    class CkptWorld(C.Structure):
        _fields_ = [("ticket", C.c_uint64), ("desc", beatbox.WorldDesc)] + [
            (k, C.c_uint32) for k in ("n", "active", "steps", "done", "pair_scale", "no_team", "reruns",
                                      "npairs", "resident", "pad")]

    n, p = 3, 1
    snap = _blob(n, p, mutate_state)
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    head_len = 8 + 16 + 8 + len(bytes(desc)) + 8
    bodies = snap[head_len:head_len + n * scenes.BODY_DESC.itemsize]
    state = snap[head_len + n * scenes.BODY_DESC.itemsize:]
    arr = np.frombuffer(bodies, dtype=scenes.BODY_DESC)
    rec = CkptWorld(ticket=7, desc=desc, n=n, active=int((arr["flags"] & scenes.BODY_DYNAMIC > 0).sum()), steps=10, done=4,
                    pair_scale=1, no_team=0, reruns=0, npairs=p, resident=1, pad=0)
    for k, v in over.items():
        setattr(rec, k, v)
    head = b"BBSRVCK2" + struct.pack("<4IQ2I", 2, 1, 0, 0, 8, 0, 0)
    return head + bytes(rec) + bodies + state


def _server_restore_status(blob: bytes) -> int:
    out = C.c_void_p()
    d = beatbox.ServerDesc(1 << 30, 4, 0, 0, 0, 0, 0)
    buf = C.create_string_buffer(blob, len(blob))
    rc = beatbox.lib().bb_server_restore(C.byref(d), buf, len(blob), C.byref(out))
    if rc == 0:
        beatbox.lib().bb_server_destroy(out)
    return rc


def test_server_restore_checks_the_blob_before_the_device():
    """bb_server_restore reads the whole checkpoint on the host first: a
    consistent blob gets as far as making the server (BB_ERR_NO_DEVICE
    here); a wrong magic, a truncated or overlong blob, a world with its
    steps all done, a wrong active count or a bad index in its state arrays
    is BB_ERR_INVALID."""
    good = _server_blob()
    assert _server_restore_status(good) == 4  # BB_ERR_NO_DEVICE
    assert _server_restore_status(b"XXSRVCK1" + good[8:]) == 1
    assert _server_restore_status(good[:-4]) == 1
    assert _server_restore_status(good + b"\0" * 4) == 1
    assert _server_restore_status(_server_blob(done=10)) == 1
    assert _server_restore_status(_server_blob(active=0)) == 1
    assert _server_restore_status(_server_blob(npairs=100000)) == 1

    def repeat(a):
        a["id"][1] = a["id"][0]

    assert _server_restore_status(_server_blob(repeat)) == 1
