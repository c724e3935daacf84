"""Python host binding of the MI355X rigid-body step (ctypes over include/bb_capi.h).

`RigidBodyManager` mirrors the reference's C++ manager
(src/rigid_body_manager.hpp:14-116): create(iterations), simulate(),
read_back_sim_config(), set/clear/get_sim_flags, set/get_sim_type,
skip_warm_starting_once(); bodies are added with the load_scene rules.
`Batch` steps many worlds together (no reference counterpart).

The library is the HIP build `libbeatbox_amd.so` next to this file.  There is
no CPU fallback: loading fails loudly when the library is missing, and
stepping fails with BB_ERR_NO_DEVICE when no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbeatbox_amd.so")

# ------------------------------------------------------------------ structs
class WorldDesc(C.Structure):
    _fields_ = [
        ("capacity_bodies", C.c_uint32),
        ("solver", C.c_uint32),
        ("iterations", C.c_uint32),
        ("flags", C.c_uint32),
        ("dt", C.c_float),
        ("gravity", C.c_float),
        ("morton_bits", C.c_uint32),
        ("quirks", C.c_uint32),
    ]

    @classmethod
    def make(cls, **kw) -> "WorldDesc":
        d = cls()
        for k, v in kw.items():
            setattr(d, k, v)
        return d


class BodyState(C.Structure):
    _fields_ = [
        ("id", C.c_uint32),
        ("flags", C.c_uint32),
        ("island_index", C.c_uint32),
        ("row", C.c_uint32),
        ("position", C.c_float * 3),
        ("rotation", C.c_float * 4),
        ("velocity", C.c_float * 3),
        ("omega", C.c_float * 3),
    ]


BODY_STATE = np.dtype(
    [("id", "<u4"), ("flags", "<u4"), ("island_index", "<u4"), ("row", "<u4"),
     ("position", "<f4", 3), ("rotation", "<f4", 4), ("velocity", "<f4", 3),
     ("omega", "<f4", 3)]
)
assert BODY_STATE.itemsize == C.sizeof(BodyState)


class StepStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "rigid_body_count", "active_rigid_body_count", "broad_phase_collision_count",
        "collision_count", "manifold_node_count", "contact_count", "island_count",
        "contact_island_count", "sleeping_count", "graph_color_count",
        "graph_color_violations", "graph_color_overflow", "avbd_color_count",
        "avbd_violations", "avbd_stick_count", "dbg_fresh", "dbg_pen", "dbg_id_sum",
        "dbg_maxv", "overflow")] + [("step_index", C.c_uint64)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class ServerDesc(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("max_worlds", C.c_uint32),
        ("max_bodies", C.c_uint32),
        ("max_resident_bodies", C.c_uint32),
        ("block", C.c_uint32),
        ("team_bodies", C.c_uint32),
        ("split_bodies", C.c_uint32),
    ]


SERVER_MAX_POOLS = 4


class ServerReport(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "submitted", "admitted", "retired", "queued", "resident_worlds", "resident_bodies",
        "bytes_in_use", "bytes_budget", "bytes_allocated", "server_steps", "world_steps",
        "body_steps")] + [(n, C.c_double) for n in (
        "device_ms", "step_ms_p50", "step_ms_p99", "queue_ms_p50", "queue_ms_p99")] + [
        ("reruns", C.c_uint64), ("pools", C.c_uint32), ("pad", C.c_uint32),
        ("pool_min_bodies", C.c_uint32 * SERVER_MAX_POOLS),
        ("pool_world_steps", C.c_uint64 * SERVER_MAX_POOLS), ("pool_step_ms_p50", C.c_double * SERVER_MAX_POOLS),
        ("pool_step_ms_p99", C.c_double * SERVER_MAX_POOLS)]

    def as_dict(self) -> dict:
        out = {n: getattr(self, n) for n, _ in self._fields_ if not n.startswith("pool") and "pad" not in n}
        out["pools"] = [dict(min_bodies=self.pool_min_bodies[p], world_steps=self.pool_world_steps[p],
                             step_ms_p50=self.pool_step_ms_p50[p], step_ms_p99=self.pool_step_ms_p99[p])
                        for p in range(self.pools)]
        return out


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_UNKNOWN, REQ_FAILED = 0, 1, 2, 3, 4


class VoxelShapeDesc(C.Structure):
    _fields_ = [("dims", C.c_uint32 * 3), ("voxel_size", C.c_float), ("density", C.c_float),
                ("occupancy", C.POINTER(C.c_uint32))]


class VoxelShapeInfo(C.Structure):
    _fields_ = [("shape_id", C.c_uint32), ("solid_count", C.c_uint32), ("surface_count", C.c_uint32),
                ("mass", C.c_float), ("inv_inertia", C.c_float * 9), ("minimum", C.c_float * 3),
                ("maximum", C.c_float * 3)]

    def as_dict(self) -> dict:
        return {"shape_id": self.shape_id, "solid_count": self.solid_count,
                "surface_count": self.surface_count, "mass": self.mass,
                "inv_inertia": np.array(self.inv_inertia, np.float32).reshape(3, 3),
                "minimum": np.array(self.minimum, np.float32), "maximum": np.array(self.maximum, np.float32)}


SHAPE_BOX, SHAPE_SPHERE, SHAPE_VOXEL = 0, 1, 2


def voxel_shape_word(shape_id: int) -> int:
    """The body shape word of voxel shape `shape_id` (BB_SHAPE_VOXEL_OF)."""
    return SHAPE_VOXEL | (int(shape_id) << 8)


def voxel_desc(solid: np.ndarray, voxel_size: float, density: float):
    """A VoxelShapeDesc from a boolean grid solid[x, y, z]; returns the
    descriptor and the occupancy words it points at (keep both alive)."""
    solid = np.asarray(solid, bool)
    dims = solid.shape
    bits = solid.flatten(order="F")  # x fastest: bit x + y*dx + z*dx*dy
    pad = (-len(bits)) % 32
    bits = np.concatenate([bits, np.zeros(pad, bool)])
    words = np.ascontiguousarray(np.packbits(bits, bitorder="little").view("<u4"))
    d = VoxelShapeDesc()
    d.dims[:] = [int(x) for x in dims]
    d.voxel_size = voxel_size
    d.density = density
    d.occupancy = words.ctypes.data_as(C.POINTER(C.c_uint32))
    return d, words

LBVH_NODE = np.dtype([("minimum", "<f4", 3), ("maximum", "<f4", 3), ("left", "<i4"),
                      ("right", "<i4"), ("primitive_idx", "<u4"), ("pad", "<u4")])
CONTACT_TAP = np.dtype([
    ("position", "<f4", 3), ("penetration", "<f4"), ("normal_impulse", "<f4"),
    ("tangent_impulse", "<f4", 2), ("mass_coefficient", "<f4"), ("tangent_mass", "<f4", 2),
    ("anchor_a", "<f4", 3), ("anchor_b", "<f4", 3), ("stick", "<u4"), ("fp", "<u4", 4),
    ("warm_start", "<u4")])
MANIFOLD_TAP = np.dtype([
    ("obb1_index", "<u4"), ("obb2_index", "<u4"), ("key", "<i4"), ("error", "<i4"),
    ("contact_count", "<i4"), ("color", "<u4"), ("normal", "<f4", 3),
    ("tangent", "<f4", (2, 3)), ("anchor_a", "<f4", 3), ("anchor_b", "<f4", 3),
    ("contacts", CONTACT_TAP, 8)])

TAP_MORTON, TAP_SORTED, TAP_LBVH, TAP_PAIRS, TAP_MANIFOLDS = 0, 1, 2, 3, 4
TAP_ISLANDS, TAP_EDGE_COLORS, TAP_BODY_COLORS, TAP_ROWS = 5, 6, 7, 8
ORDER_BY_ID, ORDER_BY_ROW = 0, 1
KERNEL_NAMES = ("k_build", "k_broad", "k_narrow", "k_post", "k_pgs", "k_avbd", "k_finish")
KERNEL_COUNT = len(KERNEL_NAMES)
BUILD_NAMES = ("keys", "sort", "hierarchy", "bounds", "climb", "reorder")
BUILD_COUNT = len(BUILD_NAMES)
AVBD_NAMES = ("colour", "group", "sweeps", "post")
AVBD_COUNT = len(AVBD_NAMES)
COUNT_NAMES = ("bodies", "active", "pairs", "manifolds", "contacts", "color_work", "max_pen_mm", "sleeping")
COUNT_FIELDS = len(COUNT_NAMES)
PARAM_BLOCK, PARAM_PGS_LDS_ROWS, PARAM_TEAM_BODIES, PARAM_LANE_BODIES, PARAM_LANE_BLOCK, PARAM_LANE_LDS_ROWS = \
    0, 1, 2, 3, 4, 5
PARAM_TEAM_SPIN = 6
PARAM_PAIR_SCALE = 7
PARAM_SORT_BLOCKS = 8
PARAM_FUSED = 9
PARAM_AVBD_LANES = 10
PARAM_PGS_PREFETCH = 11
PARAM_WALK_BUDGET = 12
PARAM_TEAM_BLOCKS = 13
PARAM_PGS_BLOCK = 14

STATUS = {0: "BB_OK", 1: "BB_ERR_INVALID", 2: "BB_ERR_CAPACITY", 3: "BB_ERR_DEVICE",
          4: "BB_ERR_NO_DEVICE", 5: "BB_ERR_OVERFLOW", 6: "BB_ERR_STATE", 7: "BB_ERR_BUDGET"}


class BBError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {STATUS.get(code, code)}")
        self.code = code


def _check(code: int, what: str) -> None:
    if code != 0:
        raise BBError(code, what)


# ----------------------------------------------------------------- loading
_LIB = None

EXPORTS = {
    # name: (restype, argtypes)
    "bb_world_desc_default": (None, [C.POINTER(WorldDesc)]),
    "bb_world_create": (C.c_int, [C.POINTER(WorldDesc), C.POINTER(C.c_void_p)]),
    "bb_world_add_bodies": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "bb_world_set_flags": (C.c_int, [C.c_void_p, C.c_uint32]),
    "bb_world_add_voxel_shape": (C.c_int, [C.c_void_p, C.POINTER(VoxelShapeDesc), C.POINTER(VoxelShapeInfo)]),
    "bb_batch_add_voxel_shape": (C.c_int, [C.c_void_p, C.POINTER(VoxelShapeDesc), C.POINTER(VoxelShapeInfo)]),
    "bb_world_clear_flags": (C.c_int, [C.c_void_p, C.c_uint32]),
    "bb_world_get_flags": (C.c_uint32, [C.c_void_p]),
    "bb_world_set_solver": (C.c_int, [C.c_void_p, C.c_uint32]),
    "bb_world_set_iterations": (C.c_int, [C.c_void_p, C.c_uint32]),
    "bb_world_skip_warm_start_once": (C.c_int, [C.c_void_p]),
    "bb_world_step": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "bb_world_sync": (C.c_int, [C.c_void_p]),
    "bb_world_set_param": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32]),
    "bb_world_read_bodies": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int]),
    "bb_world_read_stats": (C.c_int, [C.c_void_p, C.POINTER(StepStats)]),
    "bb_world_debug_tap": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t,
                                     C.POINTER(C.c_size_t)]),
    "bb_world_body_count": (C.c_uint32, [C.c_void_p]),
    "bb_world_destroy": (None, [C.c_void_p]),
    "bb_batch_create": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "bb_batch_add_world": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p,
                                     C.c_uint32, C.POINTER(C.c_uint32)]),
    "bb_batch_step": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]),
    "bb_batch_sync": (C.c_int, [C.c_void_p]),
    "bb_batch_overflow": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "bb_batch_read_world": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_int]),
    "bb_batch_world_stats": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(StepStats)]),
    "bb_batch_debug_tap": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_size_t,
                                     C.POINTER(C.c_size_t)]),
    "bb_batch_world_count": (C.c_uint32, [C.c_void_p]),
    "bb_batch_body_count": (C.c_uint32, [C.c_void_p]),
    "bb_batch_set_block_size": (C.c_int, [C.c_void_p, C.c_uint32]),
    "bb_batch_set_param": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32]),
    "bb_batch_get_param": (C.c_uint32, [C.c_void_p, C.c_int]),
    "bb_batch_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float)]),
    "bb_batch_profile": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bb_batch_energy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "bb_batch_build_split": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bb_batch_avbd_split": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bb_world_energy": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "bb_world_instance_transforms": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int]),
    "bb_world_lbvh_nodes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_int]),
    "bb_batch_lbvh_nodes": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                      C.c_int]),
    "bb_world_snapshot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "bb_world_restore": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    "bb_batch_instance_transforms": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_int]),
    "bb_batch_device_bytes": (C.c_uint64, [C.c_void_p]),
    "bb_batch_destroy": (None, [C.c_void_p]),
    "bb_batch_snapshot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "bb_batch_restore": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    "bb_server_create": (C.c_int, [C.POINTER(ServerDesc), C.POINTER(C.c_void_p)]),
    "bb_server_submit": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p, C.c_uint32, C.c_uint32,
                                   C.POINTER(C.c_uint64)]),
    "bb_server_step": (C.c_int, [C.c_void_p, C.c_uint32]),
    "bb_server_drain": (C.c_int, [C.c_void_p]),
    "bb_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32), C.c_void_p, C.c_uint32,
                                 C.POINTER(StepStats)]),
    "bb_server_report_get": (C.c_int, [C.c_void_p, C.POINTER(ServerReport)]),
    "bb_server_checkpoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "bb_server_restore": (C.c_int, [C.POINTER(ServerDesc), C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "bb_server_add_voxel_shape": (C.c_int, [C.c_void_p, C.POINTER(VoxelShapeDesc), C.POINTER(VoxelShapeInfo)]),
    "bb_server_destroy": (None, [C.c_void_p]),
    "bb_version": (C.c_char_p, []),
    "bb_copy_bandwidth": (C.c_int, [C.c_uint64, C.c_uint32, C.POINTER(C.c_float)]),
    "bb_world_step_noalloc": (C.c_int, [C.POINTER(WorldDesc), C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                        C.c_uint32, C.POINTER(StepStats), C.c_void_p]),
}

NOALLOC_MAX_BODIES = 2048
NOALLOC_SCRATCH_BYTES = 64 << 20


def lib() -> C.CDLL:
    """Loads the HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python __graft_entry__.py build` (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def copy_bandwidth(nbytes: int, reps: int = 5) -> float:
    """Mean duration (ms) of a device-to-device float4 copy of nbytes."""
    ms = C.c_float(0.0)
    _check(lib().bb_copy_bandwidth(nbytes, reps, C.byref(ms)), "bb_copy_bandwidth")
    return ms.value


def default_desc() -> WorldDesc:
    d = WorldDesc()
    lib().bb_world_desc_default(C.byref(d))
    return d


def _bodies_ptr(bodies: np.ndarray):
    from scenes import BODY_DESC  # local module next to this file
    arr = np.ascontiguousarray(bodies, dtype=BODY_DESC)
    return arr, arr.ctypes.data_as(C.c_void_p)


def step_noalloc(desc: WorldDesc, bodies: np.ndarray, nsteps: int) -> tuple[int, np.ndarray, StepStats]:
    """bb_world_step_noalloc: a fresh world of `bodies` stepped nsteps times
    in the library's static scratch (no device allocation).  Returns the
    status code (BB_OK, or BB_ERR_OVERFLOW with results), the bodies by id
    and the last step's counters; raises on any other status."""
    arr, ptr = _bodies_ptr(bodies)
    out = np.zeros(len(arr), BODY_STATE)
    stats = StepStats()
    rc = lib().bb_world_step_noalloc(C.byref(desc), ptr, len(arr), nsteps, out.ctypes.data_as(C.c_void_p),
                                     len(out), C.byref(stats), None)
    if rc not in (0, 5):
        _check(rc, "bb_world_step_noalloc")
    return rc, out, stats


# ------------------------------------------------------------ world object
class RigidBodyManager:
    """Mirror of the reference RigidBodyManager over the C-ABI."""

    def __init__(self, desc: WorldDesc | None = None, **kw):
        self._h = C.c_void_p()
        d = desc if desc is not None else default_desc()
        for k, v in kw.items():
            setattr(d, k, v)
        _check(lib().bb_world_create(C.byref(d), C.byref(self._h)), "bb_world_create")

    # SceneManager::load_scene rules
    def add_bodies(self, bodies: np.ndarray) -> np.ndarray:
        arr, ptr = _bodies_ptr(bodies)
        ids = np.zeros(len(arr), np.uint32)
        _check(lib().bb_world_add_bodies(self._h, ptr, len(arr), ids.ctypes.data_as(C.c_void_p)),
               "bb_world_add_bodies")
        return ids

    # AccelerationStructureManager::build_accel_structs, the reference's name
    # for the upload that load_scene ends with
    build_accel_structs = add_bodies

    def add_voxel_shape(self, solid: np.ndarray, voxel_size: float, density: float) -> dict:
        """SceneManager::build_voxel_shape: solid[x, y, z] cells; returns the
        shape id (bodies use shape = voxel_shape_word(id)) and mass data."""
        d, words = voxel_desc(solid, voxel_size, density)
        info = VoxelShapeInfo()
        _check(lib().bb_world_add_voxel_shape(self._h, C.byref(d), C.byref(info)), "bb_world_add_voxel_shape")
        del words
        return info.as_dict()

    def simulate(self, nsteps: int = 1) -> None:
        _check(lib().bb_world_step(self._h, nsteps, None), "bb_world_step")

    def sync(self) -> None:
        _check(lib().bb_world_sync(self._h), "bb_world_sync")

    def read_bodies(self, order: int = ORDER_BY_ID) -> np.ndarray:
        n = lib().bb_world_body_count(self._h)
        out = np.zeros(n, BODY_STATE)
        _check(lib().bb_world_read_bodies(self._h, out.ctypes.data_as(C.c_void_p), n, order),
               "bb_world_read_bodies")
        return out

    def read_back_sim_config(self) -> dict:
        s = StepStats()
        _check(lib().bb_world_read_stats(self._h, C.byref(s)), "bb_world_read_stats")
        return s.as_dict()

    def set_sim_flags(self, flags: int) -> int:
        _check(lib().bb_world_set_flags(self._h, flags), "bb_world_set_flags")
        return self.get_sim_flags()

    def clear_sim_flags(self, flags: int) -> int:
        _check(lib().bb_world_clear_flags(self._h, flags), "bb_world_clear_flags")
        return self.get_sim_flags()

    def get_sim_flags(self) -> int:
        return lib().bb_world_get_flags(self._h)

    def set_sim_type(self, solver: int) -> None:
        _check(lib().bb_world_set_solver(self._h, solver), "bb_world_set_solver")

    def set_param(self, p: int, value: int) -> None:
        _check(lib().bb_world_set_param(self._h, p, value), "bb_world_set_param")

    def set_iterations(self, it: int) -> None:
        _check(lib().bb_world_set_iterations(self._h, it), "bb_world_set_iterations")

    def skip_warm_starting_once(self) -> None:
        _check(lib().bb_world_skip_warm_start_once(self._h), "bb_world_skip_warm_start_once")

    def tap(self, kind: int, dtype) -> np.ndarray:
        return _tap(lambda dst, cap, wr: lib().bb_world_debug_tap(self._h, kind, dst, cap, wr), dtype)

    def energy(self) -> float:
        """Mechanical energy of the current state (float64)."""
        e = C.c_double(0.0)
        _check(lib().bb_world_energy(self._h, C.byref(e)), "bb_world_energy")
        return e.value

    def snapshot(self) -> bytes:
        """Checkpoint of the world (bb_world_snapshot)."""
        size = C.c_size_t(0)
        _check(lib().bb_world_snapshot(self._h, None, 0, C.byref(size)), "bb_world_snapshot")
        buf = C.create_string_buffer(size.value)
        _check(lib().bb_world_snapshot(self._h, buf, size.value, C.byref(size)), "bb_world_snapshot")
        return buf.raw[: size.value]

    @classmethod
    def restore(cls, blob: bytes, warm_start: bool = True) -> "RigidBodyManager":
        """A world created from a checkpoint (bb_world_restore)."""
        w = cls.__new__(cls)
        w._h = C.c_void_p()
        _check(lib().bb_world_restore(blob, len(blob), 1 if warm_start else 0, C.byref(w._h)), "bb_world_restore")
        return w

    def instance_transforms(self, n: int, device_ptr: int | None = None):
        """Per body id, the 3x4 affine transform [R | t] of its pose (TLAS
        instance layout).  Returns a (n, 3, 4) float32 array, or, with a device
        pointer to n*12 floats, queues the write there and returns None."""
        if device_ptr is not None:
            _check(lib().bb_world_instance_transforms(self._h, C.c_void_p(device_ptr), n, 1),
                   "bb_world_instance_transforms")
            return None
        out = np.zeros((n, 3, 4), np.float32)
        _check(lib().bb_world_instance_transforms(self._h, out.ctypes.data_as(C.c_void_p), n, 0),
               "bb_world_instance_transforms")
        return out

    def lbvh_nodes(self, device_ptr: int | None = None, cap: int = 0):
        """The last step's BVH in the reference's LBVHNode layout
        (bb_world_lbvh_nodes): a LBVH_NODE array on the host, or, with a
        device pointer to `cap` 40-byte records, the node count after
        queueing the write there."""
        cnt = C.c_uint32(0)
        _check(lib().bb_world_lbvh_nodes(self._h, None, 0, C.byref(cnt), 0), "bb_world_lbvh_nodes")
        if device_ptr is not None:
            _check(lib().bb_world_lbvh_nodes(self._h, C.c_void_p(device_ptr), cap, C.byref(cnt), 1),
                   "bb_world_lbvh_nodes")
            return cnt.value
        out = np.zeros(cnt.value, LBVH_NODE)
        _check(lib().bb_world_lbvh_nodes(self._h, out.ctypes.data_as(C.c_void_p), cnt.value, C.byref(cnt), 0),
               "bb_world_lbvh_nodes")
        return out

    def destroy(self) -> None:
        if self._h:
            lib().bb_world_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _tap(call, dtype) -> np.ndarray:
    written = C.c_size_t(0)
    call(None, 0, C.byref(written))
    nbytes = written.value
    buf = np.zeros(max(nbytes, 1), np.uint8)
    code = call(buf.ctypes.data_as(C.c_void_p), nbytes, C.byref(written))
    _check(code, "debug_tap")
    return buf[:nbytes].view(dtype).copy()


class Batch:
    """Many independent worlds stepped together (segmented, one workgroup per world)."""

    def __init__(self, capacity_worlds: int = 0, capacity_bodies: int = 0):
        self._h = C.c_void_p()
        _check(lib().bb_batch_create(capacity_worlds, capacity_bodies, C.byref(self._h)),
               "bb_batch_create")

    def add_voxel_shape(self, solid: np.ndarray, voxel_size: float, density: float) -> dict:
        """A voxel shape in the batch's pools (shared by all its worlds)."""
        d, words = voxel_desc(solid, voxel_size, density)
        info = VoxelShapeInfo()
        _check(lib().bb_batch_add_voxel_shape(self._h, C.byref(d), C.byref(info)), "bb_batch_add_voxel_shape")
        del words
        return info.as_dict()

    def add_world(self, desc: WorldDesc, bodies: np.ndarray) -> int:
        arr, ptr = _bodies_ptr(bodies)
        w = C.c_uint32(0)
        _check(lib().bb_batch_add_world(self._h, C.byref(desc), ptr, len(arr), C.byref(w)),
               "bb_batch_add_world")
        return w.value

    def step(self, nsteps: int = 1, use_graph: bool = False) -> None:
        _check(lib().bb_batch_step(self._h, nsteps, None, 1 if use_graph else 0), "bb_batch_step")

    def sync(self, allow_overflow: bool = False) -> None:
        """Waits for the steps.  Raises BBError(BB_ERR_OVERFLOW) while a
        world's overflow mark is uncollected, unless allow_overflow."""
        rc = lib().bb_batch_sync(self._h)
        if rc == 5 and allow_overflow:
            return
        _check(rc, "bb_batch_sync")

    def overflow(self, cap: int | None = None) -> list:
        """Indices of the worlds whose steps flagged an overflow since the
        last call (bb_batch_overflow); clears the marks."""
        cap = self.world_count if cap is None else cap
        out = np.zeros(max(cap, 1), np.uint32)
        cnt = C.c_uint32(0)
        _check(lib().bb_batch_overflow(self._h, out.ctypes.data_as(C.c_void_p), cap, C.byref(cnt)),
               "bb_batch_overflow")
        return out[:min(cnt.value, cap)].tolist()

    def profile(self, nsteps: int, counts: bool = False):
        """Eager steps timed with hipEvents on the library stream: returns
        (per-step ms array, per-stage summed ms array) and, with counts=True,
        a (nsteps, COUNT_FIELDS) uint64 array of each step's batch totals."""
        step_ms = np.zeros(nsteps, np.float32)
        kern_ms = np.zeros(KERNEL_COUNT, np.float32)
        cnt = np.zeros((nsteps, COUNT_FIELDS), np.uint64) if counts else None
        _check(lib().bb_batch_profile(self._h, nsteps, step_ms.ctypes.data_as(C.c_void_p),
                                      kern_ms.ctypes.data_as(C.c_void_p),
                                      None if cnt is None else cnt.ctypes.data_as(C.c_void_p)),
               "bb_batch_profile")
        return (step_ms, kern_ms, cnt) if counts else (step_ms, kern_ms)

    def build_split(self) -> np.ndarray:
        """bb_batch_build_split: ms of the build stage's parts (BUILD_NAMES)
        inside the slowest workgroup team, summed over the last profile()
        call's steps (zeros without teams)."""
        out = np.zeros(BUILD_COUNT, np.float32)
        _check(lib().bb_batch_build_split(self._h, out.ctypes.data_as(C.c_void_p)), "bb_batch_build_split")
        return out

    def avbd_split(self) -> np.ndarray:
        """bb_batch_avbd_split: ms of the AVBD stage's parts (AVBD_NAMES)
        inside the slowest workgroup team, summed over the last profile()
        call's steps (zeros without an AVBD team)."""
        out = np.zeros(AVBD_COUNT, np.float32)
        _check(lib().bb_batch_avbd_split(self._h, out.ctypes.data_as(C.c_void_p)), "bb_batch_avbd_split")
        return out

    def energy(self) -> np.ndarray:
        """Mechanical energy of each world's current state (float64)."""
        out = np.zeros(max(self.world_count, 1), np.float64)
        _check(lib().bb_batch_energy(self._h, out.ctypes.data_as(C.c_void_p), len(out)), "bb_batch_energy")
        return out[: self.world_count]

    def instance_transforms(self, w: int, n: int, device_ptr: int | None = None):
        """RigidBodyManager.instance_transforms for world w."""
        if device_ptr is not None:
            _check(lib().bb_batch_instance_transforms(self._h, w, C.c_void_p(device_ptr), n, 1),
                   "bb_batch_instance_transforms")
            return None
        out = np.zeros((n, 3, 4), np.float32)
        _check(lib().bb_batch_instance_transforms(self._h, w, out.ctypes.data_as(C.c_void_p), n, 0),
               "bb_batch_instance_transforms")
        return out

    @property
    def device_bytes(self) -> int:
        return lib().bb_batch_device_bytes(self._h)

    def set_block_size(self, threads: int) -> None:
        _check(lib().bb_batch_set_block_size(self._h, threads), "bb_batch_set_block_size")

    def set_param(self, p: int, value: int) -> None:
        _check(lib().bb_batch_set_param(self._h, p, value), "bb_batch_set_param")

    def get_param(self, p: int) -> int:
        return lib().bb_batch_get_param(self._h, p)

    def autotune(self, reps: int = 2) -> float:
        """Times every launch-parameter candidate on replays of the next step
        (state unchanged afterwards) and keeps the fastest; returns its ms."""
        ms = C.c_float(0.0)
        _check(lib().bb_batch_autotune(self._h, reps, C.byref(ms)), "bb_batch_autotune")
        return ms.value

    @property
    def world_count(self) -> int:
        return lib().bb_batch_world_count(self._h)

    @property
    def body_count(self) -> int:
        return lib().bb_batch_body_count(self._h)

    def read_world(self, w: int, n: int, order: int = ORDER_BY_ID) -> np.ndarray:
        out = np.zeros(n, BODY_STATE)
        _check(lib().bb_batch_read_world(self._h, w, out.ctypes.data_as(C.c_void_p), n, order),
               "bb_batch_read_world")
        return out

    def stats(self, w: int) -> dict:
        s = StepStats()
        _check(lib().bb_batch_world_stats(self._h, w, C.byref(s)), "bb_batch_world_stats")
        return s.as_dict()

    def tap(self, w: int, kind: int, dtype) -> np.ndarray:
        return _tap(lambda dst, cap, wr: lib().bb_batch_debug_tap(self._h, w, kind, dst, cap, wr),
                    dtype)

    def snapshot(self) -> bytes:
        """Checkpoint of every world of the batch (bb_batch_snapshot)."""
        size = C.c_size_t(0)
        _check(lib().bb_batch_snapshot(self._h, None, 0, C.byref(size)), "bb_batch_snapshot")
        buf = C.create_string_buffer(size.value)
        _check(lib().bb_batch_snapshot(self._h, buf, size.value, C.byref(size)), "bb_batch_snapshot")
        return buf.raw[: size.value]

    @classmethod
    def restore(cls, blob: bytes, warm_start: bool = True) -> "Batch":
        """A batch created from a checkpoint (bb_batch_restore)."""
        b = cls.__new__(cls)
        b._h = C.c_void_p()
        _check(lib().bb_batch_restore(blob, len(blob), 1 if warm_start else 0, C.byref(b._h)),
               "bb_batch_restore")
        return b

    def destroy(self) -> None:
        if self._h:
            lib().bb_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Server:
    """Serving loop over the C-ABI (bb_server_*): continuous batching of
    worlds admitted by HBM budget, retired when their steps are done."""

    def __init__(self, budget_bytes: int, max_worlds: int = 1024, max_bodies: int = 0,
                 max_resident_bodies: int = 0, block: int = 0, team_bodies: int = 0, split_bodies: int = 0):
        self._h = C.c_void_p()
        d = ServerDesc(budget_bytes, max_worlds, max_bodies, max_resident_bodies, block, team_bodies,
                       split_bodies)
        _check(lib().bb_server_create(C.byref(d), C.byref(self._h)), "bb_server_create")

    def add_voxel_shape(self, solid: np.ndarray, voxel_size: float, density: float) -> dict:
        """A voxel shape for the requests submitted afterwards (bb_server_add_voxel_shape)."""
        d, words = voxel_desc(solid, voxel_size, density)
        info = VoxelShapeInfo()
        _check(lib().bb_server_add_voxel_shape(self._h, C.byref(d), C.byref(info)), "bb_server_add_voxel_shape")
        del words
        return info.as_dict()

    def submit(self, desc: WorldDesc, bodies: np.ndarray, steps: int) -> int:
        arr, ptr = _bodies_ptr(bodies)
        t = C.c_uint64(0)
        _check(lib().bb_server_submit(self._h, C.byref(desc), ptr, len(arr), steps, C.byref(t)),
               "bb_server_submit")
        return t.value

    def step(self, nsteps: int = 1) -> None:
        _check(lib().bb_server_step(self._h, nsteps), "bb_server_step")

    def drain(self) -> None:
        _check(lib().bb_server_drain(self._h), "bb_server_drain")

    def checkpoint(self) -> bytes:
        """The serving loop's checkpoint (bb_server_checkpoint): resident
        worlds with their state, queued requests, unpolled results."""
        size = C.c_size_t(0)
        _check(lib().bb_server_checkpoint(self._h, None, 0, C.byref(size)), "bb_server_checkpoint")
        buf = (C.c_char * size.value)()
        _check(lib().bb_server_checkpoint(self._h, buf, size.value, C.byref(size)), "bb_server_checkpoint")
        return bytes(buf)

    @classmethod
    def restore(cls, blob: bytes, budget_bytes: int, max_worlds: int = 1024, max_bodies: int = 0,
                max_resident_bodies: int = 0, block: int = 0, team_bodies: int = 0,
                split_bodies: int = 0) -> "Server":
        """A new server holding a checkpoint's worlds (bb_server_restore)."""
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        d = ServerDesc(budget_bytes, max_worlds, max_bodies, max_resident_bodies, block, team_bodies,
                       split_bodies)
        buf = (C.c_char * len(blob)).from_buffer_copy(blob)
        _check(lib().bb_server_restore(C.byref(d), buf, len(blob), C.byref(self._h)), "bb_server_restore")
        return self

    def poll(self, ticket: int, n: int = 0):
        """(state, bodies or None, stats or None); results are released once
        returned."""
        st = C.c_uint32(0)
        out = np.zeros(max(n, 1), BODY_STATE)
        stats = StepStats()
        _check(lib().bb_server_poll(self._h, ticket, C.byref(st), out.ctypes.data_as(C.c_void_p), max(n, 1),
                                    C.byref(stats)), "bb_server_poll")
        if st.value in (REQ_DONE, REQ_FAILED):
            return st.value, out[:n], stats.as_dict()
        return st.value, None, None

    def report(self) -> dict:
        r = ServerReport()
        _check(lib().bb_server_report_get(self._h, C.byref(r)), "bb_server_report_get")
        return r.as_dict()

    def destroy(self) -> None:
        if self._h:
            lib().bb_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
