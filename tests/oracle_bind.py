"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker for the HIP step, never the
thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "beat-box_amd"))

import beatbox  # noqa: E402
from beatbox import (BODY_STATE, MANIFOLD_TAP, LBVH_NODE, StepStats, VoxelShapeDesc, VoxelShapeInfo,  # noqa: E402
                     WorldDesc, voxel_desc)
from scenes import BODY_DESC  # noqa: E402

ORACLE_SO = os.path.join(REPO, "oracle", "_build", "liboracle.so")


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)


_OL = None


def olib() -> C.CDLL:
    global _OL
    if _OL is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = C.CDLL(ORACLE_SO)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(WorldDesc)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_add_bodies.restype = C.c_int
        L.oracle_add_bodies.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
        L.oracle_step.argtypes = [C.c_void_p, C.c_uint32]
        L.oracle_set_flags.argtypes = [C.c_void_p, C.c_uint32]
        L.oracle_set_solver.argtypes = [C.c_void_p, C.c_uint32]
        L.oracle_set_iterations.argtypes = [C.c_void_p, C.c_uint32]
        L.oracle_skip_warm_start_once.argtypes = [C.c_void_p]
        L.oracle_read_bodies.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int]
        L.oracle_read_stats.argtypes = [C.c_void_p, C.POINTER(StepStats)]
        L.oracle_tap.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.oracle_morton30.restype = C.c_uint32
        L.oracle_morton30.argtypes = [C.c_float] * 3
        L.oracle_morton63.restype = C.c_uint64
        L.oracle_morton63.argtypes = [C.c_float] * 3
        L.oracle_expand10.restype = C.c_uint32
        L.oracle_expand10.argtypes = [C.c_uint32]
        L.oracle_expand21.restype = C.c_uint64
        L.oracle_expand21.argtypes = [C.c_uint64]
        L.oracle_sincos.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.oracle_order_map.restype = C.c_uint32
        L.oracle_order_map.argtypes = [C.c_float]
        L.oracle_order_unmap.restype = C.c_float
        L.oracle_order_unmap.argtypes = [C.c_uint32]
        L.oracle_solve6.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_add_voxel_shape.restype = C.c_int
        L.oracle_add_voxel_shape.argtypes = [C.c_void_p, C.POINTER(VoxelShapeDesc), C.POINTER(VoxelShapeInfo)]
        L.oracle_collide.restype = C.c_int
        L.oracle_collide.argtypes = [C.POINTER(WorldDesc), C.c_void_p, C.c_void_p, C.c_void_p]
        _OL = L
    return _OL


class OracleWorld:
    def __init__(self, desc: WorldDesc):
        self.desc = desc
        self._h = C.c_void_p(olib().oracle_create(C.byref(desc)))

    def add_bodies(self, bodies: np.ndarray) -> None:
        arr = np.ascontiguousarray(bodies, dtype=BODY_DESC)
        if olib().oracle_add_bodies(self._h, arr.ctypes.data_as(C.c_void_p), len(arr)) != 0:
            raise ValueError("oracle_add_bodies: invalid body")
        self.n = getattr(self, "n", 0) + len(arr)

    def step(self, n: int = 1) -> None:
        olib().oracle_step(self._h, n)

    def set_flags(self, f: int) -> None:
        olib().oracle_set_flags(self._h, f)

    def add_voxel_shape(self, solid: np.ndarray, voxel_size: float, density: float) -> dict:
        d, words = voxel_desc(solid, voxel_size, density)
        info = VoxelShapeInfo()
        rc = olib().oracle_add_voxel_shape(self._h, C.byref(d), C.byref(info))
        del words
        if rc != 0:
            raise ValueError(f"oracle_add_voxel_shape: {rc}")
        return info.as_dict()

    def set_solver(self, s: int) -> None:
        olib().oracle_set_solver(self._h, s)

    def skip_warm_start_once(self) -> None:
        olib().oracle_skip_warm_start_once(self._h)

    def read_bodies(self, order: int = 0) -> np.ndarray:
        out = np.zeros(self.n, BODY_STATE)
        olib().oracle_read_bodies(self._h, out.ctypes.data_as(C.c_void_p), self.n, order)
        return out

    def stats(self) -> dict:
        s = StepStats()
        olib().oracle_read_stats(self._h, C.byref(s))
        return s.as_dict()

    def tap(self, kind: int, dtype) -> np.ndarray:
        return beatbox._tap(lambda dst, cap, wr: olib().oracle_tap(self._h, kind, dst, cap, wr), dtype)

    def __del__(self):
        try:
            olib().oracle_destroy(self._h)
        except Exception:
            pass


def collide(desc: WorldDesc, a: np.ndarray, b: np.ndarray):
    out = np.zeros(1, MANIFOLD_TAP)
    aa = np.ascontiguousarray(a, dtype=BODY_DESC)
    bb = np.ascontiguousarray(b, dtype=BODY_DESC)
    have = olib().oracle_collide(C.byref(desc), aa.ctypes.data_as(C.c_void_p),
                                 bb.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out[0] if have else None


__all__ = ["OracleWorld", "olib", "collide", "build_oracle", "LBVH_NODE", "MANIFOLD_TAP"]
