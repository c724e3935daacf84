"""Shared helpers of the GPU-vs-oracle parity tests (test infrastructure)."""
from __future__ import annotations

import numpy as np

import beatbox
from beatbox import (LBVH_NODE, MANIFOLD_TAP, TAP_EDGE_COLORS, TAP_ISLANDS, TAP_LBVH,
                     TAP_MANIFOLDS, TAP_MORTON, TAP_PAIRS, TAP_SORTED, TAP_BODY_COLORS)

STATE_FIELDS = ("position", "rotation", "velocity", "omega", "flags", "id")
STAT_KEYS = ("rigid_body_count", "active_rigid_body_count", "broad_phase_collision_count",
             "collision_count", "manifold_node_count", "contact_count", "island_count",
             "contact_island_count", "sleeping_count", "graph_color_count",
             "graph_color_violations", "graph_color_overflow", "avbd_color_count",
             "avbd_violations", "avbd_stick_count", "dbg_fresh", "dbg_pen", "dbg_id_sum",
             "dbg_maxv")


def morton_dtype(bits: int):
    return np.uint32 if bits == 30 else np.uint64


def compare_step(gpu, ora, bits: int, step: int, check_taps: bool = True) -> list:
    """Returns a list of mismatch descriptions (empty = bit-exact)."""
    bad = []
    if check_taps:
        taps = [(TAP_MORTON, morton_dtype(bits), "morton"), (TAP_SORTED, np.uint32, "sorted"),
                (TAP_LBVH, LBVH_NODE, "lbvh"), (TAP_PAIRS, np.uint32, "pairs"),
                (TAP_ISLANDS, np.uint32, "islands"), (TAP_EDGE_COLORS, np.uint32, "edge_colors"),
                (TAP_BODY_COLORS, np.uint32, "body_colors"), (TAP_MANIFOLDS, MANIFOLD_TAP, "manifolds")]
        for kind, dt, name in taps:
            g = gpu.tap(kind, dt)
            o = ora.tap(kind, dt)
            if g.shape != o.shape:
                bad.append(f"step {step} {name}: shape {g.shape} vs oracle {o.shape}")
            elif g.tobytes() != o.tobytes():
                bad.append(f"step {step} {name}: differs at {first_diff(g, o)}")
    g = gpu.read_bodies()
    o = ora.read_bodies()
    for f in STATE_FIELDS:
        if g[f].tobytes() != o[f].tobytes():
            idx = np.nonzero(np.any((g[f] != o[f]).reshape(len(g), -1), axis=1))[0]
            bad.append(f"step {step} bodies.{f}: {len(idx)} bodies differ, first id {idx[:5]}")
    gs, os_ = gpu.read_back_sim_config(), ora.stats()
    for k in STAT_KEYS:
        if gs[k] != os_[k]:
            bad.append(f"step {step} stats.{k}: {gs[k]} vs oracle {os_[k]}")
    return bad


def first_diff(g: np.ndarray, o: np.ndarray):
    gb = g.view(np.uint8).reshape(len(g), -1) if g.ndim == 1 and g.dtype.fields else g.reshape(len(g), -1)
    ob = o.view(np.uint8).reshape(len(o), -1) if o.ndim == 1 and o.dtype.fields else o.reshape(len(o), -1)
    rows = np.nonzero(np.any(gb != ob, axis=1))[0]
    return f"rows {rows[:8].tolist()} (of {len(rows)})"
