"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

The reference ships no test vectors; these fixtures freeze the oracle's
canonical outputs on small seeded scenes so both the oracle (CPU suite) and
the HIP step (GPU suite) are checked against committed data:
  sat_cases.npz     two-body contact cases (face, rotated face, edge-edge,
                    box-sphere, sphere-sphere, separated)
  pile64.npz        64 random cubes + ground: taps after steps 1 and 6
  config1.npz       the 1,000-cube grid: discrete outputs of steps 0-3 and the
                    body state after step 3
  freefall.npz      one falling box, PGS_SOFT and AVBD trajectories
Run: python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "beat-box_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

import beatbox  # noqa: E402
import scenes  # noqa: E402
from beatbox import (LBVH_NODE, MANIFOLD_TAP, TAP_EDGE_COLORS, TAP_ISLANDS, TAP_LBVH,  # noqa: E402
                     TAP_MANIFOLDS, TAP_MORTON, TAP_PAIRS, TAP_SORTED, WorldDesc)
from oracle_bind import OracleWorld, collide  # noqa: E402


def box(pos, quat=(0, 0, 0, 1), half=0.5, dynamic=True):
    b = np.zeros(1, scenes.BODY_DESC)
    b["shape"] = 0
    b["flags"] = 6 if dynamic else 0
    b["position"] = pos
    b["rotation"] = quat
    b["minimum"] = -half
    b["maximum"] = half
    b["mass"] = 1.0 if dynamic else 0.0
    b["friction"] = 0.5
    return b


def sphere(pos, r):
    b = np.zeros(1, scenes.BODY_DESC)
    b["shape"] = 1
    b["flags"] = 6
    b["position"] = pos
    b["rotation"] = (0, 0, 0, 1)
    b["radius"] = r
    b["mass"] = 1.0
    b["friction"] = 0.5
    return b


def sat_cases():
    g = np.zeros(1, scenes.BODY_DESC)
    scenes._ground(g[0])
    s8, c8 = math.sin(math.pi / 8), math.cos(math.pi / 8)
    q = scenes.shoemake(7, np.array([0], np.uint64))[0]
    cases = [
        (g, box((0.3, 0.49, 0.2))),
        (g, box((0.0, 0.45, 0.0), (0, s8, 0, c8))),
        (g, box((1.0, 0.6, -1.0), tuple(q))),
        (box((0, 0, 0), (s8, 0, 0, c8)), box((0, 0, 1.35), (0, s8, 0, c8))),
        (box((0, 0.5, 0)), box((0.3, 1.45, 0.1), tuple(q))),
        (box((0, 0.5, 0)), sphere((0.8, 1.3, 0.8), 0.6)),
        (sphere((0, 1, 0), 0.5), sphere((0.6, 1.8, 0), 0.5)),
        (g, box((0.0, 0.6, 0.0))),
    ]
    desc = WorldDesc.make(**scenes.world_desc())
    a = np.concatenate([c[0] for c in cases])
    b = np.concatenate([c[1] for c in cases])
    out = np.zeros(len(cases), MANIFOLD_TAP)
    have = np.zeros(len(cases), np.int32)
    for i, (x, y) in enumerate(cases):
        m = collide(desc, x, y)
        if m is not None:
            out[i] = m
            have[i] = 1
    return dict(a=a, b=b, manifolds=out, have=have)


def taps(w, bits):
    return dict(
        morton=w.tap(TAP_MORTON, np.uint32 if bits == 30 else np.uint64),
        sorted=w.tap(TAP_SORTED, np.uint32),
        lbvh=w.tap(TAP_LBVH, LBVH_NODE),
        pairs=w.tap(TAP_PAIRS, np.uint32),
        islands=w.tap(TAP_ISLANDS, np.uint32),
        colors=w.tap(TAP_EDGE_COLORS, np.uint32),
        manifolds=w.tap(TAP_MANIFOLDS, MANIFOLD_TAP),
        bodies=w.read_bodies(),
    )


def pile64():
    bodies = scenes.random_boxes(0xF064, 64, extent=2.5, y0=0.55, y1=4.0)
    desc = WorldDesc.make(**scenes.world_desc(iterations=8))
    w = OracleWorld(desc)
    w.add_bodies(bodies)
    out = {"scene": bodies}
    w.step(1)
    out.update({f"s1_{k}": v for k, v in taps(w, 30).items()})
    w.step(5)
    out.update({f"s6_{k}": v for k, v in taps(w, 30).items()})
    return out


def config1():
    bodies = scenes.cubes_grid()
    desc = WorldDesc.make(**scenes.world_desc(iterations=8))
    w = OracleWorld(desc)
    w.add_bodies(bodies)
    out = {}
    for s in range(4):
        w.step(1)
        t = taps(w, 30)
        m = t["manifolds"]
        out[f"s{s}_morton"] = t["morton"]
        out[f"s{s}_order"] = t["sorted"].reshape(-1, 4)[:, 2].copy()
        out[f"s{s}_pairs"] = t["pairs"]
        out[f"s{s}_islands"] = t["islands"]
        out[f"s{s}_colors"] = t["colors"]
        out[f"s{s}_mkeys"] = np.stack([m["obb1_index"], m["obb2_index"], m["key"].astype(np.uint32),
                                       m["contact_count"].astype(np.uint32)], 1)
    out["s3_bodies"] = w.read_bodies()
    return out


def freefall():
    out = {}
    for name, solver in (("pgs", scenes.SOLVER_PGS_SOFT), ("avbd", scenes.SOLVER_AVBD)):
        b = box((0.0, 50.0, 0.0))
        b["velocity"] = (1.0, 0.5, -0.25)
        b["omega"] = (0.3, 1.1, -0.7)
        desc = WorldDesc.make(**scenes.world_desc(solver=solver))
        w = OracleWorld(desc)
        w.add_bodies(b)
        traj = []
        for _ in range(30):
            w.step(1)
            traj.append(w.read_bodies()[0])
        out[name] = np.array(traj)
        out[name + "_scene"] = b
    return out


def main():
    np.savez_compressed(os.path.join(HERE, "sat_cases.npz"), **sat_cases())
    np.savez_compressed(os.path.join(HERE, "pile64.npz"), **pile64())
    np.savez_compressed(os.path.join(HERE, "config1.npz"), **config1())
    np.savez_compressed(os.path.join(HERE, "freefall.npz"), **freefall())


if __name__ == "__main__":
    main()
