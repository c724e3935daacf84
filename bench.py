"""Benchmark of the MI355X rigid-body step.

Workload (the configuration the headline metric is quoted on): 4,096
independent worlds in one batch, body counts log-uniform in [16, 2048]
(seed 0xBB04), each a box pyramid (even) or 4x4 tower (odd) on its own ground
box, PGS_SOFT with 8 iterations, FP32, 63-bit Morton keys, hipGraph replay,
240 timed steps after 10 untimed ones.  Beside it: per-stage kernel durations
over the same step window (hipEvents on the library stream, eager launches of
a fresh batch that reaches the same states), the plain single world (1,000
cubes + ground, PGS_SOFT I=8, parity mode) over the same step window on the
same GPU, the configuration's own 240-step window when another one is timed,
the worlds that flagged an overflow, the batch energy change and maximum
penetration, and the CPU oracle on a bounded sample of the batch.

Prints one JSON line; value = batch body-steps/s with all state resident in
HBM.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "beat-box_amd"))

import beatbox  # noqa: E402
import scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0
# pair slots of each batch world, times the default 8n + 64: a batch world
# cannot grow within bb_batch_step, and a few stacked pyramids need more than
# 8n + 64 pairs in the configuration's window
PAIR_SCALE = 2
# the batched-worlds configuration's own window: 240 timed steps after 10
CONFIG_STEPS, CONFIG_WARMUP = 240, 10
TRAFFIC_FILE = os.path.join(REPO, "profiles", "bench_batch4096_traffic.json")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_bytes(counts: np.ndarray, iters: int, kw: int) -> dict:
    """Algorithmic (compulsory) HBM bytes of each stage kernel, summed over
    the profiled steps, from the per-unit figures of the byte model and each
    step's device-reported batch totals: N bodies, A active, P pairs,
    M manifolds, C contacts, G = sum over worlds of colours used x M,
    S = 84-byte compact state row, kw = key bytes."""
    c = counts.astype(np.float64)
    N, A, P, M, Cc, G = (c[:, i].sum() for i in range(6))
    S = 84
    passes = 4 if kw == 4 else 8
    return {
        "k_build": N * (12 + kw + 4) + passes * N * (3 * kw + 8) + N * (40 + 2 * kw)
        + (2 * N - len(c)) * 36 + N * 72 + 2 * N * S,
        "k_broad": N * 24 + P * 8,
        "k_narrow": P * 96 + M * 48 + Cc * 112,
        "k_post": M * 8 + A * 16 + G * 20 + N * 4,
        "k_pgs": (1 + 2 * iters) * (M * (36 + 2 * S) + Cc * 60) + A * 112,
        "k_avbd": 0.0,
        "k_finish": N * 16,
    }


def build_bytes(n: int, kw: int) -> dict:
    """Algorithmic HBM bytes of the build stage's parts for one world of n
    bodies (kw = key bytes): keys read the previous positions (twice for
    63-bit keys, whose box is the bodies' own bounds) and write key + index;
    each of the 8 (63-bit) or 4 (30-bit) LSD passes reads the keys for the
    histogram and reads and writes key + index in the scatter; the hierarchy
    reads each leaf's index, id, pose and shape bounds and writes its node,
    and per internal node reads two keys and writes children, parents, the
    visit flag and the leaf range; the bounds pass reads a node and its two
    children and writes it (the lead workgroup's climb does the same for the
    about 2n/512 nodes above the 512-leaf runs); the reorder moves every row
    (id, flags, sleep timer and five float4 rows) and writes the id -> row
    map and two per-row lookups."""
    passes = 8 if kw == 8 else 4
    return {
        "keys": n * ((32 if kw == 8 else 16) + kw + 4),
        "sort": passes * n * (3 * kw + 8),
        "hierarchy": n * 104 + (n - 1) * 68,
        "bounds": (n - 1) * 136,
        "climb": (2 * n // 512) * 136,
        "reorder": n * 212,
    }


def lane_launches(kernel: str, params: dict, n_wide: int, worlds: int) -> list:
    """The dispatches of `kernel` in one step of the batch, as the
    (workgroup size, workgroups) of each lane: the first lane steps the
    worlds below the wide lane's threshold with BB_PARAM_BLOCK workgroups
    (its colour sweeps with BB_PARAM_PGS_BLOCK, when set), the wide lane
    the n_wide others with BB_PARAM_LANE_BLOCK ones."""
    first = params["block"]
    if kernel == "k_pgs" and params.get("pgs_block"):
        first = params["pgs_block"]
    out = [(first, worlds - n_wide)]
    if n_wide:
        out.append((params["lane_block"], n_wide))
    return out


def profiled_traffic(kernel: str, params: dict, n_wide: int, worlds: int):
    """HBM bytes per step (read + write) of `kernel` over the whole batch,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this same
    command (copy-kernel calibrated): the sum over the step's dispatches of
    the kernel (one per lane, named by workgroup size and count), or None
    when the profile has no such dispatches (other launch parameters)."""
    try:
        with open(TRAFFIC_FILE) as f:
            k = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    total = 0.0
    for blk, grid in lane_launches(kernel, params, n_wide, worlds):
        v = k.get(f"bb::{kernel}<{blk}>@{grid}")
        if v is None or v["read_bytes"] is None:
            return None
        total += float(v["read_bytes"]) + float(v["write_bytes"] or 0.0)
    return total


def traffic_fields(kernel: str, args, model_per_launch: float, params: dict, n_wide: int) -> dict:
    """roofline.traffic: HBM bytes per launch of `kernel` from the committed
    rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE, copy-kernel
    calibrated), reported only when they were taken over this run's step
    window; otherwise null, with the profiled window's figures beside it,
    each set against the algorithmic bytes of its own window."""
    try:
        with open(TRAFFIC_FILE) as f:
            meta = json.load(f)
    except (OSError, ValueError):
        meta = {}
    prof = profiled_traffic(kernel, params, n_wide, args.worlds)
    window = meta.get("window")  # {"steps": S, "warmup": W} of the counter runs
    model = meta.get("model_bytes_per_launch", {}).get(kernel)
    # the same step window and, where the profile records them, the same launch parameters
    same = (bool(window) and (window.get("steps"), window.get("warmup")) == (args.steps, args.warmup)
            and meta.get("launch_params") in (None, params))
    out = {"traffic": prof if same else None,
           "traffic_over_model": (prof / model_per_launch) if (same and prof) else None,
           "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py (copy-kernel calibrated), "
                             "profiles/bench_batch4096_traffic.json"}
    if prof is not None and not same:
        out["traffic_profiled"] = dict(window=window, launch_params=meta.get("launch_params"),
                                       bytes_per_launch=prof, model_bytes_per_launch=model,
                                       over_model=(prof / model) if model else None)
    return out


def make_batch(worlds: int, seed: int, lo: int, hi: int):
    t0 = time.time()
    sizes = scenes.batch_world_sizes(seed, worlds, lo, hi)
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    b = beatbox.Batch()
    b.set_param(beatbox.PARAM_PAIR_SCALE, PAIR_SCALE)
    total = 0
    b.world_bodies = []
    for w, n in enumerate(sizes):
        bodies = scenes.batch_world(seed, w, int(n))
        b.add_world(desc, bodies)
        total += len(bodies)
        b.world_bodies.append(len(bodies))
    log(f"batch: {worlds} worlds, {total} bodies, built in {time.time() - t0:.1f}s")
    return b, total


# Launch parameters the autotuner chooses; re-made batches (per-stage profile,
# configuration window) are given the same values.
TUNED = ("BLOCK", "PGS_LDS_ROWS", "PGS_BLOCK", "FUSED", "PGS_PREFETCH", "WALK_BUDGET", "LANE_BODIES",
         "LANE_BLOCK")


def apply_params(b, params: dict) -> None:
    for name in TUNED:
        b.set_param(getattr(beatbox, "PARAM_" + name), params[name.lower()])


def time_batch(b, steps: int, warmup: int, tune: bool, fixed: dict | None = None):
    """Graph replay of `steps` steps after `warmup` untimed ones (and the
    autotuner, which replays the next step per candidate and leaves the state
    unchanged, or the `fixed` launch parameters); returns the wall time, the
    batch energy at step 0 and after the timed steps and the launch
    parameters used."""
    e0 = float(b.energy().sum())
    b.step(warmup, use_graph=True)
    b.sync(allow_overflow=True)
    if fixed:
        apply_params(b, fixed)
    elif tune:
        b.autotune(2)
    params = {name.lower(): b.get_param(getattr(beatbox, "PARAM_" + name)) for name in TUNED}
    t0 = time.perf_counter()
    b.step(steps, use_graph=True)
    b.sync(allow_overflow=True)
    dt = time.perf_counter() - t0
    overflowed = b.overflow()
    return dt, e0, float(b.energy().sum()), params, overflowed


def plain_world(steps: int, warmup: int):
    """The plain single world (1,000 cubes + ground, PGS_SOFT I=8, parity
    mode) over the same step window as the batch: hipGraph replay of
    `steps` steps after `warmup` untimed ones."""
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    w = beatbox.Batch()
    bodies = scenes.cubes_grid()
    w.add_world(desc, bodies)
    w.step(warmup, use_graph=True)
    w.sync(allow_overflow=True)
    t0 = time.perf_counter()
    w.step(steps, use_graph=True)
    w.sync(allow_overflow=True)
    dt = time.perf_counter() - t0
    return len(bodies) * steps / dt, dt / steps * 1e3


def single_world(name: str, desc, bodies, steps: int, warmup: int, profile: int, tune: bool = True) -> dict:
    """One large world stepped alone (hipGraph replay of `steps` steps after
    `warmup`), then `profile` eager steps timed per stage; worlds at or above
    the team threshold run as workgroup teams (k_team, whose time the profile
    splits by stage with the team's own clock marks)."""
    t0 = time.time()
    b = beatbox.Batch()
    b.add_world(desc, bodies)
    b.step(warmup, use_graph=True)
    b.sync(allow_overflow=True)
    build_s = time.time() - t0
    if tune:
        # the tuner replays the next step per candidate (the sort's sub-team
        # and the team launch's size for a team world) and leaves the state
        b.autotune(2)
    params = {n: b.get_param(getattr(beatbox, "PARAM_" + n.upper()))
              for n in ("sort_blocks", "team_blocks", "block", "avbd_lanes")}
    t0 = time.perf_counter()
    b.step(steps, use_graph=True)
    b.sync(allow_overflow=True)
    dt = time.perf_counter() - t0
    st, km, cnt = b.profile(profile, counts=True)
    split = b.build_split() / profile
    av_split = b.avbd_split() / profile
    s = b.stats(0)
    team = len(bodies) >= b.get_param(beatbox.PARAM_TEAM_BODIES) > 0
    kw = 8 if desc.morton_bits == 63 else 4
    bb = build_bytes(len(bodies), kw)
    build = {n: dict(ms=round(float(split[i]), 4), bytes=bb[n],
                     gbs=round(bb[n] / (float(split[i]) * 1e-3) / 1e9, 1) if split[i] > 0 else None)
             for i, n in enumerate(beatbox.BUILD_NAMES)} if team else None
    out = dict(bodies=len(bodies), steps=steps, warmup=warmup, ms_per_step=dt / steps * 1e3,
               body_steps_per_s=len(bodies) * steps / dt, team=team,
               profiled_step_ms_p50=float(np.percentile(st, 50)),
               stages_ms={n: round(float(km[i]) / profile, 3) for i, n in enumerate(beatbox.KERNEL_NAMES)},
               manifolds=s["collision_count"], contacts=s["contact_count"], pairs=s["broad_phase_collision_count"],
               max_penetration_mm=int(cnt[:, beatbox.COUNT_NAMES.index("max_pen_mm")].max()),
               overflow=s["overflow"], build_parts=build, launch_params=params,
               avbd_parts={n: round(float(av_split[i]), 3) for i, n in enumerate(beatbox.AVBD_NAMES)}
               if team and av_split.sum() > 0 else None)
    b.destroy()
    log(f"{name}: {out['ms_per_step']:.2f} ms/step, {out['body_steps_per_s'] / 1e6:.2f} M body-steps/s "
        f"(setup {build_s:.1f}s), {out}")
    return out


def batch_run(name: str, worlds, shapes=(), steps: int = 60, warmup: int = 10, profile: int = 10) -> dict:
    """A batch of (desc, bodies) worlds (voxel `shapes` added first): hipGraph
    replay of `steps` steps after `warmup`, then `profile` eager steps timed
    per stage."""
    b = beatbox.Batch()
    b.set_param(beatbox.PARAM_PAIR_SCALE, PAIR_SCALE)
    for cells in shapes:
        b.add_voxel_shape(cells, scenes.VOXEL_SIZE, scenes.VOXEL_DENSITY)
    total = 0
    for desc, bodies in worlds:
        b.add_world(desc, bodies)
        total += len(bodies)
    b.step(warmup, use_graph=True)
    b.sync(allow_overflow=True)
    t0 = time.perf_counter()
    b.step(steps, use_graph=True)
    b.sync(allow_overflow=True)
    dt = time.perf_counter() - t0
    _, km, cnt = b.profile(profile, counts=True)
    manifolds = int(cnt[:, beatbox.COUNT_NAMES.index("manifolds")].mean())
    overflow = len(b.overflow())
    b.destroy()
    out = dict(worlds=len(worlds), bodies=total, steps=steps, warmup=warmup, ms_per_step=dt / steps * 1e3,
               body_steps_per_s=total * steps / dt, manifolds=manifolds, overflow_worlds=overflow,
               pair_scale=PAIR_SCALE,
               stages_ms={n: round(float(km[i]) / profile, 3) for i, n in enumerate(beatbox.KERNEL_NAMES)})
    log(f"{name}: {out['ms_per_step']:.2f} ms/step, {out['body_steps_per_s'] / 1e6:.2f} M body-steps/s, {out}")
    return out


def next_rows(seed: int) -> dict:
    """Throughput of the paths beside the headline one: the per-island solver
    (plain PGS, USE_GRAPH_COLORING off) against the colour sweeps on the same
    1,024 worlds, voxel bodies, and checkpoint / render hand-off costs."""
    out = {}
    sizes = scenes.batch_world_sizes(seed, 4096, 16, 2048)[:1024]
    bodies = [scenes.batch_world(seed, w, int(n)) for w, n in enumerate(sizes)]
    colour = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    island = beatbox.WorldDesc.make(**scenes.world_desc(
        solver=scenes.SOLVER_PGS, iterations=8, morton_bits=63, quirks=0,
        flags=scenes.SIM_DEFAULT_FLAGS & ~scenes.SIM_USE_GRAPH_COLORING))
    c = batch_run("colour sweeps, 1,024 worlds", [(colour, x) for x in bodies])
    i = batch_run("per-island plain PGS, 1,024 worlds", [(island, x) for x in bodies])
    out["per_island"] = dict(workload="worlds 0-1023 of the headline batch, PGS I=8; plain PGS on the per-island "
                                      "path against PGS_SOFT colour sweeps", colour=c, island=i,
                             island_over_colour=i["body_steps_per_s"] / c["body_steps_per_s"])
    shapes, _ = scenes.voxel_mixed_pile(0, 1)
    avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    vworlds = [(avbd, scenes.voxel_mixed_pile(0xBB70 + k, 200)[1]) for k in range(64)]
    out["voxel"] = dict(workload="64 worlds of 200 bodies (3/5 voxel Ls, crosses and frames, 1/5 boxes, 1/5 "
                                 "spheres), AVBD 10 iterations",
                        **batch_run("voxel piles", vworlds, shapes=[cl for _, cl in shapes]))
    # checkpoint and render hand-off of one 20,000-body world
    w = beatbox.RigidBodyManager(avbd)
    pile = scenes.serving_pile(0xBB71, 20000)
    w.add_bodies(pile)
    w.simulate(30)
    w.sync()
    t0 = time.perf_counter()
    blob = w.snapshot()
    t1 = time.perf_counter()
    r = beatbox.RigidBodyManager.restore(blob, warm_start=True)
    r.sync()
    t2 = time.perf_counter()
    xf = w.instance_transforms(len(pile))
    t3 = time.perf_counter()
    out["checkpoint"] = dict(bodies=len(pile), blob_mb=len(blob) / 1e6, snapshot_ms=(t1 - t0) * 1e3,
                             restore_ms=(t2 - t1) * 1e3)
    out["handoff_host"] = dict(bodies=len(pile), bytes=int(xf.nbytes), ms=(t3 - t2) * 1e3,
                               note="host destination: device kernel plus the PCIe copy")
    r.destroy()
    w.destroy()
    log(f"checkpoint: {out['checkpoint']}, hand-off: {out['handoff_host']}")
    return out


def serving_stream(seed: int, duration: float):
    """The stream's arrival times and request bodies (generated once per
    seed and duration, and kept for the next run)."""
    key = (seed, duration)
    if getattr(serving_stream, "key", None) != key:
        arr, nb, hz, sd = scenes.serving_requests(seed, duration=duration)
        serving_stream.key = key
        serving_stream.arrivals = arr
        serving_stream.horizons = hz
        serving_stream.bodies = [scenes.serving_pile(int(sd[k]) & 0xFFFFFFFF, int(nb[k])) for k in range(len(arr))]
    return serving_stream.arrivals, serving_stream.bodies


def serving(duration: float, budget_gib: float, resident_cap: int, team_bodies: int, split_bodies: int,
            seed: int = 0xBB05) -> dict:
    """Serving stream: Poisson arrivals at 500 worlds/s for `duration` seconds
    (random piles of 64-20,000 bodies, AVBD 10 iterations, horizons 60-600
    steps), admitted by the HBM budget (and a resident-body cap unless 0)
    into pools by size that step at their own cadence (split_bodies 0: one
    lockstep pool); reports the server's counters and latency percentiles,
    overall and per pool, at the end of the arrival window.  The requests'
    bodies are generated before the clock starts, as a server receives them
    ready-made."""
    desc = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                      morton_bits=63, quirks=0))
    arr, bodies = serving_stream(seed, duration)
    hz = serving_stream.horizons
    srv = beatbox.Server(budget_bytes=int(budget_gib * 2**30), max_worlds=4096,
                         max_resident_bodies=resident_cap, team_bodies=team_bodies, split_bodies=split_bodies)
    t0 = time.perf_counter()
    k = 0
    while True:
        now = time.perf_counter() - t0
        if now >= duration:
            break
        while k < len(arr) and arr[k] <= now:
            srv.submit(desc, bodies[k], int(hz[k]))
            k += 1
        srv.step(1)
    rep = srv.report()
    wall = time.perf_counter() - t0
    srv.destroy()
    rep.update(duration_s=duration, wall_s=wall, arrival_rate=500.0, resident_body_cap=resident_cap,
               team_bodies=team_bodies, split_bodies=split_bodies,
               body_steps_per_s=rep["body_steps"] / (rep["device_ms"] * 1e-3) if rep["device_ms"] else 0.0)
    log(f"serving: {rep}")
    return rep


def cpu_baseline(seed: int, budget_s: float, threads: int):
    """The CPU oracle (C++, -O3, one world per call) on a bounded sample of
    the batch: the first 32*threads worlds of the same seeded batch, dealt
    round-robin to `threads` host threads (the foreign calls release the
    interpreter lock), each stepping its worlds from step 0 until the
    budget."""
    import threading

    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle_bind import OracleWorld

    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    sizes = scenes.batch_world_sizes(seed, 4096)
    nw = min(4096, 32 * threads)
    groups = [[] for _ in range(threads)]
    nb = 0
    for w in range(nw):
        o = OracleWorld(desc)
        bodies = scenes.batch_world(seed, w, int(sizes[w]))
        o.add_bodies(bodies)
        groups[w % threads].append((o, len(bodies)))
        nb += len(bodies)
    done = [0] * threads
    steps = [0] * threads
    t0 = time.perf_counter()

    def run(k):
        while time.perf_counter() - t0 < budget_s:
            for o, n in groups[k]:
                o.step(1)
                done[k] += n
            steps[k] += 1

    ts = [threading.Thread(target=run, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - t0
    return dict(value=sum(done) / dt, unit="body-steps/s", cores=threads, kind="port",
                sample=f"CPU oracle, {threads} threads: worlds 0-{nw - 1} of the seeded 4096-world batch "
                       f"({nb} bodies), {min(steps)}-{max(steps)} steps per thread from step 0, {dt:.1f}s")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0xBB04)
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch parameters, no autotuner")
    ap.add_argument("--params", default="", metavar="JSON",
                    help="launch parameters of the batch (the names of `launch_params`), instead of the "
                         "autotuner: pins one configuration across separate profiler runs")
    ap.add_argument("--no-extra", action="store_true", help="skip the large-pile, sphere-gas and serving runs")
    ap.add_argument("--extra", default="large_pile,sphere_gas,serving,next_rows",
                    help="comma-separated extra runs (large_pile, sphere_gas, serving, next_rows)")
    ap.add_argument("--no-config-window", action="store_true",
                    help="do not time the configuration's own 240-step window beside another one "
                         "(counter runs, whose per-kernel figures must come from the timed window alone)")
    ap.add_argument("--serving-seconds", type=float, default=10.0)
    ap.add_argument("--serving-budget-gib", type=float, default=200.0)
    ap.add_argument("--serving-cap", type=int, default=200_000, help="resident-body cap of the serving run")
    ap.add_argument("--serving-team-bodies", type=int, default=4096,
                    help="worlds this large are stepped as workgroup teams in the serving run")
    ap.add_argument("--serving-split-bodies", type=int, default=256,
                    help="serving pools: worlds below this size s, in [s, 4s), up to the team size, and "
                         "above step at their own cadences (0: one lockstep pool)")
    ap.add_argument("--profile-steps", type=int, default=0, help="0: the timed steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel on MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (worlds shard "
              "across GPUs without exchange; this build measures one GPU's share)", file=sys.stderr)
        return 2

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        ms = beatbox.copy_bandwidth(nbytes, 5)
        log(f"copy {args.copy_calibration} MiB: {ms:.3f} ms, {2 * nbytes / ms / 1e6:.1f} GB/s (read+write)")
        return 0

    copy_ms = beatbox.copy_bandwidth(1 << 30, 5)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"copy kernel: {copy_gbs:.0f} GB/s (practical HBM ceiling)")

    # 1) throughput: hipGraph replay of the timed steps
    batch, total = make_batch(args.worlds, args.seed, 16, 2048)
    fixed = json.loads(args.params) if args.params else None
    t, e0, e1, params, overflowed = time_batch(batch, args.steps, args.warmup, not args.no_tune, fixed)
    if overflowed:
        log(f"batch: {len(overflowed)} worlds flagged an overflow: {overflowed[:16]}")
    log(f"launch parameters: {params}")
    # worlds in the wide lane (no world of this batch reaches the team threshold)
    n_wide = sum(1 for n in batch.world_bodies if params["lane_bodies"] and n >= params["lane_bodies"])
    dev_gib = batch.device_bytes / 2**30
    batch.destroy()
    value = total * args.steps / t
    ms_per_step = t / args.steps * 1e3
    drift = (e1 - e0) / abs(e0) if e0 else 0.0
    log(f"batch: {ms_per_step:.3f} ms/step, {value / 1e6:.2f} M body-steps/s, {dev_gib:.2f} GiB, "
        f"energy {e0:.4e} J at step 0 -> {e1:.4e} J (rel. change {drift:+.3e})")

    # 2) per-stage durations over the same step window: a fresh batch of the
    # same seed reaches the same states (the step is deterministic), then
    # eager steps are timed with hipEvents on the library stream
    batch, _ = make_batch(args.worlds, args.seed, 16, 2048)
    apply_params(batch, params)
    batch.step(args.warmup, use_graph=True)
    batch.sync(allow_overflow=True)
    psteps = args.profile_steps or args.steps
    step_ms, kern_ms, counts = batch.profile(psteps, counts=True)
    batch.destroy()
    p50 = float(np.percentile(step_ms, 50))
    p99 = float(np.percentile(step_ms, 99))
    sb = stage_bytes(counts, 8, 8)
    per_launch_ms = {n: float(kern_ms[i]) / psteps for i, n in enumerate(beatbox.KERNEL_NAMES)}
    dominant = max(per_launch_ms, key=per_launch_ms.get)
    achieved = sb[dominant] / (float(kern_ms[beatbox.KERNEL_NAMES.index(dominant)]) * 1e-3) / 1e9
    stages = {n: dict(ms=round(per_launch_ms[n], 4), model_bytes=float(sb[n]) / psteps,
                      gbs=round(float(sb[n]) / (float(kern_ms[i]) * 1e-3) / 1e9, 1) if kern_ms[i] > 0 else None)
              for i, n in enumerate(beatbox.KERNEL_NAMES)}
    avg = {n: float(counts[:, i].mean()) for i, n in enumerate(beatbox.COUNT_NAMES)}
    max_pen = int(counts[:, beatbox.COUNT_NAMES.index("max_pen_mm")].max())
    log(f"stages (avg per launch over {psteps} eager steps): {stages}")
    log(f"eager step p50 {p50:.3f} ms, p99 {p99:.3f} ms; avg counts {avg}")

    plain_bs, plain_ms = plain_world(args.steps, args.warmup)
    log(f"plain (steps {args.warmup}..{args.warmup + args.steps}): {plain_ms:.3f} ms/step, "
        f"{plain_bs / 1e6:.3f} M body-steps/s, ratio {value / plain_bs:.1f}x")

    # the configuration's own window (steps 10..250) when the timed window is
    # another one: batch and plain world again, same launch parameters
    config_window = None
    if (args.steps, args.warmup) != (CONFIG_STEPS, CONFIG_WARMUP) and not args.no_config_window:
        cb, _ = make_batch(args.worlds, args.seed, 16, 2048)
        apply_params(cb, params)
        ct, _, _, _, c_ovf = time_batch(cb, CONFIG_STEPS, CONFIG_WARMUP, False)
        cb.destroy()
        c_plain_bs, c_plain_ms = plain_world(CONFIG_STEPS, CONFIG_WARMUP)
        cv = total * CONFIG_STEPS / ct
        config_window = dict(steps=CONFIG_STEPS, warmup=CONFIG_WARMUP, value=cv, ms_per_step=ct / CONFIG_STEPS * 1e3,
                             plain_body_steps_per_s=c_plain_bs, plain_ms_per_step=c_plain_ms,
                             batch_over_plain=cv / c_plain_bs, overflow_worlds=len(c_ovf))
        log(f"config window: {config_window}")

    extra = {}
    runs = set() if args.no_extra else {x.strip() for x in args.extra.split(",") if x.strip()}
    if "large_pile" in runs:
        avbd = beatbox.WorldDesc.make(**scenes.world_desc(solver=scenes.SOLVER_AVBD, iterations=10,
                                                          morton_bits=63, quirks=0))
        extra["large_pile"] = dict(workload="100,000 bodies (50/50 boxes/spheres, 0.25-1 m) falling in a "
                                            "60x200x60 m column, AVBD 10 iterations, 63-bit keys",
                                   # tuned on its second step: the tuner picks the team launch's size
                                   # (all resident workgroups, 11.2 -> 10.8 ms/step over the 300 steps)
                                   **single_world("large pile", avbd, scenes.large_pile(), 300, 1, 5))
    if "sphere_gas" in runs:
        gas = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0, gravity=0.0))
        extra["sphere_gas"] = dict(workload="1,000,000 spheres (r 0.1-0.5 m) in a 1000 m cube, v ~ N(0, 2 m/s), "
                                            "no gravity, PGS_SOFT 8 iterations, 63-bit keys",
                                   **single_world("sphere gas", gas, scenes.sphere_gas(), 100, 1, 5,
                                                  tune=not args.no_tune))
    if "serving" in runs:
        extra["serving"] = serving(args.serving_seconds, args.serving_budget_gib, args.serving_cap,
                                   args.serving_team_bodies, args.serving_split_bodies)
        extra["serving_budget_only"] = serving(args.serving_seconds, args.serving_budget_gib, 0,
                                               args.serving_team_bodies, args.serving_split_bodies)
    if "next_rows" in runs:
        extra["next_rows"] = next_rows(args.seed)

    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    cpu = None if args.no_cpu else cpu_baseline(args.seed, args.cpu_budget, threads)
    if cpu:
        log(f"cpu baseline: {cpu['value'] / 1e6:.3f} M body-steps/s")

    out = {
        "metric": "body-steps/s batched vs plain at fixed pose tol; p50/p99 step latency",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded box pyramids and towers)",
        "config": {
            "workload": f"batched-worlds: {args.worlds} worlds, bodies log-uniform [16,2048], box pyramids/towers, PGS_SOFT I=8, hipGraph, steps {args.warmup}..{args.warmup + args.steps}",
            "worlds": args.worlds,
            "bodies": total,
            "avg_pairs": avg["pairs"],
            "avg_manifolds": avg["manifolds"],
            "avg_contacts": avg["contacts"],
            "max_penetration_mm": max_pen,
            "energy_rel_change": drift,
            "device_gib": dev_gib,
            "plain_body_steps_per_s": plain_bs,
            "plain_ms_per_step": plain_ms,
            "plain_window": f"steps {args.warmup}..{args.warmup + args.steps}, the batch's window",
            "batch_over_plain": value / plain_bs,
            "overflow_worlds": len(overflowed),
            "pair_scale": PAIR_SCALE,
            "config_window": config_window,
            "eager_step_ms_p50": p50,
            "eager_step_ms_p99": p99,
            "copy_kernel_gbs": copy_gbs,
            "launch_params": params,
            **extra,
            "stages": stages,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            **traffic_fields(dominant, args, sb[dominant] / psteps, params, n_wide),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
