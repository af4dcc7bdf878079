"""Workload definitions of the batched forms (configs written from the 3,125-neuron network's
parameters, configs/net3125_stdp.json).

- ``uneven_plastic_configs``: independent plastic networks of uneven size.  N_E is drawn from
  {400, 900, 1600, 2500} by Python's Mersenne Twister seeded with ``seed`` (the recorded draw);
  N_I = N_E / 4; E on a sqrt(N_E) grid at spacing 1, I on a sqrt(N_I) grid at spacing 2, side
  sqrt(N_E), Gaussian spread S = 25, no distance cut-off; network k uses rng_seed = k (its own
  connectivity and noise).  Triplet STDP on E->E and iSTDP on I->E as in the 3,125-neuron net.
- ``adjoint_workload``: the differentiable window on the 3,125-neuron network with plasticity
  off: the state after a warm-up run, 50 recorded E neurons (every N_E/50-th id), targets
  (smoothed rates and recorded voltages) produced by the same window with the E->E weights
  scaled by 1.1, surrogate slope 1/(1 mV), checkpoints every 100 steps.
- ``shared_batch_instances``: the shared-weight batch's instances (noise key = instance, a
  200 pA x U(0,1) pulse on 100 E neurons from step 2,000 to 4,000, the engine's recorded
  per-instance draw).
"""
from __future__ import annotations

import copy
import json
import math
import os
import random

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIZES = (400, 900, 1600, 2500)


def base_config() -> dict:
    return json.load(open(os.path.join(REPO, "configs", "net3125_stdp.json")))


def network_config(n_e: int, seed: int, steps: int = 10001) -> dict:
    """The 3,125-neuron network's model on a sqrt(n_e) x sqrt(n_e) excitatory grid."""
    side = int(round(math.sqrt(n_e)))
    if side * side != n_e or side % 2:
        raise ValueError("n_e must be the square of an even number")
    c = copy.deepcopy(base_config())
    c.update({"num_excitatory_neurons": n_e, "num_inhibitory_neurons": n_e // 4,
              "excitatory_side_points": side, "inhibitory_side_points": side // 2,
              "excitatory_location_spacing": 1.0, "inhibitory_location_spacing": 2.0,
              "side_length": float(side), "rng_seed": seed, "simulation_steps": steps,
              "output_directory": f"outputs/uneven_{n_e}_{seed}",
              "voltage_recording_neuron_ids": [0]})
    for p in ("e_to_e", "e_to_i", "i_to_e", "i_to_i"):
        c[p + "_x_location_length"] = float(side)
        c[p + "_y_location_length"] = float(side)
        c[p + "_max_distance"] = float(side)
        c[p + "_spread_scale"] = 25.0
    return c


def uneven_sizes(count: int = 48, seed: int = 2024) -> list[int]:
    rng = random.Random(seed)
    return [SIZES[rng.randrange(len(SIZES))] for _ in range(count)]


def uneven_plastic_configs(count: int = 48, seed: int = 2024, steps: int = 10001) -> list[dict]:
    return [network_config(n_e, k, steps) for k, n_e in enumerate(uneven_sizes(count, seed))]


def shared_batch_instances(engine, n_e: int, count: int = 64, start: int = 2000, end: int = 4000) -> list[dict]:
    out = []
    for k in range(count):
        ids, amp = engine.batch_workload_instance(k, n_e, 100, 200e-12)
        out.append({"noise_seed": k, "stimulus_ids": ids, "stimulus_amplitude": amp,
                    "stimulus_start": start, "stimulus_end": end})
    return out


def autotune_row_blocks(engine, sizes, workdir: str, candidates=(32, 64, 128, 256, 512), reps: int = 4,
                        warmup: int = 100, steps: int = 200) -> dict:
    """The batched plastic step's tile autotuner: for each network size, time the multi-network
    step of `reps` networks of that size with every candidate GEMV row block (rows summed per
    propagate partial; it also fixes the FP64 summation order, so the choice is written into
    each network's config as gemv_row_blocks and a plain run of that config reproduces the
    network bit for bit).  Returns {n_e: (best_rows, {rows: us_per_step})}."""
    out = {}
    for n_e in sorted(set(sizes)):
        times = {}
        for rb in candidates:
            if rb > n_e // 4 and rb != candidates[0]:
                continue
            paths = []
            for r in range(reps):
                cfg = network_config(n_e, 1000 + r, steps=warmup + steps + 1)
                cfg["gemv_row_blocks"] = [rb] * 4
                path = os.path.join(workdir, f"tune_{n_e}_{rb}_{r}.json")
                json.dump(cfg, open(path, "w"))
                paths.append(path)
            multi = engine.MultiSimulation(paths, flags=("lazy_normalize",))
            multi.run_timed(warmup)
            times[rb] = multi.run_timed(steps) * 1000.0 / steps
            multi.close()
        best = min(times, key=times.get)
        out[n_e] = (best, times)
    return out


def plain_row_block_candidates(default: list) -> list:
    """GEMV row blocks (EE, EI, IE, II) tried for one plain network besides the engine's default:
    longer chunks on the excitatory-pre matrices (their 2,500+ rows) and the same or shorter
    chunks on the inhibitory-pre ones.  Measured on the 3,125-neuron network in graphs of 500
    steps: [48, 48, 32, 32] 15.02, [40, 40, 10, 10] 15.15, [40, 40, 16, 16] 15.33, the default
    [32] * 4 16.30 us/step."""
    out = []
    for fe, fi in ((1.5, 1.0), (1.25, 0.3125), (1.25, 0.5), (1.5, 0.5)):
        c = [max(1, int(round(default[0] * fe))), max(1, int(round(default[1] * fe))),
             max(1, int(round(default[2] * fi))), max(1, int(round(default[3] * fi)))]
        if c != list(default) and c not in out:
            out.append(c)
    return out


def autotune_plain_row_blocks(engine, config_path: str, workdir: str, flags, settle: int = 1000,
                              steps: int = 3000, steps_per_graph: int = 500) -> tuple:
    """The plain step's tile autotuner: the network as configured (the engine's default row
    blocks unless the config names them), then each of plain_row_block_candidates of that
    default, each built, settled and timed for `steps` steps in graphs.  The row block fixes the
    FP64 summation order of the propagate, so a choice is written into a copy of the config as
    gemv_row_blocks and the oracle reproduces that run bit for bit.
    Returns (best_config_path, best_rows, {str(rows): us_per_step})."""
    cfg = json.load(open(config_path))
    times, paths = {}, {}

    def run(path):
        sim = engine.Simulation(path, flags=flags)
        sim.set_steps_per_graph(steps_per_graph)
        sim.run_timed(settle)
        us = sim.run_timed(steps) * 1000.0 / steps
        rows = [int(r) for r in sim.info.gemv_row_block]
        sim.close()
        times[str(rows)] = us
        paths[str(rows)] = (path, rows)
        return rows

    default = run(config_path)
    if "gemv_row_blocks" not in cfg:
        for rows in plain_row_block_candidates(default):
            c = dict(cfg)
            c["gemv_row_blocks"] = rows
            path = os.path.join(workdir, "rows_" + "_".join(str(r) for r in rows) + ".json")
            json.dump(c, open(path, "w"))
            run(path)
    best = min(times, key=times.get)
    return paths[best][0], paths[best][1], times


def row_block_for(n_e: int, cap: int) -> int:
    """The GEMV row block of a network with n_e excitatory neurons under a cap: the largest power
    of two <= min(cap, n_e / 4) (at least 8)."""
    limit = max(8, min(cap, n_e // 4))
    rb = 8
    while rb * 2 <= limit:
        rb *= 2
    return rb


def autotune_row_block_cap(engine, cfgs: list, workdir: str, caps=(32, 64, 128, 256), warmup: int = 100,
                           steps: int = 300, streams: int = 4, steps_per_graph: int = 50) -> tuple:
    """The batched plastic step's tile autotuner in its own context: for each cap, every network
    gets row_block_for(n_e, cap) and the whole multi-network step of all networks is timed (the
    propagate of many networks is bandwidth-bound, where single-size tuning favours small
    blocks).  The row block fixes the FP64 summation order, so it is written into each config
    (gemv_row_blocks) and a plain run of that config reproduces the network bit for bit.
    Returns (best_cap, {cap: us_per_step})."""
    times = {}
    for cap in caps:
        paths = []
        for k, c in enumerate(cfgs):
            c = dict(c)
            c["gemv_row_blocks"] = [row_block_for(c["num_excitatory_neurons"], cap)] * 4
            paths.append(os.path.join(workdir, f"cap{cap}_{k}.json"))
            json.dump(c, open(paths[-1], "w"))
        multi = engine.MultiSimulation(paths, flags=("lazy_normalize",))
        multi.set_streams(streams, steps_per_graph, False)
        multi.run_timed(warmup)
        times[cap] = multi.run_timed(steps) * 1000.0 / steps
        multi.close()
    best = min(times, key=times.get)
    return best, times


def autotune_schedule(multi, group_mib=(0, 256, 128), steps_per_graph=(10, 50, 100), streams=(1, 2, 3, 4),
                      overlap=(False, True), steps: int = 200) -> dict:
    """The batched plastic step's schedule autotuner: time `steps` steps of every network under
    each candidate — (weight MiB per network group stepped one after the other, steps per
    graph), or (concurrent network sets on their own streams, steps per graph, plasticity of
    step t-1 overlapped with the evolve of step t or not) — and keep the fastest on `multi`.
    Results do not depend on the schedule (the networks are independent)."""
    times = {}
    for spg in steps_per_graph:
        for mib in group_mib:
            if mib == 0:
                continue  # the one-group schedule is streams = 1 below
            multi.set_schedule(mib << 20, spg)
            multi.run_timed(spg)
            times[("groups", mib, spg, False)] = multi.run_timed(steps) * 1000.0 / steps
        for ns in streams:
            for ov in overlap:
                multi.set_streams(ns, spg, ov)
                multi.run_timed(spg)
                times[("streams", ns, spg, ov)] = multi.run_timed(steps) * 1000.0 / steps
    best = min(times, key=times.get)
    if best[0] == "groups":
        multi.set_schedule(best[1] << 20, best[2])
    else:
        multi.set_streams(best[1], best[2], best[3])

    def label(k):
        if k[0] == "groups":
            return f"{k[1]}MiB/{k[2]}"
        return f"{k[1]}streams/{k[2]}" + ("/overlap" if k[3] else "")

    return {"kind": best[0], "group_mib": best[1] if best[0] == "groups" else 0,
            "streams": best[1] if best[0] == "streams" else 1, "steps_per_graph": best[2], "overlap": best[3],
            "us_per_step": {label(k): round(v, 2) for k, v in times.items()}}


def adjoint_config(base: dict | None = None) -> dict:
    c = copy.deepcopy(base if base is not None else base_config())
    c["e_e_plasticity_enabled"] = False
    c["i_e_plasticity_enabled"] = False
    return c


def adjoint_workload(engine, config: dict | None = None, steps: int = 1000, warmup: int = 1000, n_vrec: int = 50,
                     checkpoint_interval: int = 100, surrogate_beta: float = 1e-3, target_scale: float = 1.1):
    """Returns (sim, window, base weights, targets): the window starts from the state after
    `warmup` steps and holds the targets (rates [steps][2], voltages [steps][n_vrec]) of the
    E->E x target_scale network; its parameters are back at the base weights."""
    cfg = adjoint_config(config)
    sim = engine.Simulation(cfg, flags=("no_monitors", "quiet"))
    sim.run(warmup)
    sim.synchronize()
    ne = sim.NE
    ids = [int(k * ne // n_vrec) for k in range(n_vrec)]
    win = engine.Window(sim, steps, ids, checkpoint_interval=checkpoint_interval, surrogate_beta=surrogate_beta)
    base = [win.weights(k) for k in range(4)]
    win.set_params([base[0] * target_scale, None, None, None])
    tgt = win.forward()
    win.set_params([base[0], None, None, None])
    targets = (tgt["rates"], tgt["v"].astype("float64"))
    win.set_targets(*targets)
    return sim, win, base, targets
