"""Shared test helpers: the engine package, small network configs, oracle construction."""
from __future__ import annotations

import copy
import importlib
import json
import os
import sys

import numpy as np

TESTS = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(TESTS)
for p in (REPO, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle_py  # noqa: E402  (the oracle: test infrastructure)

PKG_NAME = "example-cuda-spiking-nn_amd"
CONFIGS = os.path.join(REPO, "configs")


def engine():
    return importlib.import_module(PKG_NAME)


def load_config(name: str) -> dict:
    return json.load(open(os.path.join(CONFIGS, name)))


def small_config(n_e_side=20, n_i_side=10, steps=3001, stimulus=True, plastic=True, seed=3) -> dict:
    """The 3,125-neuron workload's model on a smaller grid (same 1 mm square scaled down)."""
    c = copy.deepcopy(load_config("net3125_stdp.json"))
    side = float(n_e_side)
    c.update({"num_excitatory_neurons": n_e_side * n_e_side, "num_inhibitory_neurons": n_i_side * n_i_side,
              "excitatory_side_points": n_e_side, "inhibitory_side_points": n_i_side,
              "inhibitory_location_spacing": side / n_i_side, "side_length": side, "rng_seed": seed,
              "simulation_steps": steps, "output_directory": "outputs/test_small"})
    for p in ("e_to_e", "e_to_i", "i_to_e", "i_to_i"):
        c[p + "_x_location_length"] = side
        c[p + "_y_location_length"] = side
        c[p + "_max_distance"] = side
        c[p + "_spread_scale"] = 6.0
        c[p + "_probability_multiplier"] = 0.5
    c["group_e_background_excitatory_scaling_factor"] = 0.7e-9
    c["group_i_background_excitatory_scaling_factor"] = 0.7e-9
    c["voltage_recording_neuron_ids"] = [0, 5, n_e_side * n_e_side - 1]
    if stimulus:
        c["stimulus_selected_input_ids"] = [1, 2, 3, 40, 41, 0]
        c["stimulus_selected_input_post_group_offsets"] = [0, 0, 0, 0, 0, n_e_side * n_e_side]
        c["stimulus_start_clock_steps"] = [20, 150]
        c["stimulus_end_clock_steps"] = [60, 170]
        c["stimulus_input_values"] = [3e-10, 6e-10]
    if not plastic:
        c["e_e_plasticity_enabled"] = False
        c["i_e_plasticity_enabled"] = False
    else:
        # Larger learning rates than the workload so that weight changes are visible in short runs.
        c["e_e_plasticity_ltp_A"] = 2e-12
        c["e_e_plasticity_ltd_A"] = 1e-12
        c["e_e_plasticity_ltp_B"] = 5e-12
        c["e_e_plasticity_ltd_B"] = 1e-12
        c["i_e_plasticity_ltp_A"] = 2e-12
        c["i_e_plasticity_ltd_A"] = 1e-12
    return c


def stimulus_schedule(cfg: dict):
    """Event tables of the square-wave schedule (generate_stimulus_schedules.cpp:107-196)."""
    T = cfg["simulation_steps"]
    ids, offs = cfg["stimulus_selected_input_ids"], cfg["stimulus_selected_input_post_group_offsets"]
    starts, ends = cfg["stimulus_start_clock_steps"], cfg["stimulus_end_clock_steps"]
    amps = list(cfg["stimulus_input_values"])
    if len(amps) == 1 and len(starts) != 1:
        amps = amps * len(starts)
    seen, sel = set(), []
    for i, o in zip(ids, offs):
        if (o, i) not in seen:
            seen.add((o, i))
            sel.append((o, i))
    ev = []
    for o, i in sel:
        for s, e, a in zip(starts, ends, amps):
            if s < T:
                ev.append((s, o, i, np.float32(a)))
            if e + 1 < T:
                ev.append((e + 1, o, i, -np.float32(a)))
    ev.sort(key=lambda t: (t[0], t[1], t[2]))
    counts = np.zeros(T, np.uint32)
    for t, *_ in ev:
        counts[t] += 1
    return (counts, np.array([e[2] for e in ev], np.uint32), np.array([e[1] for e in ev], np.uint32),
            np.array([e[3] for e in ev], np.float32))


def oracle_for(config_path: str, sim=None, precision=32, lazy=False, row_block=None):
    """Oracle on the same config, connectivity and GEMV chunking as `sim`."""
    eng = engine()
    cfg = eng.config_dump(config_path)
    if row_block is None:
        row_block = tuple(sim.get_info().gemv_row_block) if sim is not None else (0, 0, 0, 0)
    # The persistent step's orders: chunk sums in parts of 8, lazy row sums by column tile; the
    # resident step's: rows over lanes and waves, lazy row sums by blocks of columns.
    info = sim.get_info() if sim is not None else None
    persistent = info is not None and info.persistent_blocks > 0
    resident = (info.resident_columns, info.resident_e_waves) if info is not None and info.resident_blocks else (0, 0)
    p = oracle_py.params_from_config(cfg, gemv_row_block=row_block, lazy_normalize=lazy,
                                     gemv_part_chunks=(8, 8, 8, 8) if persistent else (0, 0, 0, 0),
                                     lazy_tile_order=persistent, resident=resident)
    weights = [eng.generate_weight_matrix(config_path, k) for k in range(4)]
    stim = stimulus_schedule(cfg)
    return oracle_py.OracleSim(p, weights, precision=precision, stim=stim)


def assert_state_equal(g: dict, o: dict, step: int, fields=None):
    pairs = [("membrane_potentials", "V"), ("g_excitatory", "ge"), ("g_inhibitory", "gi"), ("i_stimulus", "istim"),
             ("refractory_counter", "refr"), ("spikes", "spikes"),
             ("background_g_excitatory_synaptic_conductances", "bg_ge"),
             ("background_g_inhibitory_synaptic_conductances", "bg_gi")]
    for a, b in pairs:
        if fields and a not in fields:
            continue
        if not np.array_equal(g[a], o[b]):
            bad = np.nonzero(g[a] != o[b])[0]
            raise AssertionError(f"step {step}: {a} differs at {bad.size} neurons, first {bad[:5]}: "
                                 f"gpu {g[a][bad[:3]]} oracle {o[b][bad[:3]]}")
