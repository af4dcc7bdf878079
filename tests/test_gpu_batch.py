"""Batched shared-weight form (bn_batch_*): each instance against a plain run of the same
network with the instance's noise key and stimulus pulse.

The propagate of the batch is an FP32 matrix-core GEMM (FP32 accumulation) where the plain
step accumulates in FP64, so the bar is the floating-point one: identical spike rasters and
per-step V, g within 1e-5 relative (the state, the noise, the stimulus and every other
operation are the same arithmetic).
"""
import json
import os

import numpy as np
import pytest

from helpers import engine, load_config, small_config

pytestmark = pytest.mark.gpu
REL = 1e-5


def _write(tmp_path, cfg, name):
    p = os.path.join(str(tmp_path), name)
    json.dump(cfg, open(p, "w"))
    return p


def _instances(n_e, count, n_ids, start, end, seed0=7):
    out = []
    for k in range(count):
        ids, amp = engine().batch_workload_instance(k, n_e, n_ids, 200e-12)
        out.append({"noise_seed": seed0 + k, "stimulus_ids": ids, "stimulus_amplitude": amp,
                    "stimulus_start": start, "stimulus_end": end})
    return out


def _plain_config(base, inst):
    cfg = dict(base)
    cfg["background_noise_seed"] = int(inst["noise_seed"])
    ids = [int(i) for i in inst["stimulus_ids"]]
    cfg["stimulus_selected_input_ids"] = ids
    cfg["stimulus_selected_input_post_group_offsets"] = [0] * len(ids)
    cfg["stimulus_start_clock_steps"] = [int(inst["stimulus_start"])]
    cfg["stimulus_end_clock_steps"] = [int(inst["stimulus_end"])]
    cfg["stimulus_input_values"] = [float(np.float32(inst["stimulus_amplitude"]))]
    return cfg


def _rel(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b.astype(np.float64)), 1e-12)))


def _compare_lockstep(batch, k, plain, steps):
    worst_v = worst_g = 0.0
    for step in range(steps):
        batch.run(1)
        plain.run(1)
        b, p = batch.state(k), plain.state()
        assert np.array_equal(b["spikes"], p["spikes"]), f"instance {k}: raster differs at step {step}"
        assert np.array_equal(b["i_stimulus"], p["i_stimulus"]), f"instance {k}: stimulus differs at step {step}"
        assert np.array_equal(b["background_excitatory_postsynaptic_d_traces"],
                              p["background_excitatory_postsynaptic_d_traces"]), f"noise differs at step {step}"
        worst_v = max(worst_v, _rel(b["membrane_potentials"], p["membrane_potentials"]))
        for g in ("g_excitatory", "g_inhibitory"):
            worst_g = max(worst_g, float(np.max(np.abs(b[g].astype(np.float64) - p[g]) /
                                                 np.maximum(np.abs(p[g].astype(np.float64)), 1e-12))))
    return worst_v, worst_g


def test_batch_instances_match_plain_runs(tmp_path):
    base = small_config(steps=400, stimulus=False, plastic=False)
    path = _write(tmp_path, base, "base.json")
    n_e = base["num_excitatory_neurons"]
    insts = _instances(n_e, 5, 20, start=20, end=60)
    batch = engine().BatchSimulation(path, insts)
    assert batch.info.padded_instances == 64
    plains = [engine().Simulation(_write(tmp_path, _plain_config(base, insts[k]), f"p{k}.json"),
                                  flags=("no_monitors", "quiet")) for k in (0, 4)]
    # Lock-step instance 0, then instance 4 on a second batch (the state download is per instance).
    wv, wg = _compare_lockstep(batch, 0, plains[0], 150)
    assert wv <= REL and wg <= REL, (wv, wg)
    batch.close()
    batch = engine().BatchSimulation(path, insts)
    wv, wg = _compare_lockstep(batch, 4, plains[1], 150)
    assert wv <= REL and wg <= REL, (wv, wg)
    for s in plains:
        s.close()
    batch.close()


def test_batch_instances_are_independent(tmp_path):
    # The same instance in a batch of 1 and at another column of a batch of 70 (two n-pairs):
    # identical state, since every column of the GEMM is summed in the same order.
    base = small_config(steps=400, stimulus=False, plastic=False)
    path = _write(tmp_path, base, "base.json")
    insts = _instances(base["num_excitatory_neurons"], 70, 10, start=5, end=40)
    big = engine().BatchSimulation(path, insts)
    one = engine().BatchSimulation(path, [insts[66]])
    assert big.info.padded_instances == 128 and one.info.padded_instances == 64
    big.run(230)  # two graphs of 100 steps and an eager remainder
    one.run(230)
    a, b = big.state(66), one.state(0)
    for key in ("membrane_potentials", "g_excitatory", "g_inhibitory", "spikes", "i_stimulus"):
        assert np.array_equal(a[key], b[key]), key
    big.close()
    one.close()


def test_batch_graphs_equal_eager_steps(tmp_path):
    # Graph-captured steps (graphs of 100) and the same steps launched one by one.
    base = small_config(steps=400, stimulus=False, plastic=False)
    path = _write(tmp_path, base, "base.json")
    insts = _instances(base["num_excitatory_neurons"], 6, 20, start=30, end=150)
    a = engine().BatchSimulation(path, insts)
    b = engine().BatchSimulation(path, insts)
    a.run(205)
    for _ in range(205):
        b.run(1)
    for k in (0, 5):
        sa, sb = a.state(k), b.state(k)
        for key in ("membrane_potentials", "g_excitatory", "g_inhibitory", "spikes", "i_stimulus"):
            assert np.array_equal(sa[key], sb[key]), (k, key)
    a.close()
    b.close()


@pytest.mark.parametrize("rows", [8, 24, 512])
def test_batch_gemm_tilings(tmp_path, rows):
    base = small_config(steps=200, stimulus=False, plastic=False)
    path = _write(tmp_path, base, "base.json")
    insts = _instances(base["num_excitatory_neurons"], 3, 20, start=10, end=50)
    batch = engine().BatchSimulation(path, insts)
    batch.set_gemm_rows(rows)
    assert batch.info.gemm_rows_per_wave == rows
    plain = engine().Simulation(_write(tmp_path, _plain_config(base, insts[2]), "p.json"),
                                flags=("no_monitors", "quiet"))
    wv, wg = _compare_lockstep(batch, 2, plain, 80)
    assert wv <= REL and wg <= REL, (wv, wg)
    plain.close()
    batch.close()


def test_batch_workload_3125_two_instances(tmp_path):
    # The shared-weight workload's network (the 3,125-neuron network, plasticity off, 64
    # instances, 100 stimulated E neurons each): instances 0 and 63 over 200 steps, with the
    # pulse moved to steps 50-120 so that it falls inside the window.
    base = load_config("net3125_stdp.json")
    base["e_e_plasticity_enabled"] = False
    base["i_e_plasticity_enabled"] = False
    path = _write(tmp_path, base, "base.json")
    insts = _instances(base["num_excitatory_neurons"], 64, 100, start=50, end=120, seed0=0)
    for k in (0, 63):
        batch = engine().BatchSimulation(path, insts)
        plain = engine().Simulation(_write(tmp_path, _plain_config(base, insts[k]), f"p{k}.json"),
                                    flags=("no_monitors", "quiet"))
        wv, wg = _compare_lockstep(batch, k, plain, 200)
        assert wv <= REL and wg <= REL, (k, wv, wg)
        plain.close()
        batch.close()


def test_batch_rejects_plastic_networks(tmp_path):
    path = _write(tmp_path, small_config(steps=50), "plastic.json")
    with pytest.raises(engine().BnError, match="frozen weights"):
        engine().BatchSimulation(path, [{"noise_seed": 1}])
