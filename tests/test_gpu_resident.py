"""GPU: the resident step (BN_SIM_RESIDENT) against the oracle in its summation orders.

One cooperative launch per run of steps with every weight in registers: each block owns 16 post
columns, all pre rows of them, and evolves those neurons; rows' traces are advanced from the
spike bits by every thread that holds the row.  The oracle restates the resident orders (rows
over lanes and waves, lane butterfly, lazy row sums by blocks of 16 columns), so the bar is
bit-identical: the raster of every step (device spike ring), the state after every launch,
traces and weights, for launches of 1 to 1,000 steps.
"""
import json
import os

import numpy as np
import pytest

from helpers import assert_state_equal, engine, load_config, oracle_for, small_config

pytestmark = pytest.mark.gpu


def _write(tmp_path, cfg, name="cfg.json"):
    p = os.path.join(tmp_path, name)
    json.dump(cfg, open(p, "w"))
    return p


def _resident(path, lazy=True):
    flags = ("quiet", "resident") + (("lazy_normalize",) if lazy else ())
    sim = engine().Simulation(path, flags=flags)
    info = sim.get_info()
    assert info.resident_blocks > 0 and info.resident_columns == 16 and info.kernels_per_step == 1
    print(f"resident: {info.num_excitatory_neurons} E + {info.num_inhibitory_neurons} I, {info.resident_blocks} "
          f"blocks, {info.resident_rows} rows per thread, {info.resident_e_waves} waves of E rows", flush=True)
    return sim


def _check_traces_and_weights(sim, orc):
    for k in range(4):
        for a, b, name in zip(sim.connection_traces(k), orc.connection_traces(k), ("r", "d", "x")):
            assert np.array_equal(a, b), f"connection {k} trace {name}"
    for off in range(sim.info.num_plasticities):
        for a, b, name in zip(sim.plasticity_traces(off), orc.plasticity_traces(off),
                              ("pre", "post", "triplet_pre", "triplet_post")):
            assert np.array_equal(a, b), f"plasticity {off} trace {name}"
    for k in range(4):
        w_gpu, w_orc = sim.weights(k), orc.weights(k)
        assert np.array_equal(w_gpu, w_orc), f"weights {k}: {np.count_nonzero(w_gpu != w_orc)} entries differ"


def _launches(sim, orc, sizes, first_step=1):
    """Launches of the given sizes; after each, the raster of its steps and the state."""
    step = first_step
    spikes = 0
    for n in sizes:
        sim.run(n)
        raster = []
        for _ in range(n):
            orc.step(1)
            raster.append(orc.state()["spikes"])
        raster = np.array(raster)
        got = sim.spike_raster(step, n)
        if not np.array_equal(got, raster):
            bad = np.nonzero((got != raster).any(axis=1))[0]
            raise AssertionError(f"raster differs at steps {step + bad[:5]} (launch of {n} from step {step})")
        step += n
        spikes += int(raster.sum())
        assert_state_equal(sim.state(), orc.state(), step - 1)
    return spikes


@pytest.mark.parametrize("which", ["plastic", "burst", "no-plasticity", "istdp-only", "pair-stdp"])
def test_resident_small_bit_exact(tmp_path, which):
    cfg = small_config(steps=800, plastic=which != "no-plasticity")
    if which == "burst":
        cfg["stimulus_selected_input_ids"] = list(range(0, 400, 2))
        cfg["stimulus_selected_input_post_group_offsets"] = [0] * 200
        cfg["stimulus_start_clock_steps"] = [30, 95]
        cfg["stimulus_end_clock_steps"] = [40, 101]
        cfg["stimulus_input_values"] = [4e-9, 3e-9]
    if which == "istdp-only":
        cfg["e_e_plasticity_enabled"] = False
    if which == "pair-stdp":
        cfg["e_e_plasticity_rule"] = "pair"  # no normalised matrix
    path = _write(tmp_path, cfg)
    print(f"case {which}", flush=True)
    sim = _resident(path)
    orc = oracle_for(path, sim, lazy=True)
    sim.run(1)  # step 0 is a record step and runs as a launch of its own
    orc.step(1)
    assert_state_equal(sim.state(), orc.state(), 0)
    spikes = _launches(sim, orc, [1, 2, 3, 7, 50, 100, 137, 200])
    _check_traces_and_weights(sim, orc)
    assert spikes > 0
    assert sim.get_info().graph_launches == 0
    sim.close()


def test_resident_ordinary_launch_bit_exact(tmp_path):
    # The same kernel launched without the cooperative API (the device is otherwise idle, so
    # every block is resident): the same raster, state, traces and weights.
    path = _write(tmp_path, small_config(steps=800))
    sim = _resident(path)
    sim.set_resident_cooperative(False)
    orc = oracle_for(path, sim, lazy=True)
    sim.run(1)
    orc.step(1)
    spikes = _launches(sim, orc, [1, 20, 200, 137])
    _check_traces_and_weights(sim, orc)
    assert spikes > 0
    sim.close()


@pytest.mark.parametrize("n_e_side", [20, 30, 40, 50])
def test_resident_network_sizes_bit_exact(tmp_path, n_e_side):
    # The sizes of the uneven batch (400 .. 2,500 E neurons, 2 .. 7 rows per thread).
    cfg = small_config(n_e_side=n_e_side, n_i_side=n_e_side // 2, steps=400)
    ne = n_e_side * n_e_side
    cfg["stimulus_selected_input_ids"] = list(range(0, ne, 3))
    cfg["stimulus_selected_input_post_group_offsets"] = [0] * len(cfg["stimulus_selected_input_ids"])
    cfg["stimulus_start_clock_steps"] = [20, 140]
    cfg["stimulus_end_clock_steps"] = [35, 150]
    cfg["stimulus_input_values"] = [4e-9, 3e-9]
    path = _write(tmp_path, cfg)
    sim = _resident(path)
    orc = oracle_for(path, sim, lazy=True)
    sim.run(1)
    orc.step(1)
    spikes = _launches(sim, orc, [120, 79])
    _check_traces_and_weights(sim, orc)
    assert spikes > 0
    sim.close()


def test_resident_workload_3125(tmp_path):
    # The benchmarked network: launches of 1,000 and 499 steps, every step's raster.
    path = _write(tmp_path, load_config("net3125_stdp.json"))
    sim = _resident(path)
    orc = oracle_for(path, sim, lazy=True)
    sim.run(1)
    orc.step(1)
    spikes = _launches(sim, orc, [1000, 499])
    _check_traces_and_weights(sim, orc)
    assert spikes > 1000
    sim.close()


def test_resident_snapshot_resume(tmp_path):
    path = _write(tmp_path, small_config(steps=600))
    a = _resident(path)
    a.run(1)
    a.run(120)
    snap = os.path.join(tmp_path, "snap.bin")
    a.save_state(snap)
    a.run(90)
    b = _resident(path)
    b.load_state(snap)
    b.run(90)
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(a.weights(0), b.weights(0))
    # The same simulation restored to the earlier step (its exchange words are cleared).
    a.load_state(snap)
    a.run(90)
    for k, v in a.state().items():
        assert np.array_equal(v, sb[k]), k
    a.close()
    b.close()


def test_resident_refuses_a_network_too_large(tmp_path):
    path = _write(tmp_path, load_config("net10125_stdp.json"))
    with pytest.raises(engine().BnError, match="does not fit the resident step"):
        engine().Simulation(path, flags=("quiet", "resident", "lazy_normalize"))


def test_resident_wait_limit_abandons_the_launch(tmp_path):
    # A wait limit of zero polls: a block that finds a word of another block missing gives up,
    # every block leaves, and the error surfaces at the next synchronisation (the host forces the
    # condition; every kernel still exits normally).
    path = _write(tmp_path, small_config(steps=600))
    sim = _resident(path)
    sim.run(1)
    sim.set_persistent_wait_limit(0)
    with pytest.raises(engine().BnError, match="waited past its limit"):
        for _ in range(20):
            sim.run(25)
            sim.synchronize()
    sim.close()
