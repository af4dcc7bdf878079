"""GPU: the persistent step (BN_SIM_PERSISTENT) against the oracle in its summation orders.

One cooperative launch per run of steps: weights streamed once per step (or held in registers for
the whole launch) with the plasticity applied in the same pass, evolves run by the blocks that complete a column tile, lazy row sums
in column-tile order.  The oracle restates those orders (chunk sums in parts of 8, lazy row sums
by 256-column tile), so the bar is bit-identical: the raster of every step (device spike ring),
the state after every launch, traces and weights, for launches of 1 to 1,000 steps.
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


def _check_traces_and_weights(sim, orc, plastic_only=False):
    for k in range(4):
        for a, b, name in zip(sim.connection_traces(k), orc.connection_traces(k), ("r", "d", "x")):
            assert np.array_equal(a, b), f"connection {k} trace {name}"
    for off in range(sim.info.num_plasticities):
        for a, b, name in zip(sim.plasticity_traces(off), orc.plasticity_traces(off),
                              ("pre", "post", "triplet_pre", "triplet_post")):
            assert np.array_equal(a, b), f"plasticity {off} trace {name}"
    for k in ((0, 2) if plastic_only else range(4)):
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


def _persistent(path, lazy=True, resident=True):
    flags = ("quiet", "persistent") + (("lazy_normalize",) if lazy else ())
    sim = engine().Simulation(path, flags=flags)
    if not resident:
        sim.set_persistent_resident(False)
    info = sim.get_info()
    assert info.persistent_blocks > 0 and info.gemv_part_chunks == 8
    if not resident:
        assert info.persistent_resident_rows == 0
    return sim


@pytest.mark.parametrize("resident", [True, False], ids=["resident", "streamed"])
@pytest.mark.parametrize("which", ["plastic", "burst", "no-plasticity", "istdp-only", "pair-stdp"])
def test_persistent_small_bit_exact(tmp_path, which, resident):
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
    sim = _persistent(path, resident=resident)
    if resident:
        assert sim.get_info().persistent_resident_rows > 0
    orc = oracle_for(path, sim, lazy=True)
    sim.run(1)  # step 0 is a record step and runs as a launch of its own
    orc.step(1)
    assert_state_equal(sim.state(), orc.state(), 0)
    spikes = _launches(sim, orc, [1, 2, 3, 7, 50, 100, 137, 200])
    _check_traces_and_weights(sim, orc)
    assert spikes > 0
    assert sim.get_info().graph_launches == 0
    sim.close()


def test_persistent_row_choices_bit_exact(tmp_path):
    # Other chunk heights (segments of 8 x rb rows): the oracle follows each.
    path = _write(tmp_path, small_config(steps=600))
    for rows in ([2, 2, 2, 2], [5, 3, 7, 1], [24, 24, 24, 24], [64, 64, 64, 64]):
        sim = _persistent(path)
        sim.set_persistent_rows(rows)
        info = sim.get_info()
        print("rows", rows, "->", tuple(info.gemv_row_block), info.persistent_items, "items, resident rows",
              info.persistent_resident_rows)
        orc = oracle_for(path, sim, lazy=True)
        sim.run(1)
        orc.step(1)
        _launches(sim, orc, [40, 81])
        _check_traces_and_weights(sim, orc)
        sim.close()


@pytest.mark.parametrize("resident", [True, False], ids=["resident", "streamed"])
def test_persistent_workload_3125(tmp_path, resident):
    # The benchmarked network: launches of 1,000 and 499 steps, every step's raster.
    path = _write(tmp_path, load_config("net3125_stdp.json"))
    sim = _persistent(path, resident=resident)
    info = sim.get_info()
    print(f"net3125 persistent: {info.persistent_blocks} blocks, {info.persistent_items} items, rows "
          f"{tuple(info.gemv_row_block)}, resident rows {info.persistent_resident_rows}")
    assert (info.persistent_resident_rows > 0) == resident
    orc = oracle_for(path, sim, lazy=True)
    sim.run(1)
    orc.step(1)
    spikes = _launches(sim, orc, [1000, 499])
    _check_traces_and_weights(sim, orc)
    assert spikes > 1000
    sim.close()


def test_persistent_workload_10125(tmp_path):
    path = _write(tmp_path, load_config("net10125_stdp.json"))
    sim = _persistent(path)
    info = sim.get_info()
    orc = oracle_for(path, sim, lazy=True)
    sim.run(1)
    orc.step(1)
    spikes = _launches(sim, orc, [600, 400])
    print(f"net10125 persistent: {info.persistent_blocks} blocks, {info.persistent_items} items, rows "
          f"{tuple(info.gemv_row_block)}, {spikes} spikes")
    _check_traces_and_weights(sim, orc, plastic_only=True)
    assert spikes > 0
    sim.close()


def test_persistent_snapshot_resume(tmp_path):
    # Pending segment partials and the lazy sums travel with the snapshot; the resident weights
    # are in memory once a launch has ended.
    path = _write(tmp_path, small_config(steps=600))
    a = _persistent(path)
    assert a.get_info().persistent_resident_rows > 0
    a.run(1)
    a.run(120)
    snap = os.path.join(tmp_path, "snap.bin")
    a.save_state(snap)
    a.run(90)
    b = _persistent(path)
    b.load_state(snap)
    b.run(90)
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(a.weights(0), b.weights(0))
    a.close()
    b.close()


@pytest.mark.parametrize("resident", [True, False], ids=["resident", "streamed"])
def test_persistent_wait_limit_abandons_the_launch(tmp_path, resident):
    # A wait limit of zero polls: a block that finds the previous step's evolves unfinished gives
    # up, every block leaves, and the error surfaces at the next synchronisation (the host forces
    # the condition; every kernel still exits normally).
    path = _write(tmp_path, small_config(steps=600))
    sim = _persistent(path, resident=resident)
    sim.run(1)
    sim.set_persistent_wait_limit(0)
    with pytest.raises(engine().BnError, match="waited past its limit"):
        for _ in range(20):
            sim.run(25)
            sim.synchronize()
    sim.close()
