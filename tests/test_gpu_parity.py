"""GPU parity: the HIP engine against the CPU oracle on the same config, connectivity and noise.

Bar: bit-identical per-step neuron state (V, g_e, g_i, I_stim, refractory, spikes, background
conductances), connection and plasticity traces and weights, for the fused graph-captured step,
the component (drop-in launcher) step and the lazily normalised step.  The full-size 3,125-neuron
workload is checked for identical 1,000-step rasters and per-step V/g within 1e-5 relative.
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


def _run_lockstep(sim, orc, steps, check_every=1, component=False):
    for step in range(steps):
        if component:
            sim.step_components(1)
        else:
            sim.run(1)
        orc.step(1)
        if step % check_every == 0 or step == steps - 1:
            assert_state_equal(sim.state(), orc.state(), step)


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


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_fused_step_bit_exact(tmp_path, graph):
    # graph: whole captured graphs of 20 steps (step 0, a record step, runs alone first), the state
    # compared after every graph and the raster of every step from the device spike ring;
    # eager: no graph at all, the state compared after every step.
    path = _write(tmp_path, small_config(steps=600))
    flags = ("quiet",) if graph else ("no_monitors", "quiet", "no_graph")
    sim = engine().Simulation(path, flags=flags)
    orc = oracle_for(path, sim)
    if graph:
        sim.set_steps_per_graph(20)
        sim.run(1)
        orc.step(1)
        step = 1
        for _ in range(15):
            sim.run(20)
            raster = []
            for _ in range(20):
                orc.step(1)
                raster.append(orc.state()["spikes"])
            assert np.array_equal(sim.spike_raster(step, 20), np.array(raster)), f"raster of steps {step}..{step + 19}"
            step += 20
            assert_state_equal(sim.state(), orc.state(), step - 1)
        info = sim.get_info()
        assert info.steps_per_graph == 20 and info.graph_launches == 15  # every step after 0 ran in graphs
    else:
        _run_lockstep(sim, orc, 300)
    _check_traces_and_weights(sim, orc)
    sim.close()


@pytest.mark.parametrize("mode", ["lazy", "lazy-sequential", "faithful", "burst", "lazy-parts", "faithful-parts",
                                  "burst-parts"])
def test_graph_steps_bit_exact(tmp_path, mode):
    # Whole graphs of 10 steps: the overlapped form (plasticity of step t-1 in the launch of the
    # evolve of step t, staging copies, stimulus in the propagate launch) and the sequential
    # one must both leave exactly the oracle's state after every graph; faithful normalisation
    # always runs sequentially.
    parts = mode.endswith("-parts")
    mode = mode.replace("-parts", "")
    if mode == "burst":
        cfg = small_config(n_e_side=24, n_i_side=12, steps=400)
        cfg["stimulus_selected_input_ids"] = list(range(300))
        cfg["stimulus_selected_input_post_group_offsets"] = [0] * 300
        cfg["stimulus_start_clock_steps"] = [30, 95]
        cfg["stimulus_end_clock_steps"] = [40, 101]
        cfg["stimulus_input_values"] = [4e-9, 3e-9]
    else:
        cfg = small_config(steps=600)
    path = _write(tmp_path, cfg)
    lazy = mode != "faithful"
    flags = ("no_monitors", "quiet") + (("lazy_normalize",) if lazy else ()) + \
        (("sequential_graph",) if mode == "lazy-sequential" else ()) + (("gemv_part_sums",) if parts else ())
    sim = engine().Simulation(path, flags=flags)
    # Part sums: the propagate adds each part's chunk sums in the order the evolve would (same
    # results, one partial row per part).
    assert sim.get_info().gemv_part_sums == (1 if parts else 0)
    sim.set_steps_per_graph(10)
    orc = oracle_for(path, sim, lazy=lazy)
    for chunk in range(25):
        sim.run(10)
        orc.step(10)
        assert_state_equal(sim.state(), orc.state(), 10 * chunk + 9)
        if chunk % 8 == 7:
            _check_traces_and_weights(sim, orc)
    sim.run(3)  # a remainder of eager steps after the graphs, then graphs again
    orc.step(3)
    sim.run(20)
    orc.step(20)
    assert_state_equal(sim.state(), orc.state(), 272)
    _check_traces_and_weights(sim, orc)
    sim.close()


@pytest.mark.parametrize("spg", [100, 7, 2], ids=["default-100", "odd-7-sequential", "2"])
def test_steps_per_graph_bit_exact(tmp_path, spg):
    # The default graphs of 100 steps (overlapped form), an odd count (sequential form) and the
    # shortest overlapped graph, each with an eager remainder, against the oracle; then the
    # length changed mid-run.
    path = _write(tmp_path, small_config(steps=600))
    sim = engine().Simulation(path, flags=("no_monitors", "quiet", "lazy_normalize"))
    if spg != 100:
        sim.set_steps_per_graph(spg)
    orc = oracle_for(path, sim, lazy=True)
    sim.run(2 * spg + 3)
    orc.step(2 * spg + 3)
    assert_state_equal(sim.state(), orc.state(), 2 * spg + 2)
    sim.set_steps_per_graph(10)
    sim.run(40)
    orc.step(40)
    assert_state_equal(sim.state(), orc.state(), 2 * spg + 42)
    _check_traces_and_weights(sim, orc)
    with pytest.raises(engine().BnError):
        sim.set_steps_per_graph(0)
    sim.close()


@pytest.mark.parametrize("row_blocks,nontemporal,parts", [([8, 8, 16, 8], 1, False), ([3, 5, 7, 1], 0, False),
                                                          ([512, 512, 1, 2], 1, False), ([8, 8, 16, 8], 1, True),
                                                          ([6, 5, 7, 3], 0, True)],
                         ids=["rb8-nt", "rb-odd", "rb-mixed-nt", "rb8-nt-part-sums", "rb-odd-part-sums"])
def test_gemv_tilings_bit_exact(tmp_path, row_blocks, nontemporal, parts):
    # Row blocks change the number of chunk partials per post neuron (1 to 134 here, up to 17
    # per combination part) and so the order of the FP64 sums; the oracle follows the same
    # tiling.  Non-temporal weight loads must not change any value.
    cfg = small_config(steps=300)
    cfg["gemv_row_blocks"] = row_blocks
    cfg["gemv_nontemporal_loads"] = nontemporal
    path = _write(tmp_path, cfg)
    sim = engine().Simulation(path, flags=("no_monitors", "quiet", "lazy_normalize") + (("gemv_part_sums",) if parts else ()))
    assert tuple(sim.info.gemv_row_block) == tuple(row_blocks)
    assert sim.get_info().gemv_part_sums == (1 if parts else 0)
    orc = oracle_for(path, sim, lazy=True)
    _run_lockstep(sim, orc, 150, check_every=5)
    _check_traces_and_weights(sim, orc)
    sim.close()


@pytest.mark.parametrize("form", ["graph", "eager", "component"])
def test_pair_stdp_bit_exact(tmp_path, form):
    # E->E under the reference's pair-STDP component (e_e_plasticity_rule "pair") with iSTDP on
    # I->E: the overlapped graphs, the eager fused step and the component path
    # (bn_stdp_plasticity_update) against the oracle's pair rule, with a burst so that spiking
    # rows and columns meet.
    cfg = small_config(n_e_side=24, n_i_side=12, steps=600)
    cfg["e_e_plasticity_rule"] = "pair"
    cfg["e_e_plasticity_ltp_A"] = 3e-12
    cfg["e_e_plasticity_ltd_A"] = 2e-12
    cfg["stimulus_selected_input_ids"] = list(range(0, 576, 3))
    cfg["stimulus_selected_input_post_group_offsets"] = [0] * 192
    cfg["stimulus_start_clock_steps"] = [30]
    cfg["stimulus_end_clock_steps"] = [45]
    cfg["stimulus_input_values"] = [4e-9]
    path = _write(tmp_path, cfg)
    flags = {"graph": ("quiet",), "eager": ("no_monitors", "quiet", "no_graph"),
             "component": ("no_monitors", "quiet", "component_path")}[form]
    sim = engine().Simulation(path, flags=flags)
    assert sim.info.num_plasticities == 2
    orc = oracle_for(path, sim)
    w0 = sim.weights(0)
    if form == "graph":
        sim.set_steps_per_graph(20)
        sim.run(1)
        orc.step(1)
        for k in range(12):
            sim.run(20)
            orc.step(20)
            assert_state_equal(sim.state(), orc.state(), 20 * k + 20)
    else:
        _run_lockstep(sim, orc, 200, check_every=5, component=form == "component")
    _check_traces_and_weights(sim, orc)
    assert np.count_nonzero(sim.weights(0) != w0) > 1000, "the burst should move many E->E weights"
    sim.close()


def test_component_path_bit_exact(tmp_path):
    path = _write(tmp_path, small_config(steps=600))
    sim = engine().Simulation(path, flags=("no_monitors", "quiet", "component_path"))
    orc = oracle_for(path, sim)
    _run_lockstep(sim, orc, 250, component=True)
    _check_traces_and_weights(sim, orc)
    sim.close()


def test_lazy_normalization_bit_exact(tmp_path):
    path = _write(tmp_path, small_config(steps=600))
    sim = engine().Simulation(path, flags=("no_monitors", "quiet", "lazy_normalize"))
    orc = oracle_for(path, sim, lazy=True)
    _run_lockstep(sim, orc, 300)
    _check_traces_and_weights(sim, orc)
    sim.close()


def test_lazy_burst_bit_exact(tmp_path):
    # A synchronous burst: hundreds of spiking posts and pres in the same steps.
    cfg = small_config(n_e_side=24, n_i_side=12, steps=400)
    cfg["stimulus_selected_input_ids"] = list(range(300))
    cfg["stimulus_selected_input_post_group_offsets"] = [0] * 300
    cfg["stimulus_start_clock_steps"] = [30]
    cfg["stimulus_end_clock_steps"] = [40]
    cfg["stimulus_input_values"] = [4e-9]
    path = _write(tmp_path, cfg)
    sim = engine().Simulation(path, flags=("no_monitors", "quiet", "lazy_normalize"))
    orc = oracle_for(path, sim, lazy=True)
    _run_lockstep(sim, orc, 120, check_every=5)
    _check_traces_and_weights(sim, orc)
    assert orc.state()["spikes"].sum() >= 0
    sim.close()


@pytest.mark.parametrize("sequential", [False, True], ids=["overlapped", "sequential"])
def test_spike_list_overflow_bit_exact(tmp_path, sequential):
    # Every E neuron spikes in the same step (a 300 nA pulse moves V by 150 mV per step): 4,900
    # spiking pre neurons overflow the 4,096-entry list of spiking rows, and 4,900 spiking posts
    # overflow the list of the quiet I->E rows, so both fall back to scanning the spike bits.
    n_e = 70 * 70
    cfg = small_config(n_e_side=70, n_i_side=35, steps=200)
    cfg["stimulus_selected_input_ids"] = list(range(n_e))
    cfg["stimulus_selected_input_post_group_offsets"] = [0] * n_e
    cfg["stimulus_start_clock_steps"] = [12]
    cfg["stimulus_end_clock_steps"] = [14]
    cfg["stimulus_input_values"] = [300e-9]
    path = _write(tmp_path, cfg)
    flags = ("no_monitors", "quiet", "lazy_normalize") + (("sequential_graph",) if sequential else ())
    sim = engine().Simulation(path, flags=flags)
    sim.set_steps_per_graph(10)
    orc = oracle_for(path, sim, lazy=True)
    most = 0
    for chunk in range(5):  # whole graphs; the burst falls in the second
        sim.run(10)
        for _ in range(10):
            orc.step(1)
            most = max(most, int(orc.state()["spikes"][:n_e].sum()))
        assert_state_equal(sim.state(), orc.state(), 10 * chunk + 9)
    _check_traces_and_weights(sim, orc)
    assert most > 4096, most
    sim.close()


def _fp64_bars(g, o64, o64_spikes, ok, v_t):
    """Per-step comparison with the FP64 oracle: (V relative error over the neurons in the
    contracting regime, g_e and g_i errors relative to the step's largest |g|).  `ok` marks the
    neurons whose FP64 V stayed below V_T since their last spike (or the start); it is updated."""
    v, ref_v = g["membrane_potentials"].astype(np.float64), o64["V"]
    ok &= (ref_v < v_t) | (o64_spikes == 1)
    v_rel = float(np.max(np.abs(v - ref_v)[ok] / np.abs(ref_v[ok]))) if ok.any() else 0.0
    ok |= o64_spikes == 1  # a spike restarts the trajectory from the same capped / reset values
    g_rel = []
    for a, b in (("g_excitatory", "ge"), ("g_inhibitory", "gi")):
        ref = o64[b]
        scale = max(float(np.max(np.abs(ref))), 1e-30)
        g_rel.append(float(np.max(np.abs(g[a].astype(np.float64) - ref))) / scale)
    return v_rel, g_rel[0], g_rel[1]


@pytest.mark.parametrize("normalization", ["faithful", "lazy", "resident"])
def test_workload_3125_raster_1000_steps(tmp_path, normalization):
    # The 3,125-neuron workload stepped one step at a time for 1,000 steps: every step bit-exact
    # with the FP32 oracle, and against the FP64 oracle (an independent accuracy reference):
    # identical rasters, g_e and g_i within 1e-5 of the step's largest conductance, V within 1e-5
    # relative for every neuron whose FP64 V stayed below V_T since its last spike.  Above V_T the
    # EIF exponential multiplies a difference by about 1 + dt/C g_L exp((V - V_T)/Delta_T) per
    # step (300 at V = -30 mV), so an FP32 evaluation of the step drifts from FP64 there by up to
    # ~1e-2 relative before the spike caps V at the same v_threshold; those neuron-steps are
    # counted, not bounded.  "resident": the form the benchmark times (the resident step, lazy
    # normalisation, in its own summation orders), here one launch per step.
    cfg = load_config("net3125_stdp.json")
    path = _write(tmp_path, cfg)
    lazy = normalization in ("lazy", "resident")
    flags = ("no_monitors", "quiet") + (("lazy_normalize",) if lazy else ()) + (
        ("resident",) if normalization == "resident" else ())
    sim = engine().Simulation(path, flags=flags)
    assert (sim.get_info().resident_blocks > 0) == (normalization == "resident")
    orc = oracle_for(path, sim, lazy=lazy)
    o64 = oracle_for(path, sim, precision=64)
    n_spikes = 0
    ok = np.ones(sim.N, bool)
    worst = [0.0, 0.0, 0.0]
    checked = 0
    for step in range(1000):
        sim.run(1)
        orc.step(1)
        o64.step(1)
        g, o = sim.state(), orc.state()
        assert_state_equal(g, o, step, fields=("membrane_potentials", "g_excitatory", "g_inhibitory", "spikes",
                                               "refractory_counter"))
        sp64 = o64.state()["spikes"]
        assert np.array_equal(g["spikes"], sp64), f"raster differs from FP64 at step {step}"
        bars = _fp64_bars(g, o64.state64(), sp64, ok, cfg["v_effective_spike_threshold"])
        worst = [max(a, b) for a, b in zip(worst, bars)]
        checked += int(ok.sum())
        n_spikes += int(g["spikes"].sum())
    print(f"FP64 bars over 1,000 steps: V {worst[0]:.3g} (on {checked / (1000 * sim.N):.2f} of neuron-steps), "
          f"g_e {worst[1]:.3g}, g_i {worst[2]:.3g}; {n_spikes} spikes")
    assert worst[0] <= 1e-5 and worst[1] <= 1e-5 and worst[2] <= 1e-5, worst
    assert checked >= 0.5 * 1000 * sim.N
    assert n_spikes > 0
    w_gpu, w_orc = sim.weights(0), orc.weights(0)
    assert np.array_equal(w_gpu, w_orc)
    sim.close()


def _graph_lockstep(sim, orc, first_step, graphs, spg):
    """`graphs` whole graphs of `spg` steps from `first_step`: the raster of every step from the
    device spike ring and the state after every graph, bit for bit against the oracle."""
    step = first_step
    spikes = 0
    for _ in range(graphs):
        sim.run(spg)
        raster = []
        for _ in range(spg):
            orc.step(1)
            raster.append(orc.state()["spikes"])
        raster = np.array(raster)
        got = sim.spike_raster(step, spg)
        if not np.array_equal(got, raster):
            bad = np.nonzero((got != raster).any(axis=1))[0]
            raise AssertionError(f"raster differs at steps {step + bad[:5]}")
        spikes += int(raster.sum())
        step += spg
        assert_state_equal(sim.state(), orc.state(), step - 1)
    return spikes


@pytest.mark.parametrize("spg,parts", [(500, False), (100, False), (500, True)],
                         ids=["500", "100", "500-part-sums"])
def test_bench_form_3125_graphs(tmp_path, spg, parts):
    # The form the benchmark times: lazy normalisation, overlapped graphs of 500 (and the library
    # default 100) steps with the monitors' spike ring, 1,000 steps bit-exact against the FP32
    # lazy oracle (raster of every step, state after every graph, traces and weights at the end).
    cfg = load_config("net3125_stdp.json")
    path = _write(tmp_path, cfg)
    sim = engine().Simulation(path, flags=("quiet", "lazy_normalize") + (("gemv_part_sums",) if parts else ()))
    assert sim.get_info().gemv_part_sums == (1 if parts else 0)
    sim.set_steps_per_graph(spg)
    assert sim.get_info().kernels_per_step == 2  # the overlapped form
    orc = oracle_for(path, sim, lazy=True)
    sim.run(1)  # step 0 is a record step and runs alone
    orc.step(1)
    spikes = _graph_lockstep(sim, orc, 1, 1000 // spg, spg)
    assert sim.get_info().graph_launches == 1000 // spg
    _check_traces_and_weights(sim, orc)
    assert spikes > 0
    sim.close()


@pytest.mark.parametrize("parts", [False, True], ids=["chunk-partials", "part-sums"])
def test_net10125_graphs(tmp_path, parts):
    # The 10,125-neuron workload (410 MB of weights, non-temporal weight loads, long row chunks):
    # lazy normalisation, overlapped graphs of 250 steps, 1,001 steps against the FP32 lazy oracle.
    cfg = load_config("net10125_stdp.json")
    path = _write(tmp_path, cfg)
    sim = engine().Simulation(path, flags=("quiet", "lazy_normalize") + (("gemv_part_sums",) if parts else ()))
    assert sim.get_info().gemv_part_sums == (1 if parts else 0)
    sim.set_steps_per_graph(250)
    assert sim.get_info().kernels_per_step == 2
    orc = oracle_for(path, sim, lazy=True)
    sim.run(1)
    orc.step(1)
    spikes = _graph_lockstep(sim, orc, 1, 4, 250)
    assert sim.get_info().graph_launches == 4
    print(f"net10125: {spikes} spikes in steps 1..1000; row blocks {tuple(sim.info.gemv_row_block)}")
    assert spikes > 0
    for k in range(4):
        for a, b, name in zip(sim.connection_traces(k), orc.connection_traces(k), ("r", "d", "x")):
            assert np.array_equal(a, b), f"connection {k} trace {name}"
    for off in range(sim.info.num_plasticities):
        for a, b in zip(sim.plasticity_traces(off), orc.plasticity_traces(off)):
            assert np.array_equal(a, b)
    for k in (0, 2):  # the plastic matrices
        assert np.array_equal(sim.weights(k), orc.weights(k)), f"weights {k}"
    sim.close()


def test_read_bandwidth_probe():
    # The streaming-read probe the bench reports beside the propagate: one-pass and multi-pass
    # grids, the non-temporal form above 192 MB, and the argument checks.
    eng = engine()
    for nbytes in (1 << 20, 39_062_500 // 16 * 16, 256 << 20):
        gbs = eng.measure_read_bandwidth(nbytes, 5)
        assert 10.0 < gbs < 20000.0, (nbytes, gbs)
    for nbytes in (4096, 39_062_500 // 16 * 16, 410_062_500 // 16 * 16):
        gbs = eng.measure_read_bandwidth_one_launch(nbytes, 20)
        assert 1.0 < gbs < 40000.0, (nbytes, gbs)
    with pytest.raises(eng.BnError):
        eng.measure_read_bandwidth(8, 1)
    with pytest.raises(eng.BnError):
        eng.measure_read_bandwidth(1 << 20, 0)
    with pytest.raises(eng.BnError):
        eng.measure_read_bandwidth_one_launch(1 << 20, 0)
