"""Pins of the CPU oracle (CPU only).

Each check is chosen so that a plausible slip in the oracle (a dropped term, a wrong index or
sign, a transposed operand, a wrong summation order) fails it: published known-answer vectors,
an independent implementation (rocRAND), closed forms, scipy, and hand-computed cases.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

import oracle_py as O
from helpers import TESTS, small_config

GOLDEN = os.path.join(TESTS, "golden")


# --------------------------------------------------------------------------------------
# Philox4x32-10
# --------------------------------------------------------------------------------------
def test_philox_published_known_answers():
    # Known-answer vectors of the method's authors (Random123 distribution, philox4x32 10 rounds).
    kats = [([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
            ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
            ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
             [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1])]
    for ctr, key, expected in kats:
        assert list(O.philox(ctr, key)) == expected


def test_philox_matches_rocrand_engine():
    vectors = json.load(open(os.path.join(GOLDEN, "philox_rocrand.json")))
    assert len(vectors) == 300
    for v in vectors:
        s = v["seed"]
        got = O.philox([v["step"], 0, v["neuron"], v["instance"]], [s & 0xFFFFFFFF, s >> 32])
        assert list(got) == v["out"], v


# --------------------------------------------------------------------------------------
# expf
# --------------------------------------------------------------------------------------
def _ulp_err(got: np.ndarray, x: np.ndarray) -> np.ndarray:
    exact = np.exp(x.astype(np.float64))
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / ulp


def test_expf_accuracy():
    rng = np.random.default_rng(1)
    # EIF arguments (V - V_T)/Delta_T lie in about [-20, 20]; also sweep the whole finite range.
    xs = np.concatenate([np.linspace(-20, 20, 20001), rng.uniform(-87, 88, 20000),
                         np.array([0.0, -0.0, 1e-30, -1e-30, 0.5, -0.5, 88.7, -87.3, -103.9, -100.0])]).astype(np.float32)
    got = np.array([O.expf(float(x)) for x in xs], np.float32)
    normal = np.abs(np.exp(xs.astype(np.float64))) > 1.2e-38
    assert np.max(_ulp_err(got[normal], xs[normal])) <= 2.0
    sub = ~normal & (np.exp(xs.astype(np.float64)) > 0)
    assert np.all(np.abs(got[sub].astype(np.float64) - np.exp(xs[sub].astype(np.float64))) <= 2 * 1.4e-45)
    assert O.expf(float("inf")) == float("inf") and O.expf(89.0) == float("inf")
    assert O.expf(-104.0) == 0.0 and math.isnan(O.expf(float("nan")))


# --------------------------------------------------------------------------------------
# Poisson inverse-CDF table and sampling
# --------------------------------------------------------------------------------------
@pytest.mark.parametrize("lam", [0.02, 0.1, 1.5, 7.0, 40.0])
def test_poisson_table_matches_scipy(lam):
    t = O.poisson_table(lam).astype(np.float64)
    k = np.arange(t.size)
    expected = np.floor(np.minimum(stats.poisson.cdf(k, np.float32(lam)) * 2.0 ** 32, 2.0 ** 32 - 1))
    assert np.all(np.abs(t - expected) <= 2)
    assert t[-1] == 2.0 ** 32 - 1 and np.all(np.diff(t) >= 0)
    assert O.poisson_table(0.0).size == 0 and O.poisson_table(-1.0).size == 0


def _oracle(cfg, **kw):
    p = O.params_from_config(cfg, gemv_row_block=(32, 32, 32, 32), **kw)
    ne, ni = cfg["num_excitatory_neurons"], cfg["num_inhibitory_neurons"]
    shapes = [(ne, ne), (ne, ni), (ni, ne), (ni, ni)]
    return p, shapes


def _cfg_flat(cfg):
    """The engine's parsed-config field names from a raw config (no weight-generation keys)."""
    return cfg


def _draws(seed, step, n, instance=0):
    s = seed & 0xFFFFFFFFFFFFFFFF
    return O.philox([step, 0, n, instance], [s & 0xFFFFFFFF, s >> 32])


def test_background_counts_statistics():
    # Counter (step, 0, neuron, 0), word x -> excitatory count by table inversion.
    tab = O.poisson_table(0.1)
    counts = np.array([int(np.searchsorted(tab, _draws(5, t, n)[0], side="right"))
                       for t in range(40) for n in range(2000)])
    assert abs(counts.mean() - 0.1) < 0.005
    assert abs(counts.var() - 0.1) < 0.006
    assert abs((counts == 0).mean() - math.exp(-0.1)) < 0.005
    assert abs((counts == 1).mean() - 0.1 * math.exp(-0.1)) < 0.005


def test_background_conductance_uses_the_counter_layout():
    # Background recurrence of kernel_neuron_group_evolve.cu:11-83 in FP32, fed with counts drawn
    # from the counter layout above: pins which counter/output word and table the oracle uses.
    cfg = small_config(n_e_side=10, n_i_side=5, stimulus=False, plastic=False)
    cfg["group_e_mean_background_inhibitory_spikes"] = 2.5
    p, shapes = _oracle(cfg)
    sim = O.OracleSim(p, [np.zeros(s, np.float32) for s in shapes])
    f32 = np.float32
    g = {k: f32(cfg["group_e_" + k]) for k in (
        "background_excitatory_postsynaptic_r_trace_decay", "background_excitatory_postsynaptic_d_trace_decay",
        "background_inhibitory_postsynaptic_r_trace_decay", "background_inhibitory_postsynaptic_d_trace_decay",
        "background_excitatory_scaling_factor", "background_inhibitory_scaling_factor")}
    tab_e, tab_i = O.poisson_table(0.1), O.poisson_table(2.5)
    ne = 100
    re, de, ri, di = (np.zeros(ne, np.float32) for _ in range(4))
    flush = lambda a: np.where(a < f32(1e-30), f32(0), a).astype(np.float32)  # noqa: E731
    for step in range(30):
        sim.step(1)
        st = sim.state()
        re = flush(re * g["background_excitatory_postsynaptic_r_trace_decay"])
        ri = flush(ri * g["background_inhibitory_postsynaptic_r_trace_decay"])
        de = flush(de * g["background_excitatory_postsynaptic_d_trace_decay"])
        di = flush(di * g["background_inhibitory_postsynaptic_d_trace_decay"])
        d = [_draws(cfg["rng_seed"], step, n) for n in range(ne)]
        ke = np.array([np.searchsorted(tab_e, x[0], side="right") for x in d], np.float32)
        ki = np.array([np.searchsorted(tab_i, x[1], side="right") for x in d], np.float32)
        re, de, ri, di = re + ke, de + ke, ri + ki, di + ki
        assert np.array_equal(st["bg_ge"][:ne], (de - re) * g["background_excitatory_scaling_factor"])
        assert np.array_equal(st["bg_gi"][:ne], (di - ri) * g["background_inhibitory_scaling_factor"])


# --------------------------------------------------------------------------------------
# Closed forms: connection traces, EIF step, GEMV
# --------------------------------------------------------------------------------------
def _quiet_config(ne_side=4, ni_side=2, steps=200):
    cfg = small_config(n_e_side=ne_side, n_i_side=ni_side, steps=steps, stimulus=False, plastic=False)
    for g in ("e", "i"):
        cfg[f"group_{g}_mean_background_excitatory_spikes"] = 0.0
        cfg[f"group_{g}_mean_background_inhibitory_spikes"] = 0.0
    return cfg


def _stim(cfg, schedule):
    """schedule: list of (neuron global id, step) -> one-step 1 uA pulse applied at `step`."""
    T = cfg["simulation_steps"]
    counts = np.zeros(T, np.uint32)
    ev = sorted((t, nid) for nid, t in schedule)
    ids, offs, amps = [], [], []
    per = {}
    for t, nid in ev:
        per.setdefault(t, []).append((nid, 1e-6))
        per.setdefault(t + 1, []).append((nid, -1e-6))
    for t in sorted(per):
        for nid, a in sorted(per[t]):
            counts[t] += 1
            ids.append(nid)
            offs.append(0)
            amps.append(a)
    return counts, np.array(ids, np.uint32), np.array(offs, np.uint32), np.array(amps, np.float32)


def test_connection_traces_closed_form():
    cfg = _quiet_config()
    p, shapes = _oracle(cfg)
    sim = O.OracleSim(p, [np.zeros(s, np.float32) for s in shapes], stim=_stim(cfg, [(0, 0)]))
    K = cfg["postsynaptic_kernel_normalization_value"]
    dr = float(np.exp(np.float32(-1.0) * np.float32(cfg["timestep"]) / np.float32(cfg["tau_rise_e"])))
    dd = float(np.exp(np.float32(-1.0) * np.float32(cfg["timestep"]) / np.float32(cfg["tau_decay_e"])))
    spike_step = None
    for step in range(150):
        sim.step(1)
        if spike_step is None and sim.state()["spikes"][0] == 1:
            spike_step = step
        r, d, x = sim.connection_traces(0)
        if spike_step is not None:
            k = step - spike_step
            exp_r, exp_d = K * dr ** k, K * dd ** k
            assert abs(r[0] - exp_r) <= 1e-5 * exp_r + 1e-30
            assert abs(d[0] - exp_d) <= 1e-5 * exp_d + 1e-30
            assert abs(x[0] - (exp_d - exp_r)) <= 1e-5 * exp_d
    assert spike_step == 1  # stimulus applied at step 0 is seen by evolve at step 1


def test_eif_steps_match_fp64():
    cfg = _quiet_config()
    cfg["group_e_mean_background_excitatory_spikes"] = 0.3
    p, shapes = _oracle(cfg)
    sim = O.OracleSim(p, [np.zeros(s, np.float32) for s in shapes])
    gl, C, dT, VT = cfg["g_leak"], cfg["membrane_capacitance"], cfg["depolarization_slope_factor"], cfg["v_effective_spike_threshold"]
    Ee, Ei, vr, dt = cfg["v_excitatory_reversal"], cfg["v_inhibitory_reversal"], cfg["v_rest"], cfg["timestep"]
    prev = sim.state()
    for step in range(40):
        sim.step(1)
        st = sim.state()
        v = prev["V"].astype(np.float64)
        ge, gi = prev["ge"].astype(np.float64), prev["gi"].astype(np.float64)
        dv = dt / C * (-gl * (v - vr) + gl * dT * np.exp((v - VT) / dT) - ge * (v - Ee) - gi * (v - Ei))
        expected = v + dv
        free = prev["refr"] == 0
        assert np.all(st["spikes"] == 0)
        assert np.allclose(st["V"][free], expected[free], rtol=2e-6, atol=0)
        prev = st


def test_gemv_matches_fp64():
    cfg = _quiet_config(ne_side=10, ni_side=5)
    rng = np.random.default_rng(4)
    ne, ni = 100, 25
    shapes = [(ne, ne), (ne, ni), (ni, ne), (ni, ni)]
    ws = [(rng.uniform(0, 1e-9, s) * (rng.uniform(size=s) < 0.3)).astype(np.float32) for s in shapes]
    # Chunked orders of two row blocks, and the resident step's orders (columns per block, waves
    # of excitatory rows): every term must enter each sum exactly once.
    for rb, res in (((7, 7, 7, 7), (0, 0)), ((100, 100, 25, 25), (0, 0)), ((0, 0, 0, 0), (16, 6)),
                    ((0, 0, 0, 0), (4, 1))):
        p = O.params_from_config(cfg, gemv_row_block=rb, resident=res)
        sched = [(n, 0) for n in (0, 3, 17, 50, 99)] + [(ne + 2, 0), (ne + 20, 0)]
        sim = O.OracleSim(p, ws, stim=_stim(cfg, sched))
        sim.step(6)
        st = sim.state()
        x_e = sim.connection_traces(0)[2].astype(np.float64)
        x_i = sim.connection_traces(2)[2].astype(np.float64)
        ge_e = st["bg_ge"][:ne] + ws[0].astype(np.float64).T @ x_e
        gi_e = st["bg_gi"][:ne] + ws[2].astype(np.float64).T @ x_i
        ge_i = st["bg_ge"][ne:] + ws[1].astype(np.float64).T @ x_e
        assert np.abs(x_e).sum() > 0 and np.abs(x_i).sum() > 0
        for got, exp in ((st["ge"][:ne], ge_e), (st["gi"][:ne], gi_e), (st["ge"][ne:], ge_i)):
            assert np.allclose(got, exp, rtol=1e-6, atol=1e-30)


# --------------------------------------------------------------------------------------
# Plasticity: isolated pairs (hand-computed), index self-skip, normalisation targets
# --------------------------------------------------------------------------------------
def test_istdp_isolated_pair_hand_computed():
    cfg = _quiet_config()
    cfg["i_e_plasticity_enabled"] = True
    cfg["i_e_plasticity_ltp_A"] = 3e-12
    cfg["i_e_plasticity_ltd_A"] = 2e-12
    cfg["i_e_plasticity_ltd_B"] = 0.25
    cfg["i_e_plasticity_weight_max"] = 1e-8
    ne, ni = 16, 4
    p = O.params_from_config(cfg, gemv_row_block=(8, 8, 8, 8))
    w_ie = np.full((ni, ne), 1e-9, np.float32)
    ws = [np.zeros((ne, ne), np.float32), np.zeros((ne, ni), np.float32), w_ie, np.zeros((ni, ni), np.float32)]
    # I neuron 1 (global 17) spikes at step 1, E neuron 5 at step 9.
    sim = O.OracleSim(p, ws, stim=_stim(cfg, [(ne + 1, 0), (5, 8)]))
    dpre = float(np.exp(np.float32(-1.0) * np.float32(cfg["timestep"]) / np.float32(cfg["i_e_plasticity_tau_ltp"])))
    sim.step(2)
    w = sim.weights(2)
    # LTD of pre row 1 with all post traces 0: w += -ltd_A * (0 - ltd_B), except j == 1 (index skip)
    exp = np.full(ne, 1e-9)
    exp[np.arange(ne) != 1] += 2e-12 * 0.25
    assert np.allclose(w[1], exp, rtol=1e-6, atol=0) and w[1, 1] == np.float32(1e-9)
    assert np.all(w[[0, 2, 3]] == np.float32(1e-9))
    sim.step(8)
    w2 = sim.weights(2)
    # LTP of post column 5 at step 9: w[i,5] += ltp_A * pre_i, pre_1 = dpre^(9-1)
    assert np.isclose(w2[1, 5] - w[1, 5], 3e-12 * dpre ** 8, rtol=1e-4, atol=0)
    assert np.all(w2[[0, 2, 3], 5] == np.float32(1e-9))


def test_triplet_ltp_delta_from_weight_ratios():
    cfg = _quiet_config()
    cfg["e_e_plasticity_enabled"] = True
    cfg["e_e_plasticity_ltp_A"] = 1e-11
    cfg["e_e_plasticity_ltp_B"] = 0.0
    cfg["e_e_plasticity_ltd_A"] = 0.0
    cfg["e_e_plasticity_ltd_B"] = 0.0
    ne, ni = 16, 4
    rng = np.random.default_rng(2)
    w_ee = rng.uniform(0.5e-9, 1e-9, (ne, ne)).astype(np.float32)
    np.fill_diagonal(w_ee, 0)
    ws = [w_ee, np.zeros((ne, ni), np.float32), np.zeros((ni, ne), np.float32), np.zeros((ni, ni), np.float32)]
    p = O.params_from_config(cfg, gemv_row_block=(8, 8, 8, 8))
    sim = O.OracleSim(p, ws, stim=_stim(cfg, [(3, 0), (9, 6)]))  # pre 3 at step 1, post 9 at step 7
    dpre = float(np.exp(np.float32(-1.0) * np.float32(cfg["timestep"]) / np.float32(cfg["e_e_plasticity_tau_ltp"])))
    sim.step(7)
    before = sim.weights(0).astype(np.float64)
    sim.step(1)
    after = sim.weights(0).astype(np.float64)
    # normalisation rescales whole rows, so the ratio to an untouched column isolates the LTP delta
    delta = (after[3, 9] / after[3, 1] - before[3, 9] / before[3, 1]) * before[3, 1]
    assert np.isclose(delta, 1e-11 * dpre ** 6, rtol=1e-3, atol=0)
    others = [i for i in range(ne) if i not in (1, 3, 9)]
    r_after = after[others, 9] / after[others, 1]
    r_before = before[others, 9] / before[others, 1]
    assert np.allclose(r_after, r_before, rtol=1e-5, atol=0)  # pre traces 0: no change


def _decay(cfg, tau_key):
    return float(np.exp(np.float32(-1.0) * np.float32(cfg["timestep"]) / np.float32(cfg[tau_key])))


def _row_delta(before, after, i, j, j_ref):
    """Change of entry (i, j) in the units of `before`, with row i's rescaling divided out through
    an untouched column j_ref of the same row (the row normalisation scales whole rows)."""
    return after[i, j] / after[i, j_ref] * before[i, j_ref] - before[i, j]


@pytest.mark.parametrize("lazy,tile,res", [(False, False, 0), (True, False, 0), (True, True, 0), (True, False, 4)],
                         ids=["faithful", "lazy", "lazy-tile-order", "lazy-resident-order"])
def test_triplet_ltd_and_triplet_terms_hand_computed(lazy, tile, res):
    # kernel_triplet_plasticity_update.cu:54-93 with every coefficient nonzero.  E neuron 3 spikes
    # at steps 1 and 41, E neuron 9 at step 11 (one-step 1 uA pulses one step earlier).  At
    # step 41 (traces decayed first, increments after, :16-50 and :95-115):
    #   LTD of pre row 3:  w[3][9] += -post_9 * (ltd_A + ltd_B * tpre_3)   (post_9 = y_fast of 9)
    #   LTP of post col 3: w[9][3] +=  pre_9 * (ltp_A + ltp_B * tpost_3)   (tpost_3 = y_slow of 3)
    # with post_9 = dpost^30, tpre_3 = dtpre^40, pre_9 = dpre^30, tpost_3 = dtpost^40.  The four
    # trace time constants differ, so swapping pre/post or fast/slow traces, dropping a B term or
    # using the wrong neuron's trace moves a delta by 5-60%.
    cfg = _quiet_config()
    cfg["e_e_plasticity_enabled"] = True
    cfg.update({"e_e_plasticity_ltp_A": 1e-12, "e_e_plasticity_ltp_B": 3e-12, "e_e_plasticity_ltd_A": 1e-12,
                "e_e_plasticity_ltd_B": 2e-12, "e_e_plasticity_weight_max": 1e-8})
    ne, ni = 16, 4
    rng = np.random.default_rng(5)
    w_ee = rng.uniform(0.5e-9, 1e-9, (ne, ne)).astype(np.float32)
    np.fill_diagonal(w_ee, 0)
    ws = [w_ee, np.zeros((ne, ni), np.float32), np.zeros((ni, ne), np.float32), np.zeros((ni, ni), np.float32)]
    p = O.params_from_config(cfg, gemv_row_block=(8, 8, 8, 8), lazy_normalize=lazy, lazy_tile_order=tile,
                             gemv_part_chunks=(8, 8, 8, 8) if tile else (0, 0, 0, 0), resident=(res, 1 if res else 0))
    sim = O.OracleSim(p, ws, stim=_stim(cfg, [(3, 0), (9, 10), (3, 40)]))
    sim.step(41)
    assert [int(np.nonzero(sim.state()["spikes"])[0].size)] == [0]
    before = sim.weights(0).astype(np.float64)
    sim.step(1)
    assert list(np.nonzero(sim.state()["spikes"])[0]) == [3]
    after = sim.weights(0).astype(np.float64)
    dpre, dpost = _decay(cfg, "e_e_plasticity_tau_ltp"), _decay(cfg, "e_e_plasticity_tau_ltd")
    dtpre, dtpost = _decay(cfg, "e_e_plasticity_triplet_tau_ltp"), _decay(cfg, "e_e_plasticity_triplet_tau_ltd")
    ltd = -(dpost ** 30) * (1e-12 + 2e-12 * dtpre ** 40)
    ltp = dpre ** 30 * (1e-12 + 3e-12 * dtpost ** 40)
    assert np.isclose(_row_delta(before, after, 3, 9, 1), ltd, rtol=2e-4, atol=0)
    assert np.isclose(_row_delta(before, after, 9, 3, 1), ltp, rtol=2e-4, atol=0)
    # Every other entry of rows 3 and 9 keeps its ratio to an untouched column (post traces of neurons that
    # never spiked are 0; the diagonal is skipped by index), as do all other rows.
    for i in range(ne):
        ref = 1 if i != 1 else 2
        for j in range(ne):
            if i == j or j == ref or (i, j) in ((3, 9), (9, 3)):
                continue
            assert abs(_row_delta(before, after, i, j, ref)) <= 1e-6 * before[i, j] + 1e-30, (i, j)
    # The trace increments come after the updates: the triplet traces of neuron 3 now hold the
    # decayed old value plus 1.
    pre, post, tpre, tpost = sim.plasticity_traces(0)
    assert np.isclose(tpre[3], dtpre ** 40 + 1.0, rtol=1e-5, atol=0) and np.isclose(tpost[3], dtpost ** 40 + 1.0, rtol=1e-5, atol=0)
    assert np.isclose(pre[9], dpre ** 30, rtol=1e-5, atol=0) and np.isclose(post[9], dpost ** 30, rtol=1e-5, atol=0)


def test_pair_stdp_hand_computed():
    # The reference's pair-STDP component (kernel_stdp_plasticity_update.cu:16-137) on E->E:
    # E neuron 3 spikes at steps 1 and 41, E neuron 9 at step 11.  Pre and post traces have
    # different time constants and ltp_A != ltd_A, so swapping the traces, the rates or the
    # roles of row and column moves a delta by 10-80%.
    #   step 11: LTP of column 9: w[3][9] += ltp_A * pre_3,  pre_3  = dpre^10
    #            LTD of row 9:    w[9][3] += -ltd_A * post_3, post_3 = dpost^10
    #   step 41: LTD of row 3:    w[3][9] += -ltd_A * post_9, post_9 = dpost^30
    #            LTP of column 3: w[9][3] += ltp_A * pre_9,  pre_9  = dpre^30
    # w[3][9] starts just below weight_max (the step-11 LTP clamps to it) and w[9][3] just above
    # weight_min = 0 (the step-11 LTD clamps to it); every other entry stays bit-identical (no
    # triplet terms, no normalisation, the diagonal skipped by index).
    cfg = _quiet_config()
    cfg["e_e_plasticity_enabled"] = True
    cfg.update({"e_e_plasticity_rule": "pair", "e_e_plasticity_ltp_A": 1e-12, "e_e_plasticity_ltd_A": 3e-12,
                "e_e_plasticity_ltp_B": 5e-12, "e_e_plasticity_ltd_B": 7e-12,  # unused by the pair rule
                "e_e_plasticity_tau_ltp": 0.02, "e_e_plasticity_tau_ltd": 0.005,
                "e_e_plasticity_weight_min": 0.0, "e_e_plasticity_weight_max": 1.0005e-9})
    ne, ni = 16, 4
    rng = np.random.default_rng(8)
    w_ee = rng.uniform(0.5e-9, 1e-9, (ne, ne)).astype(np.float32)
    np.fill_diagonal(w_ee, 0)
    w_ee[3, 9] = 1e-9
    w_ee[9, 3] = 1e-13
    ws = [w_ee, np.zeros((ne, ni), np.float32), np.zeros((ni, ne), np.float32), np.zeros((ni, ni), np.float32)]
    p = O.params_from_config(cfg, gemv_row_block=(8, 8, 8, 8))
    assert p.e_e_pair_stdp == 1
    sim = O.OracleSim(p, ws, stim=_stim(cfg, [(3, 0), (9, 10), (3, 40)]))
    dpre, dpost = _decay(cfg, "e_e_plasticity_tau_ltp"), _decay(cfg, "e_e_plasticity_tau_ltd")
    wmax = float(np.float32(1.0005e-9))
    sim.step(11)
    b1 = sim.weights(0)
    sim.step(1)
    assert list(np.nonzero(sim.state()["spikes"])[0]) == [9]
    a1 = sim.weights(0)
    assert a1[3, 9] == np.float32(wmax)  # 1e-9 + ltp_A * dpre^10 clamped to weight_max
    assert a1[9, 3] == 0.0  # 1e-13 - ltd_A * dpost^10 clamped to weight_min
    assert 1e-9 + 1e-12 * dpre ** 10 > wmax and 1e-13 - 3e-12 * dpost ** 10 < 0
    mask = np.ones_like(a1, bool)
    mask[3, 9] = mask[9, 3] = False
    assert np.array_equal(a1[mask], b1[mask])
    sim.step(29)
    b2 = sim.weights(0)
    sim.step(1)
    assert list(np.nonzero(sim.state()["spikes"])[0]) == [3]
    a2 = sim.weights(0).astype(np.float64)
    assert np.isclose(a2[3, 9] - wmax, -3e-12 * dpost ** 30, rtol=2e-4, atol=0)
    assert np.isclose(a2[9, 3], 1e-12 * dpre ** 30, rtol=1e-6, atol=0)
    assert np.array_equal(a2[mask], b2[mask].astype(np.float64))
    # Trace increments after the updates; no triplet traces.
    pre, post, tpre, tpost = sim.plasticity_traces(0)
    assert np.isclose(pre[3], dpre ** 40 + 1.0, rtol=1e-5, atol=0) and np.isclose(post[3], dpost ** 40 + 1.0, rtol=1e-5, atol=0)
    assert np.isclose(pre[9], dpre ** 30, rtol=1e-5, atol=0) and np.isclose(post[9], dpost ** 30, rtol=1e-5, atol=0)
    assert not np.any(tpre) and not np.any(tpost)


def test_eif_spike_cap_and_refractory_sequence():
    # kernel_neuron_group_evolve.cu:92-118.  A one-step 1 uA pulse seen at step 1 lifts V by about
    # 0.5 V: the step must cap V at v_threshold (not V + dv), set refractory = R and spike = 1;
    # for the next R steps V is held at v_reset with the counter running down R-1 .. 0 and no
    # spike; the step after that integrates freely from v_reset (FP64 hand computation).
    cfg = _quiet_config()
    p, shapes = _oracle(cfg)
    sim = O.OracleSim(p, [np.zeros(s, np.float32) for s in shapes], stim=_stim(cfg, [(0, 0)]))
    R = cfg["refractory_steps"]
    f32 = np.float32
    sim.step(1)
    st = sim.state()
    assert st["spikes"].sum() == 0 and st["V"][0] == f32(cfg["v_rest"])
    sim.step(1)
    st = sim.state()
    assert st["spikes"][0] == 1 and st["spikes"][1:].sum() == 0
    assert st["V"][0] == f32(cfg["v_threshold"]) and st["refr"][0] == R
    for k in range(1, R + 1):
        sim.step(1)
        st = sim.state()
        assert st["spikes"][0] == 0, k
        assert st["V"][0] == f32(cfg["v_reset"]), k
        assert st["refr"][0] == R - k, k
    sim.step(1)
    st = sim.state()
    gl, C, dT = cfg["g_leak"], cfg["membrane_capacitance"], cfg["depolarization_slope_factor"]
    VT, vr, dt = cfg["v_effective_spike_threshold"], cfg["v_rest"], cfg["timestep"]
    v = float(f32(cfg["v_reset"]))
    expected = v + dt / C * (-gl * (v - vr) + gl * dT * math.exp((v - VT) / dT))
    assert st["spikes"][0] == 0 and st["refr"][0] == 0
    assert abs(st["V"][0] - expected) <= 2e-6 * abs(expected)
    assert st["V"][0] != f32(cfg["v_reset"])
    # Without the pulse nothing spikes and V stays exactly at rest.
    assert np.all(st["V"][1:] == f32(cfg["v_rest"]))


def test_triplet_normalisation_restores_column_sum_targets():
    cfg = _quiet_config()
    cfg["e_e_plasticity_enabled"] = True
    ne, ni = 16, 4
    rng = np.random.default_rng(3)
    w_ee = rng.uniform(0, 1e-9, (ne, ne)).astype(np.float32) * (rng.uniform(size=(ne, ne)) < 0.5)
    w_ee = w_ee.astype(np.float32)
    ws = [w_ee, np.zeros((ne, ni), np.float32), np.zeros((ni, ne), np.float32), np.zeros((ni, ni), np.float32)]
    target = np.array([np.float32(np.sum(w_ee[:, j], dtype=np.float32)) for j in range(ne)])  # column sums
    for lazy, tile, res in ((False, False, 0), (True, False, 0), (True, True, 0), (True, False, 4)):
        p = O.params_from_config(cfg, gemv_row_block=(8, 8, 8, 8), lazy_normalize=lazy, lazy_tile_order=tile,
                                 resident=(res, 3 if res else 0))
        sim = O.OracleSim(p, ws)
        sim.step(3)
        w = sim.weights(0).astype(np.float64)
        ok = w_ee.sum(axis=1) > 0
        assert np.allclose(w.sum(axis=1)[ok], target[ok], rtol=1e-6, atol=0)


@pytest.mark.parametrize("tile,res", [(False, 0), (True, 0), (False, 16)], ids=["list-order", "tile-order", "resident-order"])
def test_lazy_row_sums_restore_targets_with_spikes(tile, res):
    # Lazy normalisation keeps R_i = sum_j M_ij incrementally (spiking rows re-summed, other rows
    # plus their LTP changes); the true weights s_i * M_ij of every row must then sum to the
    # row's target (the initial column sum, triplet_plasticity.cpp:22-30) after every step,
    # through a stimulus-driven burst with hundreds of LTP/LTD updates per step.
    cfg = small_config(n_e_side=20, n_i_side=10, steps=300)
    cfg["stimulus_selected_input_ids"] = list(range(0, 400, 3))
    cfg["stimulus_selected_input_post_group_offsets"] = [0] * len(cfg["stimulus_selected_input_ids"])
    cfg["stimulus_start_clock_steps"] = [10, 60]
    cfg["stimulus_end_clock_steps"] = [30, 90]
    cfg["stimulus_input_values"] = [2e-9, 2e-9]
    ne, ni = 400, 100
    rng = np.random.default_rng(11)
    shapes = [(ne, ne), (ne, ni), (ni, ne), (ni, ni)]
    wmax = [0.5e-9, 1e-9, 2e-9, 2e-9]
    ws = [(rng.uniform(0, w, sh) * (rng.uniform(size=sh) < 0.3)).astype(np.float32) for sh, w in zip(shapes, wmax)]
    target = np.add.accumulate(ws[0], axis=0, dtype=np.float32)[-1].astype(np.float64)
    from helpers import stimulus_schedule
    p = O.params_from_config(cfg, gemv_row_block=(4, 4, 4, 4), lazy_normalize=True, lazy_tile_order=tile,
                             gemv_part_chunks=(8, 8, 8, 8) if tile else (0, 0, 0, 0), resident=(res, 1 if res else 0))
    sim = O.OracleSim(p, ws, stim=stimulus_schedule(cfg))
    spikes = 0
    for step in range(120):
        sim.step(1)
        spikes += int(sim.state()["spikes"][:ne].sum())
        if step % 10 == 9:
            w = sim.weights(0).astype(np.float64)
            ok = target > 0
            rel = np.abs(w.sum(axis=1)[ok] - target[ok]) / target[ok]
            assert np.max(rel) <= 2e-6, (step, float(np.max(rel)))
    assert spikes > 100


def test_lazy_normalisation_tracks_faithful():
    cfg = small_config(n_e_side=16, n_i_side=8, steps=2000)
    eng_cfg = cfg
    ne, ni = 256, 64
    rng = np.random.default_rng(9)
    shapes = [(ne, ne), (ne, ni), (ni, ne), (ni, ni)]
    wmax = [0.5e-9, 1e-9, 2e-9, 2e-9]
    ws = [(rng.uniform(0, w, s) * (rng.uniform(size=s) < 0.2)).astype(np.float32) for s, w in zip(shapes, wmax)]
    stim = None
    a = O.OracleSim(O.params_from_config(eng_cfg, gemv_row_block=(32, 32, 32, 32)), ws, stim=stim)
    b = O.OracleSim(O.params_from_config(eng_cfg, gemv_row_block=(32, 32, 32, 32), lazy_normalize=True), ws, stim=stim)
    # The persistent step's orders: chunk sums in parts of 8, lazy row sums in column-tile order.
    c = O.OracleSim(O.params_from_config(eng_cfg, gemv_row_block=(4, 4, 4, 4), lazy_normalize=True,
                                         gemv_part_chunks=(8, 8, 8, 8), lazy_tile_order=True), ws, stim=stim)
    # The resident step's orders: rows over thread lanes and waves, lazy row sums by column blocks.
    d = O.OracleSim(O.params_from_config(eng_cfg, lazy_normalize=True, resident=(16, 6)), ws, stim=stim)
    worst = [0.0, 0.0, 0.0]
    for step in range(300):
        a.step(1)
        b.step(1)
        c.step(1)
        d.step(1)
        sa, sb, sc, sd = a.state(), b.state(), c.state(), d.state()
        worst[0] = max(worst[0], float(np.max(np.abs(sa["V"] - sb["V"]) / np.abs(sa["V"]))))
        worst[1] = max(worst[1], float(np.max(np.abs(sa["V"] - sc["V"]) / np.abs(sa["V"]))))
        worst[2] = max(worst[2], float(np.max(np.abs(sa["V"] - sd["V"]) / np.abs(sa["V"]))))
        assert np.array_equal(sa["spikes"], sc["spikes"]) and np.array_equal(sa["spikes"], sd["spikes"])
    wa, wb, wc, wd = (x.weights(0).astype(np.float64) for x in (a, b, c, d))
    assert np.max(np.abs(wa - wb)) <= 1e-5 * np.max(np.abs(wa))
    assert np.max(np.abs(wa - wc)) <= 1e-5 * np.max(np.abs(wa))
    assert np.max(np.abs(wa - wd)) <= 1e-5 * np.max(np.abs(wa))
    assert max(worst) <= 1e-5, worst
