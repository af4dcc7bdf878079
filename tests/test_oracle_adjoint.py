"""The oracle's FP64 adjoint of the differentiable window, pinned by FP64 central differences.

With the spike derivative switched off (spikes held fixed: the exact derivative wherever a
perturbation flips no spike), every gradient — entries of the four weight matrices, g_leak and
tau_decay_e — must match (L(p + h) - L(p - h)) / 2h, h = 1e-6 |p|, to 1e-4 relative, on a
small network with spikes, refractory periods and the background noise running.  Parameters
whose perturbation flips a spike are reported and skipped (none expected at this h).
"""
import json
import os
import sys

import numpy as np
import pytest

from helpers import REPO, engine, small_config

sys.path.insert(0, os.path.join(REPO, "oracle"))
import oracle_py  # noqa: E402


def _setup(tmp_path, seed=5):
    cfg = small_config(n_e_side=8, n_i_side=4, steps=400, stimulus=False, plastic=False, seed=seed)
    cfg["group_e_background_excitatory_scaling_factor"] = 2.0e-9
    path = os.path.join(str(tmp_path), "win.json")
    json.dump(cfg, open(path, "w"))
    dump = engine().config_dump(path)
    params = oracle_py.params_from_config(dump)
    W = [engine().generate_weight_matrix(path, k).astype(np.float64) for k in range(4)]
    NE, NI = dump["num_excitatory_neurons"], dump["num_inhibitory_neurons"]
    N = NE + NI
    rng = np.random.default_rng(seed)
    V = rng.uniform(-0.070, -0.054, N).astype(np.float32)
    V[rng.choice(N, 6, replace=False)] = -0.050  # a few neurons close to firing
    refr = np.zeros(N, np.uint32)
    refr[rng.choice(N, 4, replace=False)] = 3
    r_e = rng.uniform(0, 0.5, NE).astype(np.float32)
    r_i = rng.uniform(0, 0.5, NI).astype(np.float32)
    state = {"V": V, "refr": refr, "bg_ge": rng.uniform(0, 2e-9, N).astype(np.float32),
             "bg_gi": rng.uniform(0, 1e-9, N).astype(np.float32),
             "bre": rng.uniform(0, 0.3, N).astype(np.float32), "bde": rng.uniform(0.3, 0.6, N).astype(np.float32),
             "bri": rng.uniform(0, 0.3, N).astype(np.float32), "bdi": rng.uniform(0.3, 0.6, N).astype(np.float32),
             "r_e": r_e, "d_e": (r_e + rng.uniform(0, 1.0, NE)).astype(np.float32),
             "r_i": r_i, "d_i": (r_i + rng.uniform(0, 1.0, NI)).astype(np.float32), "clock_step": 1234}
    T = 150
    vrec = np.arange(0, NE, 7, dtype=np.uint32)
    target_rates = rng.uniform(0, 30, (T, 2))
    target_v = rng.uniform(-0.070, -0.055, (T, vrec.size))
    return dump, params, W, state, T, vrec, target_rates, target_v


def test_window_adjoint_matches_central_differences(tmp_path):
    dump, params, W, state, T, vrec, tr, tv = _setup(tmp_path)
    gl, taud = float(dump["g_leak"]), float(dump["tau_decay_e"])
    kw = dict(rate_tau=5e-3, beta=0.0, target_rates=tr, target_v=tv)
    base = oracle_py.window(params, W, gl, taud, state, T, vrec, **kw)
    assert base["spikes"].sum() > 5, "the window should contain spikes"

    def loss_at(W2, gl2, taud2):
        r = oracle_py.window(params, W2, gl2, taud2, state, T, vrec, grad=False, **kw)
        return r["loss"], np.array_equal(r["spikes"], base["spikes"])

    checks, skipped = [], []
    # g_leak and tau_decay_e
    for name, value, adj in (("g_leak", gl, base["dg_leak"]), ("tau_decay_e", taud, base["dtau_decay_e"])):
        h = 1e-6 * abs(value)
        lp, okp = loss_at(W, gl + h if name == "g_leak" else gl, taud + h if name == "tau_decay_e" else taud)
        lm, okm = loss_at(W, gl - h if name == "g_leak" else gl, taud - h if name == "tau_decay_e" else taud)
        (checks if okp and okm else skipped).append((name, (lp - lm) / (2 * h), adj))
    # weight entries: nonzero ones with the largest gradients and a few random ones
    rng = np.random.default_rng(1)
    for c in range(4):
        g = np.abs(base["dW"][c]) * (W[c] > 0)
        flat = np.argsort(g.ravel())[::-1][:4].tolist() + rng.choice(W[c].size, 3, replace=False).tolist()
        scale = np.abs(W[c][W[c] > 0]).mean()
        for f in flat:
            i, j = divmod(int(f), W[c].shape[1])
            h = 1e-6 * max(abs(W[c][i, j]), scale)
            Wp = [w.copy() for w in W]
            Wp[c][i, j] += h
            lp, okp = loss_at(Wp, gl, taud)
            Wp[c][i, j] -= 2 * h
            lm, okm = loss_at(Wp, gl, taud)
            (checks if okp and okm else skipped).append((f"W{c}[{i},{j}]", (lp - lm) / (2 * h), base["dW"][c][i, j]))
    assert not skipped, f"perturbations flipped spikes: {skipped}"
    scale = {}
    for name, fd, adj in checks:
        scale[name[:2]] = max(scale.get(name[:2], 0.0), abs(adj))
    compared = 0
    for name, fd, adj in checks:
        den = max(abs(fd), abs(adj))
        if den < 1e-7 * scale[name[:2]]:
            continue  # a (near) zero gradient entry: both sides negligible
        assert abs(fd - adj) <= 1e-4 * den, (name, fd, adj)
        compared += 1
    assert compared >= 20, compared


def test_window_surrogate_changes_only_spike_terms(tmp_path):
    # Far from threshold (a spike-free window) the fast-sigmoid derivative is still nonzero, so
    # the surrogate gradient differs from the exact one only through the spike paths; with the
    # rate terms and the connection traces' influence removed (zero weights, zero rate
    # targets and no spikes) both modes coincide.
    dump, params, W, state, T, vrec, tr, tv = _setup(tmp_path, seed=9)
    state = dict(state)
    state["V"] = np.full_like(state["V"], -0.068)
    state["bg_ge"] = np.zeros_like(state["bg_ge"])
    state["bde"] = state["bre"].copy()
    Wz = [np.zeros_like(w) for w in W]
    gl, taud = float(dump["g_leak"]), float(dump["tau_decay_e"])
    a = oracle_py.window(params, Wz, gl, taud, state, 50, vrec, beta=0.0, target_v=tv[:50])
    b = oracle_py.window(params, Wz, gl, taud, state, 50, vrec, beta=1e-3, target_v=tv[:50])
    assert a["spikes"].sum() == 0
    assert a["dg_leak"] == pytest.approx(b["dg_leak"], rel=1e-12)


@pytest.mark.parametrize("beta", [1e-3, 0.0])
def test_reverse_along_trajectory_matches_fp64_forward_adjoint(tmp_path, beta):
    # The reverse pass along a recorded trajectory (the form the GPU's FP32 run is checked
    # with) must reproduce the FP64 adjoint when the trajectory is the FP64 forward's own,
    # rounded to FP32: only the rounding of the records separates them (no re-simulation).
    dump, params, W, state, T, vrec, tr, tv = _setup(tmp_path)
    gl, taud = float(dump["g_leak"]), float(dump["tau_decay_e"])
    a = oracle_py.window(params, W, gl, taud, state, T, vrec, beta=beta, target_rates=tr, target_v=tv)
    rec = oracle_py.window_records(params, W, gl, taud, state, T, vrec)
    assert np.array_equal(rec["spikes"], a["spikes"])
    b = oracle_py.window_along(params, W, gl, taud, rec, vrec, beta=beta, target_rates=tr, target_v=tv)
    assert b["loss"] == pytest.approx(a["loss"], rel=1e-8)
    for k in range(4):
        n = np.linalg.norm(a["dW"][k])
        if n == 0:
            assert not np.any(b["dW"][k])
            continue
        assert np.linalg.norm(a["dW"][k] - b["dW"][k]) <= 2e-5 * n, k
    assert b["dg_leak"] == pytest.approx(a["dg_leak"], rel=2e-5)
    assert b["dtau_decay_e"] == pytest.approx(a["dtau_decay_e"], rel=2e-5)
