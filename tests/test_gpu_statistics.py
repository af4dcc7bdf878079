"""One simulated second of the 3,125-neuron plastic network (10,000 steps, lazy normalisation):
the engine's FP32 run against the oracle's FP64 run on the same seeded noise and connectivity.

Rasters agree exactly until FP32/FP64 rounding first flips a spike; after that the two runs are
different trajectories of the same process, so they are compared by their statistics: the E and
I population rates, the mean ISI coefficient of variation of the E neurons, and the E->E weight
distribution (mean, spread, two-sample Kolmogorov-Smirnov distance).  The yardstick for "agree"
is the spread between realisations: a run with another background-noise seed must differ from
the FP64 run by at least as much as the engine does (with small absolute floors).
"""
import json
import os

import numpy as np
import pytest

from helpers import REPO, engine, oracle_for

pytestmark = pytest.mark.gpu

T = 10000
DT = 1e-4


def _stats(raster, ne, w):
    r = np.asarray(raster)
    cvs = []
    for j in range(ne):
        t = np.nonzero(r[:, j])[0]
        if t.size >= 4:
            isi = np.diff(t)
            cvs.append(isi.std() / isi.mean())
    nz = w[w > 0]
    return {"rate_e": r[:, :ne].mean() / DT, "rate_i": r[:, ne:].mean() / DT, "cv": float(np.mean(cvs)),
            "w_mean": float(nz.mean()), "w_std": float(nz.std()), "w": nz}


def _oracle_run(path, precision, row_block, n):
    o = oracle_for(path, None, precision=precision, lazy=True, row_block=row_block)
    ras = np.zeros((T, n), np.uint8)
    for t in range(T):
        o.step(1)
        ras[t] = o.state()["spikes"]
    w = o.weights(0)
    o.close()
    return ras, w


def test_one_second_statistics_match_fp64(tmp_path):
    from scipy.stats import ks_2samp

    path = os.path.join(REPO, "configs", "net3125_stdp.json")
    sim = engine().Simulation(path, flags=("no_monitors", "quiet", "lazy_normalize"))
    ne, n = sim.NE, sim.N
    ras = np.zeros((T, n), np.uint8)
    for t in range(T):
        sim.run(1)
        ras[t] = sim.state()["spikes"]
    gpu = _stats(ras, ne, sim.weights(0))
    rb = tuple(sim.info.gemv_row_block)
    sim.close()

    r64, w64 = _oracle_run(path, 64, rb, n)
    ref = _stats(r64, ne, w64)
    same = np.all(ras == r64, axis=1)
    first_diff = int(np.argmin(same)) if not same.all() else T
    assert first_diff >= 1000, f"rasters differ from step {first_diff}"

    cfg = json.load(open(path))
    cfg["background_noise_seed"] = 777
    alt_path = os.path.join(str(tmp_path), "alt_noise.json")
    json.dump(cfg, open(alt_path, "w"))
    ra, wa = _oracle_run(alt_path, 32, rb, n)
    alt = _stats(ra, ne, wa)

    floors = {"rate_e": 0.01 * ref["rate_e"], "rate_i": 0.01 * ref["rate_i"], "cv": 0.01,
              "w_mean": 0.005 * ref["w_mean"], "w_std": 0.005 * ref["w_std"]}
    report = {}
    for key, floor in floors.items():
        d_gpu, d_alt = abs(gpu[key] - ref[key]), abs(alt[key] - ref[key])
        report[key] = (gpu[key], ref[key], alt[key])
        assert d_gpu <= max(d_alt, floor), (key, report[key])
    ks_gpu = ks_2samp(gpu["w"], ref["w"]).statistic
    ks_alt = ks_2samp(alt["w"], ref["w"]).statistic
    assert ks_gpu <= max(ks_alt, 0.01), (ks_gpu, ks_alt)
    print(f"first raster difference at step {first_diff}; (engine, FP64, other noise): {report}; "
          f"KS engine {ks_gpu:.5f}, other noise {ks_alt:.5f}", flush=True)
