"""Benchmark of the per-step hot path (one MI355X).

Workload: the 10,125-neuron plastic network (8,100 E + 2,025 I on a 1 mm square, four dense
[pre][post] FP32 matrices of 410 MB in total, triplet STDP on E->E and iSTDP on I->E, 1 kHz
Poisson background, seed 1) — configs/net10125_stdp.json.  A "step" is one pass of
System::step_the_clock (evolve -> propagate -> stimulus -> plasticity -> monitors) over the whole
network, with the weights resident in HBM.  Metric: simulated steps per second.

Prints ONE JSON line.  roofline = the propagate (GEMV) kernel's algorithmic bytes per launch
(the four weight matrices plus the pre-synaptic vectors) over its average duration, measured
with HIP events on the engine's stream; cpu_baseline = the oracle's FP32 C++/OpenMP step loop
on this host, timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
HBM_PEAK_GBS = 8000.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--config", default=os.path.join(REPO, "configs", "net10125_stdp.json"))
    ap.add_argument("--normalization", choices=("lazy", "faithful"), default="lazy")
    ap.add_argument("--kernel-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs one network on one GPU; --gpus {args.gpus} is not supported "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        return 2

    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    cfg = eng.config_dump(args.config)
    flags = ["no_monitors", "quiet"] + (["lazy_normalize"] if args.normalization == "lazy" else [])
    sim = eng.Simulation(args.config, flags=flags)
    info = sim.info
    if args.warmup:
        sim.run_timed(args.warmup)
    ms = sim.run_timed(args.steps)  # synchronises before and after, HIP events on the engine stream
    steps_per_s = args.steps / (ms / 1000.0)

    # Dominant kernel: live HIP-event timing around every propagate (GEMV) launch of real steps.
    k_ms, s_ms = sim.time_propagate_kernel(args.kernel_steps)
    k_avg_s = k_ms / 1000.0 / args.kernel_steps
    n_w = sum(int(info.weight_pre[k]) * int(info.weight_post[k]) for k in range(4))
    gemv_bytes = 4 * n_w + 4 * sum(int(info.weight_pre[k]) for k in range(4))
    achieved = gemv_bytes / k_avg_s / 1e9
    state = sim.state()
    rate_hz = float(state["spikes"].mean()) / cfg["timestep"]
    sim.close()

    traffic = None
    tpath = os.path.join(REPO, "profiles", "gemv_traffic_net10125.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        traffic = t.get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.config, cfg, args.cpu_seconds, info)

    n_total = int(info.total_neurons)
    line = {
        "metric": "sim steps/s (10,125-nrn STDP net)",
        "value": round(steps_per_s, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state/weights, f64 GEMV accumulation",
        "data": "synthetic: seeded Philox Poisson background, connectivity from the reference generator",
        "config": {"workload": "net10125_stdp (8,100 E + 2,025 I, 4 dense FP32 matrices, 410 MB, "
                               "triplet STDP + iSTDP, seed 1)",
                   "neurons": n_total, "weight_entries": n_w, "normalization": args.normalization,
                   "kernels_per_step": int(info.kernels_per_step), "parallelism": "single GPU",
                   "rate_hz_last_step": round(rate_hz, 2)},
        "roofline": {"bound": "hbm", "kernel": "k_gemv (propagate)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes_per_launch": gemv_bytes,
                     "avg_kernel_us": round(k_avg_s * 1e6, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline(path: str, cfg: dict, seconds: float, info) -> dict:
    """The oracle's FP32 OpenMP step loop on this host's cores (every GEMV row visited, as the
    reference's cuBLAS call does), on a bounded number of steps of the same network."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py

    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    weights = [eng.generate_weight_matrix(path, k) for k in range(4)]
    p = oracle_py.params_from_config(cfg, gemv_row_block=tuple(info.gemv_row_block), dense_gemv=True)
    orc = oracle_py.OracleSim(p, weights, precision=32)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc.step(1)
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        orc.step(1)
        n += 1
    dt = time.perf_counter() - t0
    orc.close()
    return {"value": round(n / dt, 3), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps of the same network from its initial state, FP32 oracle, "
                      f"{threads} OpenMP threads, {dt:.1f} s"}


if __name__ == "__main__":
    sys.exit(main())
