"""Benchmark of the per-step hot path (one MI355X).

Default workload: the 3,125-neuron plastic network (2,500 E on a 50x50 grid + 625 I on 25x25,
1 mm square, four dense [pre][post] FP32 matrices of 39 MB in total, triplet STDP on E->E and
iSTDP on I->E, 1 kHz Poisson background, seed 0) — configs/net3125_stdp.json.  A "step" is one
pass of System::step_the_clock (evolve -> propagate -> stimulus -> plasticity -> monitors) over
the whole network, with the weights resident in device memory.  Metric: simulated steps per
second.  --config configs/net10125_stdp.json runs the 10,125-neuron (410 MB) network instead.

The step runs as the resident step when the network fits the device (every weight in the
registers of one cooperative launch per run; the 3,125-neuron network does), else as the
graph-captured fused step (--step chooses).

Prints ONE JSON line.  roofline: for the resident step, its one launch is the dominant kernel:
the step's algorithmic bytes (4 x weight entries + 160 B per neuron) times the measured steps/s,
against 8 TB/s (the weights stay in registers, so this is the HBM rate a streaming step would
need to match it); for the graph step, the propagate (GEMV) kernel's algorithmic bytes per
launch over its average duration, HIP events around back-to-back launches on the engine's
stream.  cpu_baseline = the oracle's FP32 C++/OpenMP step loop on this host, timed on a bounded
sample of the same workload.
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


def host_cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    # The oracle's OpenMP threads stay close to their parent (set before its library loads).
    os.environ.setdefault("OMP_PROC_BIND", "close")
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--mode", choices=("plain", "batched-shared", "batched-uneven", "adjoint", "recording"),
                    default="plain",
                    help="plain: one network; batched-shared: 64 instances sharing frozen weights; "
                         "batched-uneven: 48 plastic networks of uneven size; adjoint: gradient of a "
                         "1,000-step window of the 3,125-neuron network (plasticity off)")
    ap.add_argument("--networks", type=int, default=48)
    ap.add_argument("--no-autotune", action="store_true",
                    help="plain: keep the configured GEMV row blocks; batched: keep the default tiles and schedule")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (plain 10000, batched 5001)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps before the timed ones (plain 100, batched 0)")
    ap.add_argument("--settle", type=int, default=2000,
                    help="plain: steps run first to bring the network from its initial state (no spikes yet) to its "
                         "settled firing, before the warm-up")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--config", default=os.path.join(REPO, "configs", "net3125_stdp.json"))
    ap.add_argument("--normalization", choices=("lazy", "faithful"), default="lazy")
    ap.add_argument("--step", choices=("auto", "resident", "graph"), default="auto",
                    help="plain: the resident step (weights in registers, one launch per run) when the network "
                         "fits, else the graph-captured fused step")
    ap.add_argument("--resident-launch", choices=("cooperative", "plain"), default="plain",
                    help="plain, resident step: an ordinary launch (default: the bench owns the GPU, and the "
                         "committed rocprofv3 profiles are of this form) or a cooperative one (the library's "
                         "default; same kernel, about 14 us more per launch)")
    ap.add_argument("--no-graph-compare", action="store_true",
                    help="plain, resident step: skip the short comparison run of the graph step")
    ap.add_argument("--gemv-row-blocks", default=None,
                    help="plain: GEMV row blocks EE,EI,IE,II (e.g. 48,48,32,32) instead of the autotuner's choice")
    ap.add_argument("--kernel-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run 5 plain device copies of this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration_mib:
        return copy_calibration(args.copy_calibration_mib)
    if args.gpus != 1:
        print(f"bench.py: this engine runs one network on one GPU; --gpus {args.gpus} is not supported "
              "(a larger network would shard by post neuron, each GPU holding W[:, its posts])", file=sys.stderr)
        return 2
    if args.mode == "batched-shared":
        args.steps = 5001 if args.steps is None else args.steps
        args.warmup = 0 if args.warmup is None else args.warmup
        return bench_batched(args)
    if args.mode == "batched-uneven":
        args.steps = 10001 if args.steps is None else args.steps
        args.warmup = 0 if args.warmup is None else args.warmup
        return bench_uneven(args)
    if args.mode == "adjoint":
        args.steps = 5 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        return bench_adjoint(args)
    if args.mode == "recording":
        args.steps = 30000 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        return bench_recording(args)
    args.steps = 10000 if args.steps is None else args.steps
    args.warmup = 100 if args.warmup is None else args.warmup
    return bench_plain(args)


def bench_plain(args) -> int:
    """One network: the resident step when the network fits the device (every weight in
    registers, one cooperative launch per run), else the graph-captured fused step."""
    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    # Monitors on: the step writes the spike-bit and voltage rings every step, as the reference's
    # step historicizes every step (record() file I/O only happens in System::run, not here).
    flags = ["quiet"] + (["lazy_normalize"] if args.normalization == "lazy" else [])
    cfg = eng.config_dump(args.config)
    name = os.path.splitext(os.path.basename(args.config))[0]
    n_e, n_i = int(cfg["num_excitatory_neurons"]), int(cfg["num_inhibitory_neurons"])
    resident = args.step in ("auto", "resident") and args.normalization == "lazy"
    sim = None
    if resident:
        try:
            sim = eng.Simulation(args.config, flags=flags + ["resident"])
            sim.set_resident_cooperative(args.resident_launch == "cooperative")
        except eng.BnError as e:
            if args.step == "resident" or "does not fit" not in str(e):
                raise
            resident = False
    config_path, tune, steps_per_graph = args.config, None, None
    if not resident:
        if args.gemv_row_blocks:
            import tempfile
            c = json.load(open(args.config))
            c["gemv_row_blocks"] = [int(r) for r in args.gemv_row_blocks.split(",")]
            config_path = os.path.join(tempfile.mkdtemp(prefix="bn_bench_rows_"), os.path.basename(args.config))
            json.dump(c, open(config_path, "w"))
        elif not args.no_autotune:
            # GEMV row blocks (the propagate's tile shape and FP64 summation order): the default
            # and a few variants, each settled and timed briefly; the timed run uses the fastest.
            import tempfile
            wl = importlib.import_module("example-cuda-spiking-nn_amd.workloads")
            tune_dir = tempfile.mkdtemp(prefix="bn_bench_rows_")
            config_path, best_rows, tune_us = wl.autotune_plain_row_blocks(eng, args.config, tune_dir, flags)
            tune = {"chosen": best_rows, "us_per_step": {k: round(v, 3) for k, v in tune_us.items()},
                    "settle_steps": 1000, "timed_steps": 3000}
        sim = eng.Simulation(config_path, flags=flags)
        # Steps per hipGraph: 500 for long runs (0.8% faster than 100 on the 3,125-neuron
        # network, 0.3% on the 10,125-neuron one); a shorter timed run is one graph of its own
        # length, rounded down to an even count (the overlapped form), any odd step eager.
        steps_per_graph = 500 if args.steps >= 500 else max(2, args.steps - args.steps % 2)
        sim.set_steps_per_graph(steps_per_graph)
    info = sim.info
    if args.settle:
        sim.run_timed(args.settle)
    if args.warmup:
        sim.run_timed(args.warmup)
    before = sim.get_info()
    first_timed = int(before.clock_step)
    ms = sim.run_timed(args.steps)  # synchronises before and after, HIP events on the engine stream
    graph_launches = int(sim.get_info().graph_launches) - int(before.graph_launches)
    steps_per_s = args.steps / (ms / 1000.0)
    # Firing over the timed window, from the device spike ring (the last 10,000 steps at most).
    window = min(args.steps, 10000)
    raster = sim.spike_raster(first_timed + args.steps - window, window)
    rate_e = float(raster[:, :n_e].mean()) / cfg["timestep"]
    rate_i = float(raster[:, n_e:].mean()) / cfg["timestep"]
    rate_all = float(raster.mean()) / cfg["timestep"]
    del raster
    n_w = sum(int(info.weight_pre[k]) * int(info.weight_post[k]) for k in range(4))
    n_total = int(info.total_neurons)
    step_bytes = 4 * n_w + 160 * n_total
    gemv_bytes = 4 * n_w + 4 * sum(int(info.weight_pre[k]) for k in range(4))
    traffic_file = os.path.join(REPO, "profiles", f"{'k_resident' if resident else 'k_gemv'}_{name}_hbm_traffic.json")
    traffic = None
    if os.path.exists(traffic_file):
        t = json.load(open(traffic_file))
        if t.get("neurons") == n_total and t.get("weight_entries") == n_w and (
                resident or t.get("gemv_row_blocks") == [int(r) for r in info.gemv_row_block]):
            traffic = t.get("hbm_bytes_per_launch") if not resident else t.get("hbm_bytes_per_step")
            if resident and traffic is not None:
                traffic = int(traffic) * args.steps   # the timed launch's steps
    graph_cmp = None
    if resident:
        # The dominant (only) kernel is the resident launch itself: the whole timed run.  Its
        # algorithmic bytes are the step's (the weights read once per step and ~40 state words per
        # neuron), although the weights stay in registers: the figure is the HBM rate a streaming
        # step would need to match it.
        achieved = step_bytes * steps_per_s / 1e9
        roofline = {"bound": "hbm", "kernel": f"k_resident (one {args.resident_launch} launch per run)",
                    "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "traffic_unit": "HBM bytes of the timed launch (rocprofv3 FETCH_SIZE + WRITE_SIZE per step of a "
                                    "10,000-step launch, times its steps; the weights stay in registers)",
                    "algorithmic_bytes_per_step": step_bytes,
                    "algorithmic_bytes_per_launch": step_bytes * args.steps,
                    "avg_kernel_us": round(ms * 1000.0, 2), "us_per_step": round(ms * 1000.0 / args.steps, 3),
                    "resident_blocks": int(info.resident_blocks), "resident_rows": int(info.resident_rows),
                    "step_bytes_B_step": step_bytes, "step_GBs": round(achieved, 1),
                    "step_frac_of_peak": round(achieved / HBM_PEAK_GBS, 4)}
        sim.close()
        if not args.no_graph_compare:
            # The graph-captured fused step on the same network (engine default row blocks), for
            # comparison: settled, 2,000 steps in graphs of 500.
            g = eng.Simulation(args.config, flags=flags)
            g.set_steps_per_graph(500)
            g.run_timed(args.settle or 1000)
            gms = g.run_timed(2000)
            graph_cmp = {"steps_per_s": round(2000 / (gms / 1000.0), 1), "us_per_step": round(gms * 1000.0 / 2000, 3),
                         "steps": 2000, "steps_per_graph": 500}
            g.close()
    else:
        # Dominant kernel: HIP events around back-to-back propagate (GEMV) launches over the
        # current weights (the simulation's pending partial sums are saved and restored).
        k_ms, s_ms = sim.time_propagate_kernel(args.kernel_steps)
        k_avg_s = k_ms / 1000.0 / args.kernel_steps
        achieved = gemv_bytes / k_avg_s / 1e9
        sim.close()
        # The applicable read bandwidth for this weight set: a plain streaming read kernel over a
        # working set of the same size (MALL-resident below 256 MB, HBM-bound above), and the
        # same working set read 200 times inside one launch (the steady rate of the level of the
        # memory hierarchy that holds it, without launch ramp and drain).
        read_gbs = eng.measure_read_bandwidth(4 * n_w, 50)
        read_gbs_one = eng.measure_read_bandwidth_one_launch(4 * n_w, 200)
        roofline = {"bound": "hbm", "kernel": "k_gemv (propagate)", "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "algorithmic_bytes_per_launch": gemv_bytes,
                    "avg_kernel_us": round(k_avg_s * 1e6, 2),
                    "measured_read_GBs_same_working_set": round(read_gbs, 1),
                    "measured_read_GBs_same_working_set_one_launch_200_passes": round(read_gbs_one, 1),
                    "frac_of_measured_read": round(achieved / read_gbs, 4),
                    "step_bytes_B_step": step_bytes,
                    "step_GBs": round(step_bytes * steps_per_s / 1e9, 1),
                    "step_frac_of_peak": round(step_bytes * steps_per_s / 1e9 / HBM_PEAK_GBS, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(config_path, cfg, args.cpu_seconds, info)

    line = {
        "metric": f"sim steps/s ({n_total:,}-nrn STDP net)",
        "value": round(steps_per_s, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_steps": args.settle,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state/weights, f64 GEMV accumulation",
        "data": "synthetic: seeded Philox Poisson background, connectivity from the reference generator",
        "config": {"workload": f"{name} ({n_e:,} E + {n_i:,} I, 4 dense FP32 matrices, {4 * n_w / 1e6:.1f} MB, "
                               f"triplet STDP + iSTDP, seed {cfg['rng_seed']}, {args.normalization} normalisation)",
                   "neurons": n_total, "weight_entries": n_w, "normalization": args.normalization,
                   "step": "resident" if resident else "graph",
                   "resident_launch": args.resident_launch if resident else None,
                   "steps_per_graph": steps_per_graph, "graph_launches_timed": graph_launches,
                   "kernels_per_step": int(info.kernels_per_step), "parallelism": "single GPU",
                   "gemv_row_blocks": None if resident else [int(r) for r in info.gemv_row_block],
                   "row_block_autotune": tune, "graph_step_same_network": graph_cmp,
                   "timed_window_rate_hz": {"all": round(rate_all, 3), "excitatory": round(rate_e, 3),
                                            "inhibitory": round(rate_i, 3), "steps": window},
                   "first_timed_step": first_timed},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_recording(args) -> int:
    """The recordings' cost on the plain workload: the reference's files (the E and I spike
    files every 10,000 steps, one voltage file per step, the E->E weight dump every 100,000 steps)
    written by the writer-thread pipeline, on the calling thread at each record step, and the
    same run without files.  Step 0 (a record step with the weight dump) runs first, untimed;
    the timed run covers `steps` steps and their record steps."""
    import shutil
    import tempfile
    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    cfg = json.load(open(args.config))
    name = os.path.splitext(os.path.basename(args.config))[0]
    flags = ["quiet"] + (["lazy_normalize", "resident"] if args.step != "graph" else ["lazy_normalize"])
    forms = {}
    for form in ("no files", "pipeline", "calling thread"):
        d = tempfile.mkdtemp(prefix="bn_rec_")
        c = dict(cfg)
        c["output_directory"] = os.path.join(d, "out")   # cleaned at creation when files are written
        path = os.path.join(d, "config.json")
        json.dump(c, open(path, "w"))
        sim = eng.Simulation(path, flags=flags + ([] if form == "no files" else ["write_files"]))
        if "resident" in flags:
            sim.set_resident_cooperative(args.resident_launch == "cooperative")
        sim.set_async_recording(form != "calling thread")
        sim.run(args.warmup)
        c0, w0, r0 = sim.record_times()
        t0 = time.perf_counter()
        sim.run(args.steps)
        sim.synchronize()
        wall = time.perf_counter() - t0
        c1, w1, r1 = sim.record_times()
        sim.close()
        files = sum(len(f) for _, _, f in os.walk(d))
        shutil.rmtree(d, ignore_errors=True)
        forms[form] = {"steps_per_s": round(args.steps / wall, 1), "wall_s": round(wall, 4),
                       "record_steps": r1 - r0, "calling_thread_s": round(c1 - c0, 4),
                       "writer_thread_s": round(w1 - w0, 4), "files": files}
    base = forms["no files"]["wall_s"]
    piped, sync = forms["pipeline"], forms["calling thread"]
    n_rec = max(1, piped["record_steps"])
    line = {
        "metric": "sim steps/s with the reference's recordings written (3,125-nrn STDP net)",
        "value": piped["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(piped["wall_s"] * 1000.0 / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 state/weights, f64 GEMV accumulation",
        "data": "synthetic: seeded Philox Poisson background; files to a temporary directory",
        "config": {"workload": f"{name}, {'resident' if 'resident' in flags else 'graph'} step, spike and voltage "
                               f"files every {cfg.get('recording_interval_steps', 10000)} steps",
                   "forms": forms,
                   "record_cost_per_interval_s": {
                       "calling_thread": round((sync["wall_s"] - base) / n_rec, 4),
                       "pipeline_wall": round((piped["wall_s"] - base) / n_rec, 4),
                       "pipeline_writer": round(piped["writer_thread_s"] / n_rec, 4)},
                   "interval_compute_s": round(base / n_rec, 4)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


MFMA_FP32_PEAK_TFLOPS = 157.3


def bench_batched(args) -> int:
    """The shared-weight batch: `instances` copies of the network with plasticity off and one
    weight set, noise key = instance index, a 2,000-4,000-step pulse of 200 pA x U(0,1) on 100
    excitatory neurons per instance (the recorded per-instance draw); `steps` steps timed
    after `warmup`.  Metric: instance-steps/s.  Beside it: the same network run as one plain
    instance on the GPU (the sequential alternative), and the GEMM's FP32 matrix-core rate."""
    import tempfile

    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    base = json.load(open(args.config))
    base["e_e_plasticity_enabled"] = False
    base["i_e_plasticity_enabled"] = False
    base["simulation_steps"] = max(int(base.get("simulation_steps", 0)), args.warmup + args.steps)
    n_e = int(base["num_excitatory_neurons"])
    insts = []
    for k in range(args.instances):
        ids, amp = eng.batch_workload_instance(k, n_e, 100, 200e-12)
        insts.append({"noise_seed": k, "stimulus_ids": ids, "stimulus_amplitude": amp,
                      "stimulus_start": 2000, "stimulus_end": 4000})
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "batched.json")
        json.dump(base, open(path, "w"))
        batch = eng.BatchSimulation(path, insts)
        tuned_rows, times = None, {}
        if not args.no_autotune:
            # GEMM autotuner: rows per wave chunk (split-K depth) by back-to-back GEMM time.
            for rows in (64, 96, 128, 152, 192, 256, 320):
                batch.set_gemm_rows(rows)
                batch.time_gemm(5)
                times[rows] = batch.time_gemm(50) / 50 * 1000.0
            tuned_rows = min(times, key=times.get)
            batch.set_gemm_rows(tuned_rows)
        info = batch.get_info()
        if args.warmup:
            batch.run_timed(args.warmup)
        ms = batch.run_timed(args.steps)
        gemm_ms = batch.time_gemm(args.kernel_steps) / args.kernel_steps
        batch.close()
        # Sequential alternative: one plain instance (same network, plasticity off), as the graph
        # step and as the resident step; the faster of the two.
        plain_cfg = dict(base)
        plain_cfg.update({"background_noise_seed": 0, "stimulus_selected_input_ids": [int(i) for i in insts[0]["stimulus_ids"]],
                          "stimulus_selected_input_post_group_offsets": [0] * 100,
                          "stimulus_start_clock_steps": [2000], "stimulus_end_clock_steps": [4000],
                          "stimulus_input_values": [float(insts[0]["stimulus_amplitude"])]})
        ppath = os.path.join(d, "plain.json")
        json.dump(plain_cfg, open(ppath, "w"))
        sim = eng.Simulation(ppath, flags=("no_monitors", "quiet"))
        graph_ms = sim.run_timed(args.steps)
        sim.close()
        sim = eng.Simulation(ppath, flags=("no_monitors", "quiet", "lazy_normalize", "resident"))
        sim.set_resident_cooperative(False)
        resident_ms = sim.run_timed(args.steps)
        sim.close()
        plain_ms = min(graph_ms, resident_ms)
    inst_steps_per_s = args.instances * args.steps / (ms / 1000.0)
    plain_inst_steps_per_s = args.steps / (plain_ms / 1000.0)
    n_w = sum(int(info.weight_pre[k]) * int(info.weight_post[k]) for k in range(4))
    B = int(info.padded_instances)
    flops = 2.0 * B * n_w
    tflops = flops / (gemm_ms / 1000.0) / 1e12
    gemm_bytes = 4 * n_w + 4 * B * sum(int(info.weight_pre[k]) for k in range(4))
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_batched(base, args.cpu_seconds)
    line = {
        "metric": "batched instance-steps/s (64 x 3,125-nrn nets, shared frozen weights)",
        "value": round(inst_steps_per_s, 1),
        "unit": "instance-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (GEMM: FP32 matrix cores, FP32 accumulation)",
        "data": "synthetic: per-instance Philox noise keys 0..63, recorded per-instance stimulus draws",
        "config": {"workload": f"batched_shared ({args.instances} instances of {os.path.basename(args.config)} "
                               "with plasticity off, one weight set)",
                   "instances": args.instances, "padded_instances": B, "neurons_per_instance": int(info.total_neurons),
                   "gemm_rows_per_wave": int(info.gemm_rows_per_wave), "kernels_per_step": int(info.kernels_per_step),
                   "gemm_autotune_us": {str(k): round(v, 2) for k, v in times.items()},
                   "parallelism": "single GPU"},
        "vs_sequential_plain": round(inst_steps_per_s / plain_inst_steps_per_s, 2),
        "sequential_plain_instance_steps_per_s": round(plain_inst_steps_per_s, 1),
        "sequential_plain_forms_steps_per_s": {"graph": round(args.steps / (graph_ms / 1000.0), 1),
                                               "resident": round(args.steps / (resident_ms / 1000.0), 1)},
        "roofline": {"bound": "mfma", "kernel": "k_batch_gemm", "achieved": round(tflops, 2),
                     "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / MFMA_FP32_PEAK_TFLOPS, 4),
                     "traffic": None, "algorithmic_flops_per_launch": flops,
                     "algorithmic_bytes_per_launch": gemm_bytes,
                     "bandwidth_GBs": round(gemm_bytes / (gemm_ms / 1000.0) / 1e9, 1),
                     "bandwidth_frac": round(gemm_bytes / (gemm_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                     "avg_kernel_us": round(gemm_ms * 1000.0, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_uneven(args) -> int:
    """48 independent plastic networks (N_E drawn from {400, 900, 1600, 2500}, triplet STDP +
    iSTDP, own seeds), one launch per phase for all of them; the GEMV row block of each size is
    chosen by the autotuner first.  Metric: network-steps/s.  Per size: the count and plain
    single-network runs of that size on the GPU, the graph step and the resident step; the faster
    of the two is the sequential alternative."""
    import tempfile

    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    wl = importlib.import_module("example-cuda-spiking-nn_amd.workloads")
    total = args.warmup + args.steps
    cfgs = wl.uneven_plastic_configs(count=args.networks, steps=total + 4000)
    sizes = [c["num_excitatory_neurons"] for c in cfgs]
    with tempfile.TemporaryDirectory() as d:
        tuned = {}
        cap_times = {}
        if not args.no_autotune:
            # Row blocks tuned on the whole multi-network step: one cap for all sizes.
            cap, cap_times = wl.autotune_row_block_cap(eng, cfgs, d)
            for c in cfgs:
                n_e = c["num_excitatory_neurons"]
                tuned[n_e] = (wl.row_block_for(n_e, cap), {})
                c["gemv_row_blocks"] = [tuned[n_e][0]] * 4
        paths = []
        for k, c in enumerate(cfgs):
            paths.append(os.path.join(d, f"net{k}.json"))
            json.dump(c, open(paths[-1], "w"))
        multi = eng.MultiSimulation(paths, flags=("lazy_normalize",))
        sched = None
        if not args.no_autotune:
            # Tuning steps advance the networks; the timed run then starts from that state.
            sched = wl.autotune_schedule(multi)
        minfo = multi.get_info()
        if args.warmup:
            multi.run_timed(args.warmup)
        ms = multi.run_timed(args.steps)
        k_ms = multi.time_propagate(args.kernel_steps) / args.kernel_steps
        n_w = 0
        for k in range(len(cfgs)):
            inf = multi.networks[k].info
            n_w += sum(int(inf.weight_pre[c]) * int(inf.weight_post[c]) for c in range(4))
        n_pre = sum(sum(int(multi.networks[k].info.weight_pre[c]) for c in range(4)) for k in range(len(cfgs)))
        rates = {}
        for k, c in enumerate(cfgs):
            rates[c["num_excitatory_neurons"]] = rates.get(c["num_excitatory_neurons"], [])
            rates[c["num_excitatory_neurons"]].append(k)
        per_size = {}
        for n_e, idx in sorted(rates.items()):
            # The sequential alternative runs each network as a plain run would: the engine's
            # default row blocks for a single network, not the batch's.
            plain_cfg = dict(cfgs[idx[0]])
            plain_cfg.pop("gemv_row_blocks", None)
            plain_path = os.path.join(d, f"plain_{n_e}.json")
            json.dump(plain_cfg, open(plain_path, "w"))
            plain_steps = min(args.steps, 2000)
            rates_of = {}
            for form, extra in (("graph", ()), ("resident", ("resident",))):
                sim = eng.Simulation(plain_path, flags=("no_monitors", "quiet", "lazy_normalize") + extra)
                if extra:
                    sim.set_resident_cooperative(False)
                sim.run_timed(500)
                pms = sim.run_timed(plain_steps)
                sim.close()
                rates_of[form] = round(plain_steps / (pms / 1000.0), 1)
            per_size[str(n_e)] = {"networks": len(idx), "plain_single_steps_per_s": max(rates_of.values()),
                                  "plain_graph_steps_per_s": rates_of["graph"],
                                  "plain_resident_steps_per_s": rates_of["resident"],
                                  "gemv_row_block": tuned.get(n_e, (None,))[0]}
        multi.close()
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_uneven(cfgs, rates, args.cpu_seconds)
    net_steps_per_s = len(cfgs) * args.steps / (ms / 1000.0)
    seq_s_per_step = sum(len(v) / per_size[str(n_e)]["plain_single_steps_per_s"] for n_e, v in rates.items())
    seq_rate = len(cfgs) / seq_s_per_step
    seq_graph = len(cfgs) / sum(len(v) / per_size[str(n_e)]["plain_graph_steps_per_s"] for n_e, v in rates.items())
    gemv_bytes = 4 * n_w + 4 * n_pre
    achieved = gemv_bytes / (k_ms / 1000.0) / 1e9
    line = {
        "metric": "batched network-steps/s (48 uneven plastic nets)",
        "value": round(net_steps_per_s, 1),
        "unit": "network-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state/weights, f64 GEMV accumulation",
        "data": "synthetic: per-network seeds 0..47, sizes from a recorded Mersenne Twister draw (seed 2024)",
        "config": {"workload": f"batched_uneven ({len(cfgs)} plastic networks, N_E in {{400,900,1600,2500}}, "
                               "triplet STDP + iSTDP, lazy normalisation)",
                   "networks": len(cfgs), "weight_entries": n_w, "kernels_per_step": int(minfo.kernels_per_step),
                   "per_size": per_size, "row_block_cap_us_per_step": {str(k): round(v, 2) for k, v in cap_times.items()},
                   "schedule": sched, "groups": int(minfo.num_groups),
                   "parallelism": "single GPU"},
        "vs_sequential_plain": round(net_steps_per_s / seq_rate, 2),
        "sequential_plain_network_steps_per_s": round(seq_rate, 1),
        "vs_sequential_plain_graph_step": round(net_steps_per_s / seq_graph, 2),
        "sequential_plain_graph_step_network_steps_per_s": round(seq_graph, 1),
        "roofline": {"bound": "hbm", "kernel": "k_gemv_multi (propagate, all networks)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": gemv_bytes, "avg_kernel_us": round(k_ms * 1000.0, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_adjoint(args) -> int:
    """Differentiable window: the 3,125-neuron network with plasticity off, 1,000 steps from the
    state after a 1,000-step warm-up, loss against the targets of the E->E x 1.1 network,
    gradient of all four weight matrices, g_leak and tau_decay_e by the hand-written adjoint
    (surrogate slope 1/(1 mV), checkpoints every 100 steps).  A "step" of the bench is one
    forward + backward pass over the window; metric: seconds per 1,000 simulated steps."""
    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    wl = importlib.import_module("example-cuda-spiking-nn_amd.workloads")
    T, S, warm = 1000, 100, 1000
    sim, win, base, _ = wl.adjoint_workload(eng, steps=T, warmup=warm, checkpoint_interval=S)
    for _ in range(args.warmup):
        win.forward(outputs=False)
        win.backward(download=False)
    # Timed: forward + backward with the gradient left in device memory (its download to the
    # host is reported separately).
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f = win.forward(outputs=False)
        g = win.backward(download=False)
    dt = time.perf_counter() - t0
    dev = win.get_info()
    t2 = time.perf_counter()
    g["dW"] = [win.gradient(k) for k in range(4)]
    dt_dl = time.perf_counter() - t2
    # Forward alone, for the split.
    t1 = time.perf_counter()
    for _ in range(args.steps):
        win.forward(outputs=False)
    dt_fwd = time.perf_counter() - t1
    s_per_1k = dt / args.steps * 1000.0 / T
    n_w = sum(w.size for w in base)
    n = sim.N
    # Algorithmic bytes per window: the weight set read by every forward step, every reverse
    # step (transposed product) and every recomputed step (all but the last segment, which the
    # forward records), plus 160 B of state per neuron per pass, and dW read-modify-written once
    # per segment.
    b_step = 4 * n_w + 160 * n
    n_seg = -(-T // S)
    recomputed = (n_seg - 1) * S
    alg_bytes = (2 * T + recomputed) * b_step + n_seg * 2 * 4 * n_w
    achieved = alg_bytes / (dt / args.steps) / 1e9
    cpu = None if args.no_cpu_baseline else cpu_baseline_adjoint(sim, win, base, args.cpu_seconds)
    gnorm = float(sum((d.astype("float64") ** 2).sum() for d in g["dW"]) ** 0.5)
    line = {
        "metric": "adjoint s/1k steps (3,125-nrn net, plasticity off)",
        "value": round(s_per_1k, 5),
        "unit": "s/1k steps",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1000.0, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state/weights, f64 adjoint reductions",
        "data": "synthetic: seeded Philox Poisson background, targets from the E->E x 1.1 network",
        "config": {"workload": f"adjoint window (net3125 {sim.NE:,} E + {sim.NI:,} I, plasticity off, {T} steps "
                               f"from the state after {warm} steps, 50 recorded E, checkpoints every {S}, "
                               "surrogate slope 1/(1 mV))",
                   "window_steps": T, "checkpoint_interval": S, "neurons": n, "weight_entries": int(n_w),
                   "forward_s_per_1k": round(dt_fwd / args.steps * 1000.0 / T, 5),
                   "device_ms_last_forward": round(dev.forward_ms, 3), "device_ms_last_backward": round(dev.backward_ms, 3),
                   "gradient_download_ms": round(dt_dl * 1000.0, 2),
                   "loss": f["loss"], "grad_norm_w": gnorm, "dg_leak": g["dg_leak"],
                   "dtau_decay_e": g["dtau_decay_e"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "whole window (forward + recompute propagate, transposed "
                                                "products, dW updates)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": int(alg_bytes)},
        "cpu_baseline": cpu,
    }
    win.close()
    sim.close()
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline_adjoint(sim, win, base, seconds: float) -> dict:
    """The oracle's FP64 OpenMP adjoint (forward + reverse, all states kept) on the window's start
    state and weights, over as many steps of the window as fit the time budget."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import oracle_py

    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    st = sim.state()
    rE, dE, _ = sim.connection_traces(0)
    rI, dI, _ = sim.connection_traces(2)
    state = {"V": st["membrane_potentials"], "refr": st["refractory_counter"],
             "bg_ge": st["background_g_excitatory_synaptic_conductances"],
             "bg_gi": st["background_g_inhibitory_synaptic_conductances"],
             "bre": st["background_excitatory_postsynaptic_r_traces"],
             "bde": st["background_excitatory_postsynaptic_d_traces"],
             "bri": st["background_inhibitory_postsynaptic_r_traces"],
             "bdi": st["background_inhibitory_postsynaptic_d_traces"],
             "r_e": rE, "d_e": dE, "r_i": rI, "d_i": dI, "clock_step": sim.get_info().clock_step}
    import tempfile

    wl = importlib.import_module("example-cuda-spiking-nn_amd.workloads")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "adj.json")
        json.dump(wl.adjoint_config(), open(path, "w"))
        cfg = eng.config_dump(path)
    p = oracle_py.params_from_config(cfg)
    W = [b.astype(np.float64) for b in base]
    ids = np.array([int(k * sim.NE // 50) for k in range(50)], np.uint32)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    steps = 20
    while True:
        t0 = time.perf_counter()
        oracle_py.window(p, W, float(cfg["g_leak"]), float(cfg["tau_decay_e"]), state, steps, ids, beta=1e-3)
        dt = time.perf_counter() - t0
        if dt * 2 > seconds or steps >= 1000:
            break
        steps = min(1000, max(steps * 2, int(steps * seconds / max(dt, 1e-3) / 2)))
    return {"value": round(dt / steps * 1000.0, 3), "unit": "s/1k steps", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the window (forward + reverse) from the same start state, FP64 oracle "
                      f"adjoint, {threads} OpenMP threads, {dt:.1f} s"}


def cpu_baseline_uneven(cfgs, by_size: dict, seconds: float) -> dict:
    """The uneven batch on the oracle's FP32 OpenMP loop: one network of each size class stepped
    from its initial state for a share of the time budget (every GEMV row visited), the networks
    of the batch then run one after the other: network-steps/s = networks / sum of the seconds per
    step of each network."""
    import tempfile

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py

    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    per_size, s_per_step, parts = {}, 0.0, []
    budget = seconds / max(1, len(by_size))
    with tempfile.TemporaryDirectory() as d:
        for n_e, idx in sorted(by_size.items()):
            c = dict(cfgs[idx[0]])
            c.pop("gemv_row_blocks", None)
            path = os.path.join(d, f"cpu_{n_e}.json")
            json.dump(c, open(path, "w"))
            cfg = eng.config_dump(path)
            weights = [eng.generate_weight_matrix(path, k) for k in range(4)]
            p = oracle_py.params_from_config(cfg, lazy_normalize=True, dense_gemv=True)
            orc = oracle_py.OracleSim(p, weights, precision=32)
            orc.step(1)
            t0 = time.perf_counter()
            n = 0
            while time.perf_counter() - t0 < budget:
                orc.step(1)
                n += 1
            dt = time.perf_counter() - t0
            orc.close()
            per_size[str(n_e)] = round(n / dt, 2)
            s_per_step += len(idx) * dt / n
            parts.append(f"N_E {n_e}: {n} steps in {dt:.1f} s")
    return {"value": round(len(cfgs) / s_per_step, 3), "unit": "network-steps/s", "cores": threads, "kind": "port",
            "per_size_steps_per_s": per_size,
            "sample": f"one network of each size from its initial state, FP32 oracle (lazy normalisation), "
                      f"{threads} OpenMP threads ({'; '.join(parts)}); the {len(cfgs)} networks one after the other"}


def cpu_baseline_batched(base: dict, seconds: float) -> dict:
    """One instance of the batch on the oracle's FP32 OpenMP loop (instances run one after the
    other on the CPU, so instance-steps/s = steps/s of one instance)."""
    import tempfile

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py

    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cpu.json")
        json.dump(base, open(path, "w"))
        cfg = eng.config_dump(path)
        weights = [eng.generate_weight_matrix(path, k) for k in range(4)]
    p = oracle_py.params_from_config(cfg, gemv_row_block=(0, 0, 0, 0), dense_gemv=True)
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
    return {"value": round(n / dt, 3), "unit": "instance-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps of one instance (plasticity off) from its initial state, FP32 oracle, "
                      f"{threads} OpenMP threads, {dt:.1f} s"}


def copy_calibration(mib: int) -> int:
    """Five device-to-device copies of `mib` MiB: the reference point that scales the HBM byte
    counters (FETCH_SIZE, WRITE_SIZE) of a profiled run to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 5}), flush=True)
    return 0


def cpu_baseline(path: str, cfg: dict, seconds: float, info) -> dict:
    """The oracle's FP32 OpenMP step loop on this host's cores (every GEMV row visited, as the
    reference's cuBLAS call does), on a bounded number of steps of the same network."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py

    eng = importlib.import_module("example-cuda-spiking-nn_amd")
    weights = [eng.generate_weight_matrix(path, k) for k in range(4)]
    p = oracle_py.params_from_config(cfg, gemv_row_block=tuple(info.gemv_row_block), dense_gemv=True)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def timed(precision, budget):
        orc = oracle_py.OracleSim(p, weights, precision=precision)
        orc.step(1)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < budget:
            orc.step(1)
            n += 1
        dt = time.perf_counter() - t0
        orc.close()
        return n, dt

    n, dt = timed(32, seconds)
    n64, dt64 = timed(64, seconds / 3)
    return {"value": round(n / dt, 3), "unit": "steps/s", "cores": threads, "kind": "port",
            "fp64_value": round(n64 / dt64, 3),
            "host": f"{host_cpu_model()}, nproc {os.cpu_count()}, OMP_NUM_THREADS {threads}, "
                    f"OMP_PROC_BIND {os.environ.get('OMP_PROC_BIND')}",
            "sample": f"{n} steps of the same network from its initial state, FP32 oracle, "
                      f"{threads} OpenMP threads, {dt:.1f} s (FP64 oracle: {n64} steps, {dt64:.1f} s)"}


if __name__ == "__main__":
    sys.exit(main())
