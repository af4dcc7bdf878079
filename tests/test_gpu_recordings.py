"""Recording files (spike_recorder.cpp:40-60, voltage_recorder.cpp:54-77, weight_recorder.cpp:25-51)
written by a 10,001-step run, checked line by line against the oracle's raster and voltages,
including the reference's quirks: the slot-0 line of a spike record labelled C-10000 holds step
C, no spike file lines before step 10,000, voltage files named with uint32 wrap-around at C=0,
the weight dump at step 0.  The fused and component paths must write identical files."""
import filecmp
import json
import os

import numpy as np
import pytest

from helpers import engine, oracle_for, small_config

pytestmark = pytest.mark.gpu
T = 10000


def _run(tmp_path, tag, flags):
    cfg = small_config(n_e_side=12, n_i_side=6, steps=T + 1)
    cfg["output_directory"] = os.path.join(tmp_path, tag)
    path = os.path.join(tmp_path, f"{tag}.json")
    json.dump(cfg, open(path, "w"))
    sim = engine().Simulation(path, flags=("write_files", "quiet") + flags)
    sim.run(T + 1)
    sim.synchronize()
    sim.close()
    return cfg, path


def _spike_lines(path):
    lines = open(path).read().split("\n")
    assert lines[0] == "clock_step, neuron_ids"
    return [ln for ln in lines[1:] if ln]


def test_recordings_match_oracle_and_paths_agree(tmp_path):
    cfg, path = _run(str(tmp_path), "fused", ())
    _, _ = _run(str(tmp_path), "component", ("component_path",))
    out_f = cfg["output_directory"]
    out_c = os.path.join(str(tmp_path), "component")
    # Oracle raster and voltages of the same run.
    probe = engine().Simulation(path, flags=("no_monitors", "quiet"))
    orc = oracle_for(path, probe)
    probe.close()
    ne = cfg["num_excitatory_neurons"]
    raster, volts = {}, {}
    vids = cfg["voltage_recording_neuron_ids"]
    for step in range(T + 1):
        orc.step(1)
        st = orc.state()
        raster[step] = np.nonzero(st["spikes"])[0]
        volts[step] = st["V"][vids].copy()
    # Spikes: 10,000 lines labelled 0..9999; line 0 holds step 10,000.
    for grp, fname, lo, hi in (("e", "spikes_e.txt", 0, ne), ("i", "spikes_i.txt", ne, orc.N)):
        lines = _spike_lines(os.path.join(out_f, "recordings", fname))
        assert len(lines) == T
        for inner, ln in enumerate(lines):
            label, ids = ln.split(", ", 1)
            assert int(label) == inner
            src = T if inner == 0 else inner
            expect = [str(n - lo) for n in raster[src] if lo <= n < hi]
            assert ids == "".join(i + " " for i in expect), (grp, inner)
        assert filecmp.cmp(os.path.join(out_f, "recordings", fname), os.path.join(out_c, "recordings", fname),
                           shallow=False)
    # Voltages: record at C=0 writes names (0-10000)+h+1 in uint32 (4294957297.. then _step0),
    # record at C=10000 writes _step1.._step10000; file K holds V of step K-1, except
    # K = C-9999 which holds step C.
    rec = os.path.join(out_f, "recordings")
    names = [f for f in os.listdir(rec) if f.startswith("voltages_step0_step")]
    assert len(names) == 2 * T  # 9,999 wrapped names + "_step0" at C=0, then _step1.._step10000
    for K in (1, 2, 500, 9999, 10000):
        txt = open(os.path.join(rec, f"voltages_step0_step{K}.txt")).read().split("\n")
        assert txt[0] == "neuron_id, voltage"
        src = T if K == 1 else K - 1
        for q, nid in enumerate(vids):
            vid, v = txt[1 + q].split(", ")
            assert int(vid) == nid
            assert float(v) == pytest.approx(float(volts[src][q]), rel=1e-5, abs=1e-9)
    wrap = (2 ** 32 - T + 1)
    assert os.path.exists(os.path.join(rec, f"voltages_step0_step{wrap}.txt"))
    for name in ("voltages_step0_step1.txt", "voltages_step0_step777.txt"):
        assert filecmp.cmp(os.path.join(rec, name), os.path.join(out_c, "recordings", name), shallow=False)
    # Weight dump at step 0: the initial E->E matrix, every entry, "%g" formatting.
    wfile = os.path.join(out_f, "recordings", "weights_e_to_e_step0.txt")
    lines = open(wfile).read().split("\n")
    assert lines[0] == "pre_neuron_id, post_neuron_id, weight"
    w0 = engine().generate_weight_matrix(path, 0)
    assert len([ln for ln in lines[1:] if ln]) == w0.size
    for k in (0, 1, 77, w0.size - 1):
        i, j, w = lines[1 + k].split(", ")
        assert (int(i), int(j)) == divmod(k, w0.shape[1])
        assert w == "%g" % w0.flat[k]
    # Input files: locations, weights and stimulus schedules exist with the reference headers.
    assert open(os.path.join(out_f, "locations", "group_e_locations.txt")).readline() == \
        "neuron_id, locationdim_0, locationdim_1\n"
    assert open(os.path.join(out_f, "weights", "weights_e_to_e_step0.txt")).readline() == \
        "pre_neuron_id, post_neuron_id, weight\n"
    assert os.path.exists(os.path.join(out_f, "stimulus_schedules", "timestep_event_counts.txt"))


@pytest.mark.parametrize("step", ["graph", "resident"])
def test_recording_pipeline_writes_the_same_files(tmp_path, step):
    # The writer-thread pipeline (default: the rings' copies queued on the stream, the files
    # written while the next interval runs) and recordings on the calling thread at each record
    # step write byte-identical files; two intervals and the record step at 0.
    flags = ("write_files", "quiet") + (("resident", "lazy_normalize") if step == "resident" else ())
    dirs = []
    for piped in (True, False):
        cfg = small_config(n_e_side=12, n_i_side=6, steps=2 * T + 1)
        cfg["output_directory"] = os.path.join(str(tmp_path), f"{step}_{piped}")
        path = os.path.join(str(tmp_path), f"{step}_{piped}.json")
        json.dump(cfg, open(path, "w"))
        sim = engine().Simulation(path, flags=flags)
        sim.set_async_recording(piped)
        sim.run(T)
        sim.run(T + 1)   # the record steps at 10,000 and 20,000 fall inside the second run
        call_s, writer_s, records = sim.record_times()
        print(f"{step} piped={piped}: {records} record steps, {call_s:.3f} s on the calling thread, "
              f"{writer_s:.3f} s in the writer")
        assert records == 3
        assert (writer_s > 0.0) == piped
        sim.close()
        dirs.append(os.path.join(cfg["output_directory"], "recordings"))
    names = sorted(os.listdir(dirs[0]))
    assert names == sorted(os.listdir(dirs[1]))
    assert len([n for n in names if n.startswith("voltages_step0_step")]) == 3 * T
    for name in names:
        assert filecmp.cmp(os.path.join(dirs[0], name), os.path.join(dirs[1], name), shallow=False), name
    lines = _spike_lines(os.path.join(dirs[0], "spikes_e.txt"))
    assert len(lines) == 2 * T
