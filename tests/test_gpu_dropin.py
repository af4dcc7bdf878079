"""GPU: a reference-side caller of the drop-in boundary against the oracle.

tests/reference_caller.py builds the reference's components on the exported device state
(bn_state_*, in place of GpuSimulationState) with the reference's unpadded weight layout
(weight_matrix_ld = 0: the 625- and 225-column E->I and I->I matrices take the scalar-load GEMV,
post % 4 != 0), and steps them through the drop-in launchers in System::step_the_clock's order
(system.cpp:43-63), historicizing into the [neuron][slot] histories of the monitors.  Bar: every
step bit-identical to the FP32 oracle (faithful normalisation, the state's GEMV row blocks), the
histories equal to the oracle's raster and voltages, the weights equal at the end.
"""
import json
import os

import numpy as np
import pytest

from helpers import load_config, oracle_for, small_config
from reference_caller import ReferenceSystem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["net3125", "small-stimulus"])
def test_reference_caller_bit_exact(tmp_path, which):
    if which == "net3125":
        cfg, steps = load_config("net3125_stdp.json"), 700
    else:
        cfg, steps = small_config(n_e_side=20, n_i_side=15, steps=400), 300
    path = os.path.join(tmp_path, "cfg.json")
    json.dump(cfg, open(path, "w"))
    ref = ReferenceSystem(path)
    n_e, n_i = ref.NE, ref.NI
    assert tuple(ref.buf.weight_matrix_ld) == (n_e, n_i, n_e, n_i)  # unpadded rows
    assert n_i % 4 != 0
    orc = oracle_for(path, row_block=tuple(ref.buf.gemv_row_block))
    raster, v_e = [], []
    for step in range(steps):
        ref.step_the_clock()
        orc.step(1)
        g, o = ref.state(), orc.state()
        for k in ("V", "ge", "gi", "istim", "refr", "spikes"):
            if not np.array_equal(g[k], o[k]):
                bad = np.nonzero(g[k] != o[k])[0]
                raise AssertionError(f"step {step}: {k} differs at {bad.size} neurons, first {bad[:5]}")
        raster.append(o["spikes"])
        v_e.append(o["V"][:n_e])
    raster = np.array(raster)
    assert raster.sum() > 0
    for k in range(4):
        assert np.array_equal(ref.weights_now(k), orc.weights(k)), f"weights {k}"
    # The monitors' histories: slot = step % interval (kernel_historicize_*.cu).
    hist_e, hist_i = ref.spike_history(0), ref.spike_history(1)
    assert np.array_equal(hist_e[:, :steps].T, raster[:, :n_e])
    assert np.array_equal(hist_i[:, :steps].T, raster[:, n_e:])
    assert not hist_e[:, steps:].any()
    assert np.array_equal(ref.voltage_history()[:, :steps].T, np.array(v_e))
    assert not ref.spike_history(0).any()  # flushed by the download
    ref.close()


def test_state_rejects_foreign_views_and_bad_shapes(tmp_path):
    cfg = small_config(n_e_side=8, n_i_side=4, steps=50)
    path = os.path.join(tmp_path, "cfg.json")
    json.dump(cfg, open(path, "w"))
    ref = ReferenceSystem(path)
    import ctypes as C

    import reference_caller as R
    from helpers import engine

    L = R.lib()
    bad = R.ConnectionView.from_buffer_copy(ref.conns[1])
    bad.pre_group_num_neurons += 1
    w = np.zeros((65, 16), np.float32)
    rc = L.bn_state_upload_connection_view(ref.h, C.byref(bad), w.ctypes.data_as(R.PF), None, None, None)
    assert rc != 0 and "shape" in engine().lib().bn_last_error().decode()
    neg = np.array(ref.weights[0], np.float32)
    neg[0, 1] = -1.0
    rc = L.bn_state_upload_connection_view(ref.h, C.byref(ref.conns[0]), neg.ctypes.data_as(R.PF), None, None, None)
    assert rc != 0 and "non-negative" in engine().lib().bn_last_error().decode()
    tab, ln = R.PU(), R.U()
    assert L.bn_state_poisson_table(ref.h, 2, 0, 0.1, C.byref(tab), C.byref(ln)) != 0
    ref.close()
