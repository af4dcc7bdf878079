"""GPU: the allocation-free propagate (bn_connection_propagate_no_alloc) against
bn_connection_propagate, through the C ABI.

Two identical reference-side callers (tests/reference_caller.py) are stepped to the same state;
then each connection's propagate runs on one through the existing entry point (the caller's
gemv_partials) and on the other through the allocation-free one (the library's per-GPU scratch):
the post conductances and the connection traces must be bit-identical.  The calls run twice in
a row on the same GPU, the second over smaller connections (the scratch keeps the first call's
contents), and an oversized request is refused without launching anything.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import reference_caller as R
from helpers import engine, small_config

pytestmark = pytest.mark.gpu


def _bind():
    L = R.lib()
    L.bn_connection_propagate_no_alloc.argtypes = [R.ConnectionView, C.c_void_p]
    L.bn_connection_propagate_no_alloc.restype = C.c_int
    return L


def _conn_outputs(ref, k):
    c = ref.conns[k]
    g = ref.state()["gi" if c.pre_group_is_inhibtory else "ge"][c.post_group_offset:c.post_group_offset +
                                                                c.post_group_num_neurons]
    n = c.pre_group_num_neurons
    traces = [ref._download(p, n, np.float32) for p in (c.postsynaptic_r_traces, c.postsynaptic_d_traces,
                                                        c.summed_postsynaptic_conductances)]
    return g, traces


def _free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_no_alloc_propagate_matches_existing_entry_point(tmp_path):
    L = _bind()
    cfg = small_config(n_e_side=40, n_i_side=15, steps=400)  # 1,600 E + 225 I; E->E partials use the scratch
    cfg["stimulus_input_values"] = [3e-9, 3e-9]  # 1.5 mV per step: the pulsed E neurons spike from step ~25
    path = os.path.join(tmp_path, "cfg.json")
    json.dump(cfg, open(path, "w"))
    a, b = R.ReferenceSystem(path), R.ReferenceSystem(path)
    for _ in range(60):  # stimulus at step 20: spikes, nonzero traces
        a.step_the_clock()
        b.step_the_clock()
    assert np.array_equal(a.state()["V"], b.state()["V"])
    free0 = _free_device_bytes()
    # Two calls in a row on this GPU: the large E->E connection, then the smaller ones.
    for order in ([0], [3, 1, 2]):
        for k in order:
            ca, cb = a.conns[k], b.conns[k]
            R.check(L.bn_connection_propagate(ca, a.stream), "bn_connection_propagate")
            R.check(L.bn_connection_propagate_no_alloc(cb, b.stream), "bn_connection_propagate_no_alloc")
            ga, ta = _conn_outputs(a, k)
            gb, tb = _conn_outputs(b, k)
            assert np.array_equal(ga, gb), f"connection {k}: g differs"
            for x, y in zip(ta, tb):
                assert np.array_equal(x, y), f"connection {k}: traces differ"
            if k in (0, 1):  # E pre neurons have spiked: their connections carry current
                assert np.abs(ga).sum() > 0 and np.abs(ta[2]).sum() > 0
    assert _free_device_bytes() == free0
    # Again with twice the row block (half the partial sums, a smaller part of the scratch than
    # the first call used): still identical.
    ca = R.ConnectionView.from_buffer_copy(a.conns[0])
    cb = R.ConnectionView.from_buffer_copy(b.conns[0])
    ca.gemv_row_block = cb.gemv_row_block = 2 * a.conns[0].gemv_row_block
    R.check(L.bn_connection_propagate(ca, a.stream), "bn_connection_propagate")
    R.check(L.bn_connection_propagate_no_alloc(cb, b.stream), "bn_connection_propagate_no_alloc")
    assert np.array_equal(_conn_outputs(a, 0)[0], _conn_outputs(b, 0)[0])
    # Oversized: 1,600 chunks of one row x 1,600 columns > 1,048,576 partial sums.
    cb.gemv_row_block = 1
    v_before = b.state()
    rc = L.bn_connection_propagate_no_alloc(cb, b.stream)
    assert rc != 0 and "exceed" in engine().lib().bn_last_error().decode()
    assert np.array_equal(b.state()["ge"], v_before["ge"])  # nothing ran
    cb.gemv_row_block = 0
    assert L.bn_connection_propagate_no_alloc(cb, b.stream) != 0
    a.close()
    b.close()
