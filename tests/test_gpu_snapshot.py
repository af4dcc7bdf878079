"""State snapshot / resume (bn_sim_save_state, bn_sim_load_state): a simulation restored from a
snapshot continues exactly as the one it was taken from (every neuron field, connection and
plasticity trace and weight, bit for bit), across graph and eager steps, stimulus pulses and both
normalisation modes; a snapshot of another network is refused."""
import json
import os

import numpy as np
import pytest

from helpers import engine, small_config

pytestmark = pytest.mark.gpu


def _same(a, b):
    sa, sb = a.state(), b.state()
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key
    for k in range(4):
        assert np.array_equal(a.weights(k), b.weights(k)), f"weights {k}"
        for x, y in zip(a.connection_traces(k), b.connection_traces(k)):
            assert np.array_equal(x, y), f"connection traces {k}"
    for off in range(a.info.num_plasticities):
        for x, y in zip(a.plasticity_traces(off), b.plasticity_traces(off)):
            assert np.array_equal(x, y), f"plasticity traces {off}"


@pytest.mark.parametrize("lazy", [True, False], ids=["lazy", "faithful"])
def test_snapshot_resume_bit_exact(tmp_path, lazy):
    cfg = small_config(steps=900)
    cfg["stimulus_start_clock_steps"] = [20, 305]
    cfg["stimulus_end_clock_steps"] = [60, 330]
    path = os.path.join(str(tmp_path), "net.json")
    json.dump(cfg, open(path, "w"))
    flags = ("no_monitors", "quiet") + (("lazy_normalize",) if lazy else ())
    snap = os.path.join(str(tmp_path), "state.bin")
    a = engine().Simulation(path, flags=flags)
    a.run(300)
    a.save_state(snap)
    a.run(207)
    b = engine().Simulation(path, flags=flags)
    b.run(13)  # any state: the snapshot replaces it
    b.load_state(snap)
    assert b.get_info().clock_step == 300
    b.run(207)
    assert b.get_info().clock_step == a.get_info().clock_step == 507
    _same(a, b)
    # A different network is refused.
    other = small_config(n_e_side=12, n_i_side=6, steps=900)
    opath = os.path.join(str(tmp_path), "other.json")
    json.dump(other, open(opath, "w"))
    c = engine().Simulation(opath, flags=flags)
    with pytest.raises(engine().BnError):
        c.load_state(snap)
    for s in (a, b, c):
        s.close()
