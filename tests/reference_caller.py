"""A reference-side caller of the drop-in boundary (test infrastructure).

Mirrors what the reference's host code does around its launchers, in Python over ctypes: one
device state object (bn_state_*, in place of GpuSimulationState), the components' views built
from the state's buffers as their make_gpu_view functions do (neuron_group.cpp:64-76,
connection.cpp:44-87, stimulus.cpp, plasticity.cpp / triplet_plasticity.cpp:14-85,
istdp_plasticity.cpp:22-56), and System::step_the_clock's launcher order (system.cpp:43-63):
evolve per group, propagate per connection, stimulus with the host cursor (stimulus.cpp:46-60),
plasticity updates, historicize per monitor, then the engine's clock advance.  The weight rows
are the reference's unpadded layout (weight_matrix_ld = 0 in every view).
"""
from __future__ import annotations

import ctypes as C
import importlib

import numpy as np

from helpers import engine, stimulus_schedule

F, U, B = C.c_float, C.c_uint32, C.c_bool
PF, PU, PD = C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.POINTER(C.c_double)
_libm = C.CDLL("libm.so.6")
_libm.expf.restype = C.c_float
_libm.expf.argtypes = [C.c_float]


def f32_decay(dt: float, tau: float) -> float:
    """std::exp(-1.0f * dt / tau) in binary32 (connection.cpp:46-47, main_buryn.cpp:295-298)."""
    return _libm.expf(float(np.float32(np.float32(-1.0) * np.float32(dt)) / np.float32(tau)))


class DeviceClock(C.Structure):
    _fields_ = [("clock_step", U), ("next_event_idx", U), ("rng_key", U * 2)]


class NeuronGroupView(C.Structure):
    _fields_ = [(n, F) for n in (
        "timestep", "g_leak", "membrane_capacitance", "depolarization_slope_factor", "v_effective_spike_threshold",
        "v_excitatory_reversal", "v_inhibitory_reversal", "v_threshold", "v_reset", "v_rest")] + [
        ("refractory_steps", U), ("is_inhibitory", B)] + [(n, F) for n in (
            "mean_background_excitatory_spikes", "mean_background_inhibitory_spikes",
            "background_excitatory_postsynaptic_r_trace_decay", "background_excitatory_postsynaptic_d_trace_decay",
            "background_inhibitory_postsynaptic_r_trace_decay", "background_inhibitory_postsynaptic_d_trace_decay",
            "background_excitatory_scaling_factor", "background_inhibitory_scaling_factor")] + [
        ("num_neurons", U), ("offset", U), ("membrane_potentials", PF), ("g_excitatory", PF), ("g_inhibitory", PF),
        ("i_stimulus", PF), ("refractory_counter", PU), ("spikes", PU)] + [(n, PF) for n in (
            "background_g_excitatory_synaptic_conductances", "background_g_inhibitory_synaptic_conductances",
            "background_excitatory_postsynaptic_r_traces", "background_excitatory_postsynaptic_d_traces",
            "background_inhibitory_postsynaptic_r_traces", "background_inhibitory_postsynaptic_d_traces")] + [
        ("rng_clock", C.POINTER(DeviceClock)), ("poisson_table_excitatory", PU), ("poisson_table_inhibitory", PU),
        ("poisson_table_excitatory_len", U), ("poisson_table_inhibitory_len", U), ("rng_instance", U)]


class ConnectionView(C.Structure):
    _fields_ = [("postsynaptic_r_trace_decay", F), ("postsynaptic_d_trace_decay", F),
                ("postsynaptic_kernel_normalization_value", F), ("pre_group_is_inhibtory", B)] + [
        (n, U) for n in ("weight_matrix_idx", "pre_group_offset", "post_group_offset", "pre_group_num_neurons",
                         "post_group_num_neurons")] + [
        (n, PF) for n in ("weight_matrix", "postsynaptic_r_traces", "postsynaptic_d_traces",
                          "summed_postsynaptic_conductances", "post_group_g_excitatory", "post_group_g_inhibitory")] + [
        ("spikes", PU), ("weight_matrix_ld", U), ("gemv_partials", PD), ("gemv_row_block", U)]


class StimulusView(C.Structure):
    _fields_ = [("g_excitatory", PF), ("g_inhibitory", PF), ("i_stimulus", PF), ("timestep_event_count", PU),
                ("stimulus_event_neuron_ids", PU), ("stimulus_event_post_group_offsets", PU),
                ("stimulus_event_amplitude_changes", PF), ("num_timesteps", U), ("num_events", U)]


class PlasticityView(C.Structure):
    _fields_ = [(n, F) for n in ("ltp_A", "ltd_A", "ltp_B", "ltd_B", "pre_trace_decay", "post_trace_decay",
                                 "triplet_pre_trace_decay", "triplet_post_trace_decay", "weight_min", "weight_max")] + [
        (n, U) for n in ("weight_matrix_idx", "plasticity_offset", "pre_group_offset", "post_group_offset",
                         "pre_group_num_neurons", "post_group_num_neurons")] + [
        (n, PF) for n in ("pre_traces", "post_traces", "triplet_pre_traces", "triplet_post_traces")] + [
        ("pre_spikes", PU), ("post_spikes", PU), ("weight_matrix", PF), ("pre_weight_sums", PF),
        ("weight_matrix_ld", U)]


class PlasticityParams(C.Structure):
    _fields_ = [(n, F) for n in ("ltp_A", "ltd_A", "ltp_B", "ltd_B", "pre_trace_decay", "post_trace_decay",
                                 "triplet_pre_trace_decay", "triplet_post_trace_decay", "weight_min", "weight_max")] + [
        ("plasticity_offset", U)]


class StateBuffers(C.Structure):
    _fields_ = [(n, U) for n in ("total_neurons", "num_excitatory_neurons", "num_inhibitory_neurons",
                                 "recording_interval_steps")] + [
        (n, PF) for n in ("membrane_potentials", "g_excitatory", "g_inhibitory", "i_stimulus")] + [
        ("refractory_counter", PU), ("spikes", PU)] + [
        (n, PF) for n in ("background_g_excitatory_synaptic_conductances",
                          "background_g_inhibitory_synaptic_conductances",
                          "background_excitatory_postsynaptic_r_traces", "background_excitatory_postsynaptic_d_traces",
                          "background_inhibitory_postsynaptic_r_traces", "background_inhibitory_postsynaptic_d_traces")] + [
        ("clock", C.POINTER(DeviceClock)), ("weight_matrix", PF * 4), ("weight_matrix_ld", U * 4),
        ("weight_pre", U * 4), ("weight_post", U * 4), ("postsynaptic_r_traces", PF * 4),
        ("postsynaptic_d_traces", PF * 4), ("summed_postsynaptic_conductances", PF * 4), ("gemv_partials", PD * 4),
        ("gemv_row_block", U * 4), ("timestep_event_count", PU), ("stimulus_event_neuron_ids", PU),
        ("stimulus_event_post_group_offsets", PU), ("stimulus_event_amplitude_changes", PF)] + [
        (n, PF) for n in ("pre_traces", "post_traces", "triplet_pre_traces", "triplet_post_traces",
                          "pre_weight_sums")] + [("spike_history", PU), ("membrane_potential_history", PF)]


BN_STATE_REFERENCE_LAYOUT = 1
BN_STATE_HISTORIES = 2

_bound = False


def lib():
    global _bound
    L = engine().lib()
    if _bound:
        return L
    P = C.c_void_p
    L.bn_state_create.argtypes = [U, U, U, U, U, C.c_uint64, U, C.POINTER(P)]
    L.bn_state_destroy.argtypes = [P]
    L.bn_state_destroy.restype = None
    L.bn_state_get_buffers.argtypes = [P, C.POINTER(StateBuffers)]
    L.bn_state_get_stream.argtypes = [P, C.POINTER(P)]
    L.bn_state_total_allocated_bytes.argtypes = [P, C.POINTER(C.c_uint64)]
    L.bn_state_poisson_table.argtypes = [P, U, U, F, C.POINTER(PU), PU]
    L.bn_state_upload_neuron_group_view.argtypes = [P, C.POINTER(NeuronGroupView), P]
    L.bn_state_upload_connection_view.argtypes = [P, C.POINTER(ConnectionView), PF, PF, PF, PF]
    L.bn_state_upload_stimulus_view.argtypes = [P, C.POINTER(StimulusView), PU, U, PU, PU, PF, U]
    L.bn_state_upload_plasticity_view.argtypes = [P, C.POINTER(PlasticityView), PF, PF, PF, PF, PF]
    L.bn_state_download_and_flush_spike_history.argtypes = [P, U, U, U, PU]
    L.bn_state_download_and_flush_membrane_potential_history.argtypes = [P, U, U, U, PF]
    L.bn_state_download_weight_matrix.argtypes = [P, U, PF]
    L.bn_state_download.argtypes = [P, P, P, C.c_size_t]
    L.bn_neuron_group_evolve.argtypes = [NeuronGroupView, P]
    L.bn_connection_propagate.argtypes = [ConnectionView, P]
    L.bn_stimulus_apply.argtypes = [StimulusView, U, U, P]
    for f in ("bn_triplet_plasticity_update", "bn_triplet_plasticity_update_normalize", "bn_istdp_plasticity_update",
              "bn_stdp_plasticity_update"):
        getattr(L, f).argtypes = [PlasticityView, PlasticityParams, P]
    L.bn_historicize_spikes.argtypes = [PU, PU, U, U, U, P]
    L.bn_historicize_membrane_potentials.argtypes = [PF, PF, U, U, U, P]
    L.bn_clock_init.argtypes = [C.POINTER(DeviceClock), C.c_uint64, P]
    L.bn_clock_advance.argtypes = [C.POINTER(DeviceClock), P]
    for f in ("bn_state_create", "bn_state_get_buffers", "bn_state_get_stream", "bn_state_total_allocated_bytes",
              "bn_state_poisson_table", "bn_state_upload_neuron_group_view", "bn_state_upload_connection_view",
              "bn_state_upload_stimulus_view", "bn_state_upload_plasticity_view",
              "bn_state_download_and_flush_spike_history", "bn_state_download_and_flush_membrane_potential_history",
              "bn_state_download_weight_matrix", "bn_state_download"):
        getattr(L, f).restype = C.c_int
    _bound = True
    return L


def _engine_module():
    return importlib.import_module(engine().__name__ + ".engine")


def check(rc: int, what: str = "") -> None:
    """cuda_check equivalent: raise with the engine's "what: message" text."""
    if rc != 0:
        raise engine().BnError(rc, f"{what}: {engine().lib().bn_last_error().decode()}")


def offset(ptr, n: int, ctype):
    """ptr + n elements (device addresses are plain integers to ctypes)."""
    return C.cast(C.c_void_p(C.cast(ptr, C.c_void_p).value + n * C.sizeof(ctype)), C.POINTER(ctype))


class ReferenceSystem:
    """The reference's main_buryn.cpp wiring of one network on the drop-in launchers."""

    def __init__(self, config_path: str, recording_interval: int = 10000):
        eng = engine()
        L = lib()
        self._L = L
        cfg = eng.config_dump(config_path)
        self.cfg = cfg
        ne, ni = cfg["num_excitatory_neurons"], cfg["num_inhibitory_neurons"]
        self.NE, self.NI, self.N = ne, ni, ne + ni
        self.T = int(cfg["simulation_steps"])
        counts, ids, offs, amps = stimulus_schedule(cfg)
        self.stim_counts = counts
        n_plast = int(cfg["e_e_plasticity_enabled"]) + int(cfg["i_e_plasticity_enabled"])
        h = C.c_void_p()
        check(L.bn_state_create(self.T, ne, ni, len(ids), n_plast, int(cfg["rng_seed"]) & 0xFFFFFFFFFFFFFFFF,
                                BN_STATE_REFERENCE_LAYOUT | BN_STATE_HISTORIES, C.byref(h)), "bn_state_create")
        self.h = h
        self.buf = StateBuffers()
        check(L.bn_state_get_buffers(h, C.byref(self.buf)), "bn_state_get_buffers")
        assert self.buf.recording_interval_steps == recording_interval
        s = C.c_void_p()
        check(L.bn_state_get_stream(h, C.byref(s)), "bn_state_get_stream")
        self.stream = s
        b = self.buf
        dt = cfg["timestep"]
        # Neuron groups (neuron_group.cpp:64-76), E at 0 and I at N_E (main_buryn.cpp:163,186).
        self.groups = []
        for g, (off, n) in enumerate(((0, ne), (ne, ni))):
            pre = "group_e_" if g == 0 else "group_i_"
            v = NeuronGroupView()
            for k in ("timestep", "g_leak", "membrane_capacitance", "depolarization_slope_factor",
                      "v_effective_spike_threshold", "v_excitatory_reversal", "v_inhibitory_reversal", "v_threshold",
                      "v_reset", "v_rest"):
                setattr(v, k, cfg[k])
            v.refractory_steps = cfg["refractory_steps"]
            v.is_inhibitory = g == 1
            for k in ("mean_background_excitatory_spikes", "mean_background_inhibitory_spikes",
                      "background_excitatory_postsynaptic_r_trace_decay",
                      "background_excitatory_postsynaptic_d_trace_decay",
                      "background_inhibitory_postsynaptic_r_trace_decay",
                      "background_inhibitory_postsynaptic_d_trace_decay", "background_excitatory_scaling_factor",
                      "background_inhibitory_scaling_factor"):
                setattr(v, k, cfg[pre + k])
            v.num_neurons, v.offset = n, off
            for k, t in (("membrane_potentials", F), ("g_excitatory", F), ("g_inhibitory", F), ("i_stimulus", F),
                         ("refractory_counter", U), ("spikes", U),
                         ("background_g_excitatory_synaptic_conductances", F),
                         ("background_g_inhibitory_synaptic_conductances", F),
                         ("background_excitatory_postsynaptic_r_traces", F),
                         ("background_excitatory_postsynaptic_d_traces", F),
                         ("background_inhibitory_postsynaptic_r_traces", F),
                         ("background_inhibitory_postsynaptic_d_traces", F)):
                setattr(v, k, offset(getattr(b, k), off, t))
            v.rng_clock = b.clock
            for kind, name in ((0, "excitatory"), (1, "inhibitory")):
                tab, ln = PU(), U()
                check(L.bn_state_poisson_table(h, g, kind, getattr(v, f"mean_background_{name}_spikes"), C.byref(tab),
                                               C.byref(ln)), "bn_state_poisson_table")
                setattr(v, f"poisson_table_{name}", tab)
                setattr(v, f"poisson_table_{name}_len", ln.value)
            v.rng_instance = 0
            init_v = np.full(n, cfg["v_rest"], np.float32)  # neuron_group.cpp:21
            host = _engine_module().NeuronStateHost(*([init_v.ctypes.data] + [None] * 11))
            check(L.bn_state_upload_neuron_group_view(h, C.byref(v), C.byref(host)), "upload_neuron_group_view")
            self.groups.append(v)
        # Connections EE, EI, IE, II (main_buryn.cpp:210-272, connection.cpp:44-87).
        self.weights = [eng.generate_weight_matrix(config_path, k) for k in range(4)]
        self.conns = []
        for k, (pg, qg) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
            exc = pg == 0
            c = ConnectionView()
            c.postsynaptic_r_trace_decay = f32_decay(dt, cfg["tau_rise_e" if exc else "tau_rise_i"])
            c.postsynaptic_d_trace_decay = f32_decay(dt, cfg["tau_decay_e" if exc else "tau_decay_i"])
            c.postsynaptic_kernel_normalization_value = cfg["postsynaptic_kernel_normalization_value"]
            c.pre_group_is_inhibtory = not exc
            c.weight_matrix_idx = k
            c.pre_group_offset, c.pre_group_num_neurons = self.groups[pg].offset, self.groups[pg].num_neurons
            c.post_group_offset, c.post_group_num_neurons = self.groups[qg].offset, self.groups[qg].num_neurons
            c.weight_matrix = b.weight_matrix[k]
            c.postsynaptic_r_traces = b.postsynaptic_r_traces[k]
            c.postsynaptic_d_traces = b.postsynaptic_d_traces[k]
            c.summed_postsynaptic_conductances = b.summed_postsynaptic_conductances[k]
            c.post_group_g_excitatory = offset(b.g_excitatory, c.post_group_offset, F)
            c.post_group_g_inhibitory = offset(b.g_inhibitory, c.post_group_offset, F)
            c.spikes = offset(b.spikes, c.pre_group_offset, U)
            c.weight_matrix_ld = 0  # the reference layout
            c.gemv_partials = b.gemv_partials[k]
            c.gemv_row_block = b.gemv_row_block[k]
            w = np.ascontiguousarray(self.weights[k], np.float32)
            check(L.bn_state_upload_connection_view(h, C.byref(c), w.ctypes.data_as(PF), None, None, None),
                  "upload_connection_view")
            self.conns.append(c)
        # Stimulus (stimulus.cpp), host cursor.
        st = StimulusView()
        st.g_excitatory, st.g_inhibitory, st.i_stimulus = b.g_excitatory, b.g_inhibitory, b.i_stimulus
        st.timestep_event_count = b.timestep_event_count
        st.stimulus_event_neuron_ids = b.stimulus_event_neuron_ids
        st.stimulus_event_post_group_offsets = b.stimulus_event_post_group_offsets
        st.stimulus_event_amplitude_changes = b.stimulus_event_amplitude_changes
        st.num_timesteps, st.num_events = len(counts), len(ids)
        self._stim_host = [np.ascontiguousarray(a) for a in (counts, ids, offs, amps)]
        c_, i_, o_, a_ = self._stim_host
        check(L.bn_state_upload_stimulus_view(h, C.byref(st), c_.ctypes.data_as(PU), len(c_), i_.ctypes.data_as(PU),
                                              o_.ctypes.data_as(PU), a_.ctypes.data_as(PF), len(i_)),
              "upload_stimulus_view")
        self.stim = st
        self.next_event_idx = 0
        self.current_timestep = 0
        # Plasticities (main_buryn.cpp:286-321; plasticity views as triplet_plasticity.cpp:35-85).
        self.plast = []
        off = 0
        for enabled, conn_idx, key, kind in ((cfg["e_e_plasticity_enabled"], 0, "e_e", "triplet"),
                                             (cfg["i_e_plasticity_enabled"], 2, "i_e", "istdp")):
            if not enabled:
                continue
            c = self.conns[conn_idx]
            p = PlasticityParams()
            p.ltp_A = cfg[f"{key}_plasticity_ltp_A"]
            p.ltd_A = cfg[f"{key}_plasticity_ltd_A"]
            p.ltp_B = cfg.get(f"{key}_plasticity_ltp_B", 0.0) if kind == "triplet" else 0.0
            p.ltd_B = cfg[f"{key}_plasticity_ltd_B"]
            p.weight_min = cfg[f"{key}_plasticity_weight_min"]
            p.weight_max = cfg[f"{key}_plasticity_weight_max"]
            p.pre_trace_decay = f32_decay(dt, cfg[f"{key}_plasticity_tau_ltp"])
            p.post_trace_decay = f32_decay(dt, cfg[f"{key}_plasticity_tau_ltd"])
            if kind == "triplet":
                p.triplet_pre_trace_decay = f32_decay(dt, cfg[f"{key}_plasticity_triplet_tau_ltp"])
                p.triplet_post_trace_decay = f32_decay(dt, cfg[f"{key}_plasticity_triplet_tau_ltd"])
            p.plasticity_offset = off
            v = PlasticityView()
            for f, _ in PlasticityParams._fields_[:10]:
                setattr(v, f, getattr(p, f))
            v.weight_matrix_idx, v.plasticity_offset = c.weight_matrix_idx, off
            v.pre_group_offset, v.post_group_offset = c.pre_group_offset, c.post_group_offset
            v.pre_group_num_neurons, v.post_group_num_neurons = c.pre_group_num_neurons, c.post_group_num_neurons
            base = off * self.N
            v.pre_traces = offset(b.pre_traces, base + v.pre_group_offset, F)
            v.post_traces = offset(b.post_traces, base + v.post_group_offset, F)
            v.triplet_pre_traces = offset(b.triplet_pre_traces, base + v.pre_group_offset, F)
            v.triplet_post_traces = offset(b.triplet_post_traces, base + v.post_group_offset, F)
            v.pre_spikes = offset(b.spikes, v.pre_group_offset, U)
            v.post_spikes = offset(b.spikes, v.post_group_offset, U)
            v.weight_matrix = c.weight_matrix
            v.pre_weight_sums = offset(b.pre_weight_sums, base + v.post_group_offset, F)
            v.weight_matrix_ld = 0
            wsums = None
            if kind == "triplet":
                # triplet_plasticity.cpp:22-30: initial column sums, FP32, sequential over pre rows.
                w = self.weights[conn_idx]
                wsums = np.add.accumulate(w, axis=0, dtype=np.float32)[-1].astype(np.float32)
            check(L.bn_state_upload_plasticity_view(h, C.byref(v), None, None, None, None,
                                                    None if wsums is None else wsums.ctypes.data_as(PF)),
                  "upload_plasticity_view")
            self.plast.append((kind, v, p))
            off += 1
        self.interval = recording_interval
        self.clock_step = 0

    def close(self):
        if getattr(self, "h", None):
            self._L.bn_state_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step_the_clock(self):
        """system.cpp:43-63 through the drop-in launchers, then the device clock advance."""
        L, s = lib(), self.stream
        for v in self.groups:
            check(L.bn_neuron_group_evolve(v, s), "kernel_neuron_group_evolve")
        for c in self.conns:
            check(L.bn_connection_propagate(c, s), "kernel_connection_propagate")
        # stimulus.cpp:46-60 (host cursor)
        if self.current_timestep < len(self.stim_counts):
            n = int(self.stim_counts[self.current_timestep])
            if n:
                check(L.bn_stimulus_apply(self.stim, self.next_event_idx, n, s), "kernel_stimulus_apply")
            self.next_event_idx += n
            self.current_timestep += 1
        for kind, v, p in self.plast:
            f = L.bn_triplet_plasticity_update_normalize if kind == "triplet" else L.bn_istdp_plasticity_update
            check(f(v, p, s), f"kernel_{kind}_plasticity_update")
        # SpikeRecorder (E, I) and VoltageRecorder (E) historicize (spike_recorder.cpp, voltage_recorder.cpp)
        b = self.buf
        for g in self.groups:
            check(L.bn_historicize_spikes(offset(b.spikes, g.offset, U),
                                          offset(b.spike_history, g.offset * self.interval, U), g.num_neurons,
                                          self.interval, self.clock_step, s), "kernel_historicize_spikes")
        check(L.bn_historicize_membrane_potentials(b.membrane_potentials, b.membrane_potential_history, self.NE,
                                                   self.interval, self.clock_step, s),
              "kernel_historicize_membrane_potentials")
        check(L.bn_clock_advance(b.clock, s), "clock advance")
        self.clock_step += 1

    def _download(self, ptr, n, dtype):
        out = np.zeros(n, dtype)
        check(lib().bn_state_download(self.h, C.cast(ptr, C.c_void_p), out.ctypes.data, out.nbytes), "download")
        return out

    def state(self) -> dict:
        b = self.buf
        return {"V": self._download(b.membrane_potentials, self.N, np.float32),
                "ge": self._download(b.g_excitatory, self.N, np.float32),
                "gi": self._download(b.g_inhibitory, self.N, np.float32),
                "istim": self._download(b.i_stimulus, self.N, np.float32),
                "refr": self._download(b.refractory_counter, self.N, np.uint32),
                "spikes": self._download(b.spikes, self.N, np.uint32)}

    def weights_now(self, idx: int) -> np.ndarray:
        k = self.conns[idx]
        out = np.zeros((k.pre_group_num_neurons, k.post_group_num_neurons), np.float32)
        check(lib().bn_state_download_weight_matrix(self.h, idx, out.ctypes.data_as(PF)), "download_weight_matrix")
        return out

    def spike_history(self, group: int) -> np.ndarray:
        """[num_neurons][interval] history of the group, slot = step % interval, zeroed after the
        download (download_and_flush_spike_history with the recorder's interval)."""
        g = self.groups[group]
        out = np.zeros((g.num_neurons, self.interval), np.uint32)
        check(lib().bn_state_download_and_flush_spike_history(self.h, g.offset, g.num_neurons, self.interval,
                                                              out.ctypes.data_as(PU)), "download spike history")
        return out

    def voltage_history(self) -> np.ndarray:
        """[N_E][interval] voltage history of the E group (download_and_flush_membrane_potential_history)."""
        out = np.zeros((self.NE, self.interval), np.float32)
        check(lib().bn_state_download_and_flush_membrane_potential_history(self.h, 0, self.NE, self.interval,
                                                                           out.ctypes.data_as(PF)),
              "download voltage history")
        return out
