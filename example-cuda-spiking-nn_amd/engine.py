"""ctypes view of the engine's C ABI (include/buryn_amd.h).

``Simulation`` mirrors what the reference's driver does with a config
(cpp_src/main_buryn.cpp:98-343) and exposes the step loop (System::run,
cpp_src/core/system.cpp:65-83) plus state downloads for parity checks.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libburyn_amd.so")

BURYN_FLAGS = {
    "write_files": 1 << 0,
    "no_monitors": 1 << 1,
    "lazy_normalize": 1 << 2,
    "component_path": 1 << 3,
    "no_graph": 1 << 4,
    "quiet": 1 << 5,
    "sequential_graph": 1 << 6,
    "reference_layout": 1 << 7,
    "persistent": 1 << 8,
    "gemv_part_sums": 1 << 9,
    "resident": 1 << 10,
}


class BnError(RuntimeError):
    """An engine call failed; carries the C-ABI code and the engine's message."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class SimInfo(C.Structure):
    _fields_ = [
        ("num_excitatory_neurons", C.c_uint32), ("num_inhibitory_neurons", C.c_uint32),
        ("total_neurons", C.c_uint32), ("clock_step", C.c_uint32), ("simulation_steps", C.c_uint32),
        ("num_plasticities", C.c_uint32), ("weight_matrix_ld", C.c_uint32 * 4),
        ("gemv_row_block", C.c_uint32 * 4), ("weight_pre", C.c_uint32 * 4), ("weight_post", C.c_uint32 * 4),
        ("flags", C.c_uint32), ("kernels_per_step", C.c_uint32), ("device_bytes", C.c_uint64),
        ("steps_per_graph", C.c_uint32), ("graph_launches", C.c_uint32),
        ("persistent_blocks", C.c_uint32), ("persistent_items", C.c_uint32), ("gemv_part_chunks", C.c_uint32),
        ("persistent_resident_rows", C.c_uint32), ("gemv_part_sums", C.c_uint32),
        ("resident_blocks", C.c_uint32), ("resident_columns", C.c_uint32), ("resident_e_waves", C.c_uint32),
        ("resident_rows", C.c_uint32),
    ]


_STATE_FIELDS = [
    ("membrane_potentials", np.float32), ("g_excitatory", np.float32), ("g_inhibitory", np.float32),
    ("i_stimulus", np.float32), ("refractory_counter", np.uint32), ("spikes", np.uint32),
    ("background_g_excitatory_synaptic_conductances", np.float32),
    ("background_g_inhibitory_synaptic_conductances", np.float32),
    ("background_excitatory_postsynaptic_r_traces", np.float32),
    ("background_excitatory_postsynaptic_d_traces", np.float32),
    ("background_inhibitory_postsynaptic_r_traces", np.float32),
    ("background_inhibitory_postsynaptic_d_traces", np.float32),
]


class NeuronStateHost(C.Structure):
    _fields_ = [(n, C.c_void_p) for n, _ in _STATE_FIELDS]


class BatchInstance(C.Structure):
    _fields_ = [("noise_seed", C.c_uint64), ("stimulus_amplitude", C.c_float), ("stimulus_start", C.c_uint32),
                ("stimulus_end", C.c_uint32), ("num_stimulus_ids", C.c_uint32),
                ("stimulus_ids", C.POINTER(C.c_uint32))]


class MultiInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("num_networks", "clock_step", "evolve_blocks", "propagate_blocks",
                                          "plasticity_blocks", "nontemporal", "kernels_per_step", "num_groups",
                                          "steps_per_graph")] + [("group_weight_bytes", C.c_uint64),
                                                                 ("num_streams", C.c_uint32),
                                                                 ("overlapped", C.c_uint32)]


class BatchInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "num_instances", "padded_instances", "num_excitatory_neurons", "num_inhibitory_neurons", "total_neurons",
        "clock_step", "gemm_rows_per_wave", "gemm_blocks")] + [
        ("gemm_groups", C.c_uint32 * 4), ("weight_pre", C.c_uint32 * 4), ("weight_post", C.c_uint32 * 4),
        ("kernels_per_step", C.c_uint32)]


class WindowSpec(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("checkpoint_interval", C.c_uint32), ("n_vrec", C.c_uint32),
                ("vrec_ids", C.POINTER(C.c_uint32)), ("rate_tau", C.c_float), ("surrogate_beta", C.c_float)]


class WindowInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("steps", "checkpoint_interval", "start_step", "total_neurons",
                                          "num_excitatory_neurons", "num_inhibitory_neurons", "n_vrec")] + [
        ("g_leak", C.c_float), ("tau_decay_e", C.c_float), ("forward_ms", C.c_float), ("backward_ms", C.c_float)]


EXPORTED_SYMBOLS = [
    "bn_last_error", "bn_poisson_table", "bn_neuron_group_evolve", "bn_connection_propagate",
    "bn_stimulus_apply", "bn_stimulus_apply_device_cursor", "bn_triplet_plasticity_update",
    "bn_triplet_plasticity_update_normalize", "bn_istdp_plasticity_update", "bn_stdp_plasticity_update",
    "bn_historicize_spikes", "bn_historicize_membrane_potentials", "bn_clock_init", "bn_clock_advance", "bn_sim_create",
    "bn_sim_create_from_json_text", "bn_sim_destroy", "bn_sim_run", "bn_sim_step_components",
    "bn_sim_synchronize", "bn_sim_get_info", "bn_sim_get_neuron_state", "bn_sim_get_weights",
    "bn_sim_get_connection_traces", "bn_sim_get_plasticity_traces", "bn_sim_run_timed",
    "bn_sim_time_propagate_kernel", "bn_sim_set_steps_per_graph", "bn_config_dump_json", "bn_generate_weight_matrix", "bn_set_connectivity_cache",
    "bn_batch_create", "bn_batch_destroy", "bn_batch_run", "bn_batch_synchronize", "bn_batch_run_timed",
    "bn_batch_time_gemm", "bn_batch_set_gemm_rows", "bn_batch_get_info", "bn_batch_get_state",
    "bn_batch_workload_instance", "bn_multi_create", "bn_multi_destroy", "bn_multi_run", "bn_multi_run_timed",
    "bn_multi_synchronize", "bn_multi_network", "bn_multi_get_info", "bn_multi_time_propagate",
    "bn_multi_set_schedule", "bn_multi_set_streams", "bn_window_create", "bn_window_destroy", "bn_window_set_params",
    "bn_window_get_weights", "bn_window_set_targets", "bn_window_forward", "bn_window_backward",
    "bn_window_get_info", "bn_window_get_gradient", "bn_measure_read_bandwidth", "bn_sim_save_state",
    "bn_sim_load_state", "bn_sim_get_spike_raster", "bn_measure_read_bandwidth_one_launch",
    "bn_sim_set_persistent_rows", "bn_sim_set_persistent_wait_limit", "bn_sim_set_persistent_resident",
    "bn_sim_set_resident_cooperative", "bn_sim_set_async_recording", "bn_sim_record_times",
]

_lib = None


def build_library() -> None:
    """Compiles the engine for gfx950 in-tree (csrc/Makefile)."""
    subprocess.run(["make", "-C", os.path.join(PKG, "csrc"), "-j8"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"engine library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, U, I = C.c_void_p, C.c_uint32, C.c_int
    fp = C.POINTER(C.c_float)
    up = C.POINTER(C.c_uint32)
    L.bn_last_error.restype = C.c_char_p
    L.bn_poisson_table.argtypes = [C.c_float, up, U, up]
    L.bn_sim_create.argtypes = [C.c_char_p, U, C.POINTER(P)]
    L.bn_sim_create_from_json_text.argtypes = [C.c_char_p, U, C.POINTER(P)]
    L.bn_sim_destroy.argtypes = [P]
    L.bn_sim_destroy.restype = None
    for f in ("bn_sim_run", "bn_sim_step_components"):
        getattr(L, f).argtypes = [P, U]
    L.bn_sim_synchronize.argtypes = [P]
    L.bn_sim_get_info.argtypes = [P, C.POINTER(SimInfo)]
    L.bn_sim_get_neuron_state.argtypes = [P, C.POINTER(NeuronStateHost)]
    L.bn_sim_get_weights.argtypes = [P, U, fp]
    L.bn_sim_get_connection_traces.argtypes = [P, U, fp, fp, fp]
    L.bn_sim_get_plasticity_traces.argtypes = [P, U, fp, fp, fp, fp]
    L.bn_sim_run_timed.argtypes = [P, U, fp]
    L.bn_sim_get_spike_raster.argtypes = [P, U, U, up]
    L.bn_sim_set_persistent_rows.argtypes = [P, up]
    L.bn_sim_set_persistent_wait_limit.argtypes = [P, U]
    L.bn_sim_set_persistent_resident.argtypes = [P, I]
    L.bn_sim_set_resident_cooperative.argtypes = [P, I]
    L.bn_sim_set_async_recording.argtypes = [P, I]
    L.bn_sim_record_times.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
    L.bn_sim_time_propagate_kernel.argtypes = [P, U, fp, fp]
    L.bn_sim_set_steps_per_graph.argtypes = [P, U]
    L.bn_measure_read_bandwidth.argtypes = [C.c_uint64, U, fp]
    L.bn_measure_read_bandwidth_one_launch.argtypes = [C.c_uint64, U, fp]
    L.bn_sim_save_state.argtypes = [P, C.c_char_p]
    L.bn_sim_load_state.argtypes = [P, C.c_char_p]
    L.bn_config_dump_json.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.bn_generate_weight_matrix.argtypes = [C.c_char_p, U, fp, up, up]
    L.bn_set_connectivity_cache.argtypes = [C.c_char_p]
    L.bn_batch_create.argtypes = [C.c_char_p, U, C.POINTER(BatchInstance), C.POINTER(P)]
    L.bn_batch_destroy.argtypes = [P]
    L.bn_batch_destroy.restype = None
    L.bn_batch_run.argtypes = [P, U]
    L.bn_batch_synchronize.argtypes = [P]
    L.bn_batch_run_timed.argtypes = [P, U, fp]
    L.bn_batch_time_gemm.argtypes = [P, U, fp]
    L.bn_batch_set_gemm_rows.argtypes = [P, U]
    L.bn_batch_get_info.argtypes = [P, C.POINTER(BatchInfo)]
    L.bn_batch_get_state.argtypes = [P, U, C.POINTER(NeuronStateHost)]
    L.bn_batch_workload_instance.argtypes = [U, U, U, C.c_float, up, fp]
    L.bn_multi_create.argtypes = [C.POINTER(C.c_char_p), U, U, C.POINTER(P)]
    L.bn_multi_destroy.argtypes = [P]
    L.bn_multi_destroy.restype = None
    L.bn_multi_run.argtypes = [P, U]
    L.bn_multi_run_timed.argtypes = [P, U, fp]
    L.bn_multi_synchronize.argtypes = [P]
    L.bn_multi_network.argtypes = [P, U, C.POINTER(P)]
    L.bn_multi_get_info.argtypes = [P, C.POINTER(MultiInfo)]
    L.bn_multi_time_propagate.argtypes = [P, U, fp]
    L.bn_multi_set_schedule.argtypes = [P, C.c_uint64, U]
    L.bn_multi_set_streams.argtypes = [P, U, U, U]
    dp = C.POINTER(C.c_double)
    L.bn_window_create.argtypes = [P, C.POINTER(WindowSpec), C.POINTER(P)]
    L.bn_window_destroy.argtypes = [P]
    L.bn_window_set_params.argtypes = [P, C.POINTER(P), C.c_float, C.c_float]
    L.bn_window_get_weights.argtypes = [P, U, fp]
    L.bn_window_set_targets.argtypes = [P, dp, dp]
    L.bn_window_forward.argtypes = [P, dp, dp, fp, C.POINTER(C.c_uint8)]
    L.bn_window_backward.argtypes = [P, C.POINTER(P), dp, dp]
    L.bn_window_get_info.argtypes = [P, C.POINTER(WindowInfo)]
    L.bn_window_get_gradient.argtypes = [P, U, fp]
    for f in EXPORTED_SYMBOLS:
        if f not in ("bn_last_error", "bn_sim_destroy", "bn_batch_destroy", "bn_multi_destroy",
                     "bn_window_destroy"):
            getattr(L, f).restype = I
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise BnError(rc, lib().bn_last_error().decode())


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


def poisson_table(mean: float, capacity: int = 256) -> np.ndarray:
    t = np.zeros(capacity, np.uint32)
    n = C.c_uint32(0)
    _check(lib().bn_poisson_table(mean, _up(t), capacity, C.byref(n)))
    return t[: n.value]


def measure_read_bandwidth(nbytes: int, repeats: int = 50) -> float:
    """GB/s of a plain streaming read over an `nbytes` device buffer, warm (see the C ABI)."""
    gbs = C.c_float(0)
    _check(lib().bn_measure_read_bandwidth(nbytes, repeats, C.byref(gbs)))
    return gbs.value


def measure_read_bandwidth_one_launch(nbytes: int, passes: int = 100) -> float:
    """GB/s of `passes` streaming passes over an `nbytes` device buffer inside one launch."""
    gbs = C.c_float(0)
    _check(lib().bn_measure_read_bandwidth_one_launch(nbytes, passes, C.byref(gbs)))
    return gbs.value


def config_dump(path: str) -> dict:
    """The engine's parsed config (defaults, overrides and path prefixing applied)."""
    need = C.c_size_t(0)
    _check(lib().bn_config_dump_json(path.encode(), None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value)
    _check(lib().bn_config_dump_json(path.encode(), buf, need.value, C.byref(need)))
    return json.loads(buf.value.decode())


def set_connectivity_cache(directory: str | None) -> None:
    """Directory of the binary connectivity cache (None or "": off); process-wide."""
    _check(lib().bn_set_connectivity_cache((directory or "").encode()))


def generate_weight_matrix(path: str, idx: int) -> np.ndarray:
    pre, post = C.c_uint32(0), C.c_uint32(0)
    _check(lib().bn_generate_weight_matrix(path.encode(), idx, None, C.byref(pre), C.byref(post)))
    out = np.zeros((pre.value, post.value), np.float32)
    _check(lib().bn_generate_weight_matrix(path.encode(), idx, _fp(out), C.byref(pre), C.byref(post)))
    return out


class Simulation:
    """One network on the GPU.  ``flags``: names from BURYN_FLAGS."""

    def __init__(self, config: str | dict, flags=("no_monitors", "quiet")):
        self.flag_bits = 0
        for f in flags:
            self.flag_bits |= BURYN_FLAGS[f]
        h = C.c_void_p()
        if isinstance(config, dict):
            _check(lib().bn_sim_create_from_json_text(json.dumps(config).encode(), self.flag_bits, C.byref(h)))
        else:
            _check(lib().bn_sim_create(config.encode(), self.flag_bits, C.byref(h)))
        self.h = h
        self.info = self.get_info()
        self.N = self.info.total_neurons
        self.NE = self.info.num_excitatory_neurons
        self.NI = self.info.num_inhibitory_neurons

    @classmethod
    def _borrowed(cls, handle, owner):
        """A view of a bn_sim owned by another object (a network of a MultiSimulation)."""
        self = cls.__new__(cls)
        self.flag_bits = 0
        self.h = handle
        self._owner = owner
        self.info = self.get_info()
        self.N = self.info.total_neurons
        self.NE = self.info.num_excitatory_neurons
        self.NI = self.info.num_inhibitory_neurons
        return self

    def close(self):
        if getattr(self, "_owner", None) is not None:
            self.h = None
            return
        if getattr(self, "h", None):
            lib().bn_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def get_info(self) -> SimInfo:
        info = SimInfo()
        _check(lib().bn_sim_get_info(self.h, C.byref(info)))
        return info

    def run(self, n: int) -> None:
        _check(lib().bn_sim_run(self.h, n))

    def step_components(self, n: int) -> None:
        _check(lib().bn_sim_step_components(self.h, n))

    def synchronize(self) -> None:
        _check(lib().bn_sim_synchronize(self.h))

    def run_timed(self, n: int) -> float:
        ms = C.c_float(0)
        _check(lib().bn_sim_run_timed(self.h, n, C.byref(ms)))
        return ms.value

    def set_steps_per_graph(self, n: int) -> None:
        """Steps per captured graph of the fused step (default 100; even: overlapped form)."""
        _check(lib().bn_sim_set_steps_per_graph(self.h, n))

    def set_persistent_rows(self, rows) -> None:
        """Rows per wave chunk of the persistent step per connection (int or 4 ints; 0 = chosen)."""
        r = np.array(rows if hasattr(rows, "__len__") else [rows] * 4, np.uint32)
        _check(lib().bn_sim_set_persistent_rows(self.h, _up(r)))
        self.info = self.get_info()

    def set_persistent_resident(self, allow: bool) -> None:
        """Allow or forbid register-resident weights in the persistent step (re-plans)."""
        _check(lib().bn_sim_set_persistent_resident(self.h, 1 if allow else 0))
        self.info = self.get_info()

    def set_resident_cooperative(self, cooperative: bool) -> None:
        """Cooperative (default) or ordinary launches of the resident step (same kernel)."""
        _check(lib().bn_sim_set_resident_cooperative(self.h, 1 if cooperative else 0))

    def set_async_recording(self, on: bool) -> None:
        """Recordings written by a writer thread while the steps run (default) or at each record step."""
        _check(lib().bn_sim_set_async_recording(self.h, 1 if on else 0))

    def record_times(self):
        """(seconds on the calling thread at record steps, seconds in the writer thread, record steps)."""
        a, b, n = C.c_double(0), C.c_double(0), C.c_uint32(0)
        _check(lib().bn_sim_record_times(self.h, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def set_persistent_wait_limit(self, polls: int) -> None:
        _check(lib().bn_sim_set_persistent_wait_limit(self.h, polls))

    def time_propagate_kernel(self, n: int):
        k, s = C.c_float(0), C.c_float(0)
        _check(lib().bn_sim_time_propagate_kernel(self.h, n, C.byref(k), C.byref(s)))
        return k.value, s.value

    def save_state(self, path: str) -> None:
        """Every device buffer of the simulation to `path` (see bn_sim_save_state)."""
        _check(lib().bn_sim_save_state(self.h, path.encode()))

    def load_state(self, path: str) -> None:
        """Resume from a snapshot taken from a simulation of the same config and flags."""
        _check(lib().bn_sim_load_state(self.h, path.encode()))
        self.info = self.get_info()

    def state(self) -> dict:
        arrays = {n: np.zeros(self.N, dt) for n, dt in _STATE_FIELDS}
        st = NeuronStateHost(*[arrays[n].ctypes.data for n, _ in _STATE_FIELDS])
        _check(lib().bn_sim_get_neuron_state(self.h, C.byref(st)))
        return arrays

    def weights(self, idx: int) -> np.ndarray:
        out = np.zeros((self.info.weight_pre[idx], self.info.weight_post[idx]), np.float32)
        _check(lib().bn_sim_get_weights(self.h, idx, _fp(out)))
        return out

    def spike_raster(self, first_step: int, num_steps: int) -> np.ndarray:
        """[num_steps][total_neurons] 0/1 spikes of the given steps from the device spike ring
        (monitors on; see bn_sim_get_spike_raster)."""
        words = (self.N + 31) // 32
        bits = np.zeros((num_steps, words), np.uint32)
        _check(lib().bn_sim_get_spike_raster(self.h, first_step, num_steps, _up(bits)))
        flat = np.unpackbits(bits.view(np.uint8).reshape(num_steps, words * 4), axis=1, bitorder="little")
        return flat[:, : self.N].astype(np.uint32)

    def connection_traces(self, idx: int):
        n = self.info.weight_pre[idx]
        r, d, x = (np.zeros(n, np.float32) for _ in range(3))
        _check(lib().bn_sim_get_connection_traces(self.h, idx, _fp(r), _fp(d), _fp(x)))
        return r, d, x

    def plasticity_traces(self, off: int):
        a = [np.zeros(self.N, np.float32) for _ in range(4)]
        _check(lib().bn_sim_get_plasticity_traces(self.h, off, *[_fp(x) for x in a]))
        return a


def batch_workload_instance(instance: int, num_excitatory: int, num_ids: int = 100,
                            max_amplitude: float = 200e-12):
    """The recorded stimulus draw of one instance of the shared-weight workload: (ids, amplitude)."""
    ids = np.zeros(num_ids, np.uint32)
    amp = C.c_float(0)
    _check(lib().bn_batch_workload_instance(instance, num_excitatory, num_ids, max_amplitude, _up(ids), C.byref(amp)))
    return ids, amp.value


class BatchSimulation:
    """B instances of one plasticity-free network sharing its weights (bn_batch_* API).

    instances: list of dicts with keys noise_seed, and optionally stimulus_ids (excitatory-local),
    stimulus_amplitude, stimulus_start, stimulus_end.
    """

    def __init__(self, config: str, instances):
        self._keep = []
        arr = (BatchInstance * len(instances))()
        for k, inst in enumerate(instances):
            ids = np.ascontiguousarray(inst.get("stimulus_ids", []), np.uint32)
            self._keep.append(ids)
            arr[k].noise_seed = int(inst["noise_seed"])
            arr[k].stimulus_amplitude = float(inst.get("stimulus_amplitude", 0.0))
            arr[k].stimulus_start = int(inst.get("stimulus_start", 0))
            arr[k].stimulus_end = int(inst.get("stimulus_end", 0))
            arr[k].num_stimulus_ids = ids.size
            arr[k].stimulus_ids = _up(ids) if ids.size else None
        h = C.c_void_p()
        _check(lib().bn_batch_create(config.encode(), len(instances), arr, C.byref(h)))
        self.h = h
        self.info = self.get_info()
        self.N = self.info.total_neurons

    def close(self):
        if getattr(self, "h", None):
            lib().bn_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_info(self) -> BatchInfo:
        info = BatchInfo()
        _check(lib().bn_batch_get_info(self.h, C.byref(info)))
        return info

    def run(self, n: int) -> None:
        _check(lib().bn_batch_run(self.h, n))

    def synchronize(self) -> None:
        _check(lib().bn_batch_synchronize(self.h))

    def run_timed(self, n: int) -> float:
        ms = C.c_float(0)
        _check(lib().bn_batch_run_timed(self.h, n, C.byref(ms)))
        return ms.value

    def time_gemm(self, n: int) -> float:
        ms = C.c_float(0)
        _check(lib().bn_batch_time_gemm(self.h, n, C.byref(ms)))
        return ms.value

    def set_gemm_rows(self, rows: int) -> None:
        _check(lib().bn_batch_set_gemm_rows(self.h, rows))
        self.info = self.get_info()

    def state(self, instance: int) -> dict:
        arrays = {n: np.zeros(self.N, dt) for n, dt in _STATE_FIELDS}
        st = NeuronStateHost(*[arrays[n].ctypes.data for n, _ in _STATE_FIELDS])
        _check(lib().bn_batch_get_state(self.h, instance, C.byref(st)))
        return arrays


class MultiSimulation:
    """Several plastic networks advanced together (bn_multi_* API); networks[k] is the
    network's own Simulation view (state, weights, traces)."""

    def __init__(self, configs, flags=("lazy_normalize",)):
        bits = 0
        for f in flags:
            bits |= BURYN_FLAGS[f]
        paths = (C.c_char_p * len(configs))(*[c.encode() for c in configs])
        h = C.c_void_p()
        _check(lib().bn_multi_create(paths, len(configs), bits, C.byref(h)))
        self.h = h
        self.networks = []
        for k in range(len(configs)):
            sh = C.c_void_p()
            _check(lib().bn_multi_network(self.h, k, C.byref(sh)))
            self.networks.append(Simulation._borrowed(sh, self))

    def close(self):
        if getattr(self, "h", None):
            for s in self.networks:
                s.close()
            lib().bn_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_info(self) -> MultiInfo:
        info = MultiInfo()
        _check(lib().bn_multi_get_info(self.h, C.byref(info)))
        return info

    def run(self, n: int) -> None:
        _check(lib().bn_multi_run(self.h, n))

    def synchronize(self) -> None:
        _check(lib().bn_multi_synchronize(self.h))

    def run_timed(self, n: int) -> float:
        ms = C.c_float(0)
        _check(lib().bn_multi_run_timed(self.h, n, C.byref(ms)))
        return ms.value

    def time_propagate(self, n: int) -> float:
        ms = C.c_float(0)
        _check(lib().bn_multi_time_propagate(self.h, n, C.byref(ms)))
        return ms.value

    def set_schedule(self, group_weight_bytes: int, steps_per_graph: int = 10) -> None:
        _check(lib().bn_multi_set_schedule(self.h, group_weight_bytes, steps_per_graph))

    def set_streams(self, streams: int, steps_per_graph: int = 10, overlap: bool = False) -> None:
        _check(lib().bn_multi_set_streams(self.h, streams, steps_per_graph, 1 if overlap else 0))

    def network(self, k: int) -> "Simulation":
        sh = C.c_void_p()
        _check(lib().bn_multi_network(self.h, k, C.byref(sh)))  # synchronises the multi stream
        return self.networks[k]


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ptr(a):
    """Address of a host numpy array or a device tensor (anything with data_ptr())."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


class Window:
    """Differentiable window of a plasticity-free network (bn_window_* API): ``steps`` steps
    from the current state of ``sim``, the rate + voltage loss against targets, and its gradient
    with respect to the four weight matrices, g_leak and tau_decay_e (hand-written adjoint with
    checkpoint recompute)."""

    def __init__(self, sim: Simulation, steps: int, vrec_ids, checkpoint_interval: int = 100,
                 rate_tau: float = 5e-3, surrogate_beta: float = 1e-3):
        self._ids = np.ascontiguousarray(vrec_ids, dtype=np.uint32)
        spec = WindowSpec(steps, checkpoint_interval, self._ids.size, _up(self._ids), rate_tau, surrogate_beta)
        h = C.c_void_p()
        _check(lib().bn_window_create(sim.h, C.byref(spec), C.byref(h)))
        self.h = h
        self.info = self.get_info()
        self.T = self.info.steps
        self.N = self.info.total_neurons
        self.NE = self.info.num_excitatory_neurons
        self.NI = self.info.num_inhibitory_neurons
        self.nv = self.info.n_vrec
        self.shapes = [(self.NE, self.NE), (self.NE, self.NI), (self.NI, self.NE), (self.NI, self.NI)]

    def close(self):
        if getattr(self, "h", None):
            lib().bn_window_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_info(self) -> WindowInfo:
        info = WindowInfo()
        _check(lib().bn_window_get_info(self.h, C.byref(info)))
        return info

    def set_params(self, weights=None, g_leak: float | None = None, tau_decay_e: float | None = None) -> None:
        """weights: None or 4 entries (None = keep), each a float32 [pre][post] numpy array or a
        contiguous float32 device tensor."""
        info = self.get_info()
        gl = info.g_leak if g_leak is None else g_leak
        td = info.tau_decay_e if tau_decay_e is None else tau_decay_e
        ptrs = None
        if weights is not None:
            keep = []
            for k, w in enumerate(weights):
                if w is not None and not hasattr(w, "data_ptr"):
                    w = np.ascontiguousarray(w, dtype=np.float32)
                    assert w.shape == self.shapes[k], (w.shape, self.shapes[k])
                keep.append(w)
            self._keep = keep
            ptrs = (C.c_void_p * 4)(*[_ptr(w) for w in keep])
        _check(lib().bn_window_set_params(self.h, ptrs, gl, td))

    def weights(self, idx: int) -> np.ndarray:
        out = np.zeros(self.shapes[idx], np.float32)
        _check(lib().bn_window_get_weights(self.h, idx, _fp(out)))
        return out

    def set_targets(self, target_rates=None, target_v=None) -> None:
        tr = None if target_rates is None else np.ascontiguousarray(target_rates, dtype=np.float64)
        tv = None if target_v is None else np.ascontiguousarray(target_v, dtype=np.float64)
        if tr is not None:
            assert tr.shape == (self.T, 2), tr.shape
        if tv is not None:
            assert tv.shape == (self.T, self.nv), tv.shape
        _check(lib().bn_window_set_targets(self.h, _dp(tr), _dp(tv)))

    def forward(self, outputs: bool = True) -> dict:
        loss = C.c_double(0)
        rates = np.zeros((self.T, 2), np.float64) if outputs else None
        v = np.zeros((self.T, self.nv), np.float32) if outputs else None
        raster = np.zeros((self.T, self.N), np.uint8) if outputs else None
        _check(lib().bn_window_forward(self.h, C.byref(loss), _dp(rates), _fp(v),
                                       None if raster is None else raster.ctypes.data_as(C.POINTER(C.c_uint8))))
        return {"loss": loss.value, "rates": rates, "v": v, "raster": raster}

    def backward(self, dW=None, download: bool = True) -> dict:
        """Gradient of the last forward's loss.  dW: None or 4 contiguous float32 device tensors
        to fill.  With dW None: host arrays (download) or nothing (the gradient stays on the
        device; gradient(k) fetches it)."""
        ptrs = None if dW is None else (C.c_void_p * 4)(*[_ptr(w) for w in dW])
        dgl, dtau = C.c_double(0), C.c_double(0)
        _check(lib().bn_window_backward(self.h, ptrs, C.byref(dgl), C.byref(dtau)))
        out = dW
        if dW is None and download:
            out = [self.gradient(k) for k in range(4)]
        return {"dW": out, "dg_leak": dgl.value, "dtau_decay_e": dtau.value}

    def gradient(self, idx: int) -> np.ndarray:
        out = np.zeros(self.shapes[idx], np.float32)
        _check(lib().bn_window_get_gradient(self.h, idx, _fp(out)))
        return out
