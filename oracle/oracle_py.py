"""ORACLE binding — test infrastructure only.

ctypes wrapper of oracle/liboracle.so (the CPU restatement of the reference step, see
oracle/oracle.h).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
and only as the reference for checks / CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class GroupParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "mean_background_excitatory_spikes", "mean_background_inhibitory_spikes",
        "background_excitatory_postsynaptic_r_trace_decay", "background_excitatory_postsynaptic_d_trace_decay",
        "background_inhibitory_postsynaptic_r_trace_decay", "background_inhibitory_postsynaptic_d_trace_decay",
        "background_excitatory_scaling_factor", "background_inhibitory_scaling_factor")]


class Params(C.Structure):
    _fields_ = [
        ("num_excitatory_neurons", C.c_uint32), ("num_inhibitory_neurons", C.c_uint32),
        ("rng_seed", C.c_int32), ("rng_instance", C.c_uint32),
        ("timestep", C.c_float), ("g_leak", C.c_float), ("membrane_capacitance", C.c_float),
        ("depolarization_slope_factor", C.c_float), ("v_effective_spike_threshold", C.c_float),
        ("v_excitatory_reversal", C.c_float), ("v_inhibitory_reversal", C.c_float),
        ("v_threshold", C.c_float), ("v_reset", C.c_float), ("v_rest", C.c_float),
        ("refractory_steps", C.c_uint32), ("group", GroupParams * 2),
        ("tau_decay_e", C.c_float), ("tau_rise_e", C.c_float), ("tau_decay_i", C.c_float), ("tau_rise_i", C.c_float),
        ("postsynaptic_kernel_normalization_value", C.c_float),
        ("e_e_plasticity_enabled", C.c_uint32),
        ("e_e_ltp_A", C.c_float), ("e_e_ltd_A", C.c_float), ("e_e_ltp_B", C.c_float), ("e_e_ltd_B", C.c_float),
        ("e_e_tau_ltp", C.c_float), ("e_e_tau_ltd", C.c_float), ("e_e_triplet_tau_ltp", C.c_float),
        ("e_e_triplet_tau_ltd", C.c_float), ("e_e_weight_min", C.c_float), ("e_e_weight_max", C.c_float),
        ("i_e_plasticity_enabled", C.c_uint32),
        ("i_e_ltp_A", C.c_float), ("i_e_ltd_A", C.c_float), ("i_e_ltd_B", C.c_float), ("i_e_tau_ltp", C.c_float),
        ("i_e_tau_ltd", C.c_float), ("i_e_weight_min", C.c_float), ("i_e_weight_max", C.c_float),
        ("gemv_row_block", C.c_uint32 * 4), ("lazy_normalize", C.c_uint32),
        ("dense_gemv", C.c_uint32), ("gemv_part_chunks", C.c_uint32 * 4), ("lazy_tile_order", C.c_uint32),
        ("e_e_pair_stdp", C.c_uint32), ("resident_columns", C.c_uint32), ("resident_e_waves", C.c_uint32),
    ]


_lib = None


class WindowState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("V", "refr", "bg_ge", "bg_gi", "bre", "bde", "bri", "bdi", "r_e", "d_e",
                                          "r_i", "d_i")] + [("clock_step", C.c_uint32)]


class WindowTrajectory(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("Vprev", "ge", "gi", "refr", "spikes", "d_e", "r_e", "d_i", "r_i",
                                          "v_rec")]


class WindowSpec(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("n_vrec", C.c_uint32), ("vrec_ids", C.c_void_p), ("rate_tau", C.c_double),
                ("surrogate_beta", C.c_double), ("target_rates", C.c_void_p), ("target_v", C.c_void_p),
                ("forced_spikes", C.c_void_p)]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, capture_output=True)
        L = C.CDLL(LIB_PATH)
        P, U, F, I = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        fp = C.POINTER(C.c_float)
        up = C.POINTER(C.c_uint32)
        L.oracle_create.restype = P
        L.oracle_create.argtypes = [C.POINTER(Params), I, fp, fp, fp, fp, up, U, up, up, fp, U]
        L.oracle_destroy.argtypes = [P]
        L.oracle_step.argtypes = [P, U]
        L.oracle_step.restype = I
        L.oracle_clock_step.argtypes = [P]
        L.oracle_clock_step.restype = U
        L.oracle_get_state.argtypes = [P, fp, fp, fp, fp, up, up, fp, fp]
        dp = C.POINTER(C.c_double)
        L.oracle_get_state64.argtypes = [P, dp, dp, dp]
        L.oracle_get_weights.argtypes = [P, U, fp]
        L.oracle_get_connection_traces.argtypes = [P, U, fp, fp, fp]
        L.oracle_get_plasticity_traces.argtypes = [P, U, fp, fp, fp, fp]
        L.oracle_philox4x32_10.argtypes = [up, up, up]
        L.oracle_expf.argtypes = [F]
        L.oracle_expf.restype = F
        L.oracle_poisson_table.argtypes = [F, up, U]
        L.oracle_poisson_table.restype = U
        L.oracle_window.argtypes = [C.POINTER(Params), C.POINTER(dp), C.c_double, C.c_double,
                                    C.POINTER(WindowState), C.POINTER(WindowSpec), dp, C.POINTER(dp), dp, dp, dp, dp,
                                    C.POINTER(C.c_uint8)]
        L.oracle_window.restype = I
        L.oracle_window_along.argtypes = [C.POINTER(Params), C.POINTER(dp), C.c_double, C.c_double,
                                          C.POINTER(WindowTrajectory), C.POINTER(WindowSpec), dp, C.POINTER(dp), dp,
                                          dp]
        L.oracle_window_along.restype = I
        fpp = C.POINTER(C.c_float)
        u8 = C.POINTER(C.c_uint8)
        L.oracle_window_records.argtypes = [C.POINTER(Params), C.POINTER(dp), C.c_double, C.c_double,
                                            C.POINTER(WindowState), C.POINTER(WindowSpec), fpp, fpp, fpp, u8, u8,
                                            fpp, fpp, fpp, fpp, fpp]
        L.oracle_window_records.restype = I
        _lib = L
    return _lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


def params_from_config(cfg: dict, gemv_row_block=(0, 0, 0, 0), lazy_normalize=False,
                       rng_instance=0, dense_gemv=False, gemv_part_chunks=(0, 0, 0, 0),
                       lazy_tile_order=False, resident=(0, 0)) -> Params:
    """Builds oracle parameters from a parsed config (the engine's config dump).  resident:
    (post columns per block, waves of excitatory rows) of the resident step's orders, (0, 0) off."""
    p = Params()
    p.num_excitatory_neurons = cfg["num_excitatory_neurons"]
    p.num_inhibitory_neurons = cfg["num_inhibitory_neurons"]
    noise_seed = cfg.get("background_noise_seed", -1)
    p.rng_seed = noise_seed if noise_seed is not None and noise_seed >= 0 else cfg["rng_seed"]
    p.rng_instance = rng_instance
    for k in ("timestep", "g_leak", "membrane_capacitance", "depolarization_slope_factor",
              "v_effective_spike_threshold", "v_excitatory_reversal", "v_inhibitory_reversal",
              "v_threshold", "v_reset", "v_rest", "tau_decay_e", "tau_rise_e", "tau_decay_i", "tau_rise_i",
              "postsynaptic_kernel_normalization_value"):
        setattr(p, k, cfg[k])
    p.refractory_steps = cfg["refractory_steps"]
    for gi, g in enumerate(("e", "i")):
        for f, _ in GroupParams._fields_:
            setattr(p.group[gi], f, cfg[f"group_{g}_{f}"])
    p.e_e_plasticity_enabled = int(cfg["e_e_plasticity_enabled"])
    p.i_e_plasticity_enabled = int(cfg["i_e_plasticity_enabled"])
    for a, b in (("e_e_ltp_A", "e_e_plasticity_ltp_A"), ("e_e_ltd_A", "e_e_plasticity_ltd_A"),
                 ("e_e_ltp_B", "e_e_plasticity_ltp_B"), ("e_e_ltd_B", "e_e_plasticity_ltd_B"),
                 ("e_e_tau_ltp", "e_e_plasticity_tau_ltp"), ("e_e_tau_ltd", "e_e_plasticity_tau_ltd"),
                 ("e_e_triplet_tau_ltp", "e_e_plasticity_triplet_tau_ltp"),
                 ("e_e_triplet_tau_ltd", "e_e_plasticity_triplet_tau_ltd"),
                 ("e_e_weight_min", "e_e_plasticity_weight_min"), ("e_e_weight_max", "e_e_plasticity_weight_max"),
                 ("i_e_ltp_A", "i_e_plasticity_ltp_A"), ("i_e_ltd_A", "i_e_plasticity_ltd_A"),
                 ("i_e_ltd_B", "i_e_plasticity_ltd_B"), ("i_e_tau_ltp", "i_e_plasticity_tau_ltp"),
                 ("i_e_tau_ltd", "i_e_plasticity_tau_ltd"), ("i_e_weight_min", "i_e_plasticity_weight_min"),
                 ("i_e_weight_max", "i_e_plasticity_weight_max")):
        setattr(p, a, cfg[b])
    for k in range(4):
        p.gemv_row_block[k] = int(gemv_row_block[k])
    p.lazy_normalize = int(lazy_normalize)
    p.dense_gemv = int(dense_gemv)
    for k in range(4):
        p.gemv_part_chunks[k] = int(gemv_part_chunks[k])
    p.lazy_tile_order = int(lazy_tile_order)
    p.e_e_pair_stdp = int(cfg.get("e_e_plasticity_rule", "triplet") == "pair")
    p.resident_columns, p.resident_e_waves = int(resident[0]), int(resident[1])
    return p


class OracleSim:
    """One network on the CPU oracle. Weights: list of 4 dense [pre][post] float32 arrays."""

    def __init__(self, params: Params, weights, precision=32, stim=None):
        self.p = params
        self.NE = params.num_excitatory_neurons
        self.NI = params.num_inhibitory_neurons
        self.N = self.NE + self.NI
        self._w = [np.ascontiguousarray(w, dtype=np.float32) for w in weights]
        if stim is None:
            counts = ids = offs = amps = None
            nt = ne = 0
        else:
            counts, ids, offs, amps = [np.ascontiguousarray(a) for a in stim]
            counts = counts.astype(np.uint32)
            ids = ids.astype(np.uint32)
            offs = offs.astype(np.uint32)
            amps = amps.astype(np.float32)
            nt, ne = len(counts), len(ids)
            if ne == 0:
                ids = offs = None
                amps = None
        self._stim = (counts, ids, offs, amps)
        self.h = lib().oracle_create(C.byref(params), precision, *[_fp(w) for w in self._w],
                                     _up(counts), nt, _up(ids), _up(offs), _fp(amps), ne)
        self.precision = precision

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n=1):
        rc = lib().oracle_step(self.h, n)
        if rc:
            raise RuntimeError("Stimulus event indexing exceeded uploaded event buffers")

    @property
    def clock_step(self):
        return lib().oracle_clock_step(self.h)

    def state(self):
        N = self.N
        out = {k: np.zeros(N, np.float32) for k in ("V", "ge", "gi", "istim", "bg_ge", "bg_gi")}
        refr = np.zeros(N, np.uint32)
        spikes = np.zeros(N, np.uint32)
        lib().oracle_get_state(self.h, _fp(out["V"]), _fp(out["ge"]), _fp(out["gi"]), _fp(out["istim"]),
                               _up(refr), _up(spikes), _fp(out["bg_ge"]), _fp(out["bg_gi"]))
        out["refr"] = refr
        out["spikes"] = spikes
        return out

    def state64(self):
        N = self.N
        v, ge, gi = (np.zeros(N, np.float64) for _ in range(3))
        dp = C.POINTER(C.c_double)
        lib().oracle_get_state64(self.h, v.ctypes.data_as(dp), ge.ctypes.data_as(dp), gi.ctypes.data_as(dp))
        return {"V": v, "ge": ge, "gi": gi}

    def weights(self, idx):
        pre, post = self._w[idx].shape
        out = np.zeros((pre, post), np.float32)
        lib().oracle_get_weights(self.h, idx, _fp(out))
        return out

    def connection_traces(self, idx):
        n = self._w[idx].shape[0]
        r, d, x = (np.zeros(n, np.float32) for _ in range(3))
        lib().oracle_get_connection_traces(self.h, idx, _fp(r), _fp(d), _fp(x))
        return r, d, x

    def plasticity_traces(self, off):
        a = [np.zeros(self.N, np.float32) for _ in range(4)]
        lib().oracle_get_plasticity_traces(self.h, off, *[_fp(x) for x in a])
        return a


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(_up(c), _up(k), _up(out))
    return out


def expf(x: float) -> float:
    return lib().oracle_expf(x)


def poisson_table(mean: float, cap=256):
    t = np.zeros(cap, np.uint32)
    n = lib().oracle_poisson_table(mean, _up(t), cap)
    return t[:n]


def window(params: Params, W, g_leak: float, tau_decay_e: float, state: dict, steps: int, vrec_ids,
           rate_tau: float = 5e-3, beta: float = 1e-3, target_rates=None, target_v=None, forced_spikes=None,
           grad: bool = True) -> dict:
    """FP64 forward + reverse pass of the differentiable window (oracle/adjoint.cpp).

    W: four float64 arrays [pre][post] (EE, EI, IE, II).  state: dict of float32/uint32 arrays
    V, refr, bg_ge, bg_gi, bre, bde, bri, bdi (length N), r_e, d_e (N_E), r_i, d_i (N_I) and
    clock_step.  Returns loss, rates [steps][2], v [steps][n_vrec], spikes [steps][N] and, with
    grad, dW (list of 4), dg_leak, dtau_decay_e."""
    NE, NI = params.num_excitatory_neurons, params.num_inhibitory_neurons
    N = NE + NI
    Wd = [np.ascontiguousarray(w, np.float64) for w in W]
    keep = []

    def ptr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data

    st = WindowState()
    for name in ("V", "bg_ge", "bg_gi", "bre", "bde", "bri", "bdi", "r_e", "d_e", "r_i", "d_i"):
        setattr(st, name, ptr(state[name], np.float32))
    st.refr = ptr(state["refr"], np.uint32)
    st.clock_step = int(state["clock_step"])
    vrec = np.ascontiguousarray(vrec_ids, np.uint32)
    sp = WindowSpec()
    sp.steps, sp.n_vrec, sp.vrec_ids = steps, vrec.size, vrec.ctypes.data
    sp.rate_tau, sp.surrogate_beta = rate_tau, beta
    sp.target_rates = None if target_rates is None else ptr(target_rates, np.float64)
    sp.target_v = None if target_v is None else ptr(target_v, np.float64)
    sp.forced_spikes = None if forced_spikes is None else ptr(forced_spikes, np.uint8)
    dp = C.POINTER(C.c_double)
    Wp = (dp * 4)(*[w.ctypes.data_as(dp) for w in Wd])
    loss = C.c_double(0)
    rates = np.zeros((steps, 2))
    v = np.zeros((steps, vrec.size))
    spikes = np.zeros((steps, N), np.uint8)
    out = {"loss": None, "rates": rates, "v": v, "spikes": spikes}
    dW = [np.zeros_like(w) for w in Wd] if grad else None
    dWp = (dp * 4)(*[d.ctypes.data_as(dp) for d in dW]) if grad else None
    dgl, dtau = C.c_double(0), C.c_double(0)
    rc = lib().oracle_window(C.byref(params), Wp, g_leak, tau_decay_e, C.byref(st), C.byref(sp), C.byref(loss), dWp,
                             C.byref(dgl), C.byref(dtau), rates.ctypes.data_as(dp), v.ctypes.data_as(dp),
                             spikes.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc:
        raise RuntimeError("oracle_window failed")
    out["loss"] = loss.value
    if grad:
        out.update({"dW": dW, "dg_leak": dgl.value, "dtau_decay_e": dtau.value})
    return out


def _window_state(state: dict, keep: list) -> "WindowState":
    def ptr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data

    st = WindowState()
    for name in ("V", "bg_ge", "bg_gi", "bre", "bde", "bri", "bdi", "r_e", "d_e", "r_i", "d_i"):
        setattr(st, name, ptr(state[name], np.float32))
    st.refr = ptr(state["refr"], np.uint32)
    st.clock_step = int(state["clock_step"])
    return st


def window_records(params: Params, W, g_leak: float, tau_decay_e: float, state: dict, steps: int, vrec_ids) -> dict:
    """The FP64 forward's per-step records rounded to FP32, in window_along's trajectory layout."""
    NE, NI = params.num_excitatory_neurons, params.num_inhibitory_neurons
    N = NE + NI
    keep = []
    st = _window_state(state, keep)
    vrec = np.ascontiguousarray(vrec_ids, np.uint32)
    sp = WindowSpec()
    sp.steps, sp.n_vrec, sp.vrec_ids = steps, vrec.size, vrec.ctypes.data
    sp.rate_tau, sp.surrogate_beta = 5e-3, 0.0
    dp = C.POINTER(C.c_double)
    Wd = [np.ascontiguousarray(w, np.float64) for w in W]
    Wp = (dp * 4)(*[w.ctypes.data_as(dp) for w in Wd])
    out = {k: np.zeros((steps, N), np.float32) for k in ("Vprev", "ge", "gi")}
    out.update({k: np.zeros((steps, N), np.uint8) for k in ("refr", "spikes")})
    out.update({k: np.zeros((steps, NE), np.float32) for k in ("d_e", "r_e")})
    out.update({k: np.zeros((steps, NI), np.float32) for k in ("d_i", "r_i")})
    out["v_rec"] = np.zeros((steps, vrec.size), np.float32)
    f = lambda k: out[k].ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    u = lambda k: out[k].ctypes.data_as(C.POINTER(C.c_uint8))  # noqa: E731
    if lib().oracle_window_records(C.byref(params), Wp, g_leak, tau_decay_e, C.byref(st), C.byref(sp), f("Vprev"),
                                   f("ge"), f("gi"), u("refr"), u("spikes"), f("d_e"), f("r_e"), f("d_i"), f("r_i"),
                                   f("v_rec")):
        raise RuntimeError("oracle_window_records failed")
    return out


def window_along(params: Params, W, g_leak: float, tau_decay_e: float, traj: dict, vrec_ids,
                 rate_tau: float = 5e-3, beta: float = 1e-3, target_rates=None, target_v=None) -> dict:
    """Loss and FP64 reverse pass of the window along a recorded trajectory (oracle/adjoint.cpp
    oracle_window_along).  traj: Vprev, ge, gi [T][N] float32; refr, spikes [T][N] uint8;
    d_e, r_e [T][N_E], d_i, r_i [T][N_I] float32 (traces before each step); v_rec [T][n_vrec]
    float32 (V after each step)."""
    T = traj["Vprev"].shape[0]
    keep = []

    def ptr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data

    tr = WindowTrajectory()
    for name in ("Vprev", "ge", "gi", "d_e", "r_e", "d_i", "r_i", "v_rec"):
        setattr(tr, name, ptr(traj[name], np.float32))
    tr.refr = ptr(traj["refr"], np.uint8)
    tr.spikes = ptr(traj["spikes"], np.uint8)
    vrec = np.ascontiguousarray(vrec_ids, np.uint32)
    sp = WindowSpec()
    sp.steps, sp.n_vrec, sp.vrec_ids = T, vrec.size, vrec.ctypes.data
    sp.rate_tau, sp.surrogate_beta = rate_tau, beta
    sp.target_rates = None if target_rates is None else ptr(target_rates, np.float64)
    sp.target_v = None if target_v is None else ptr(target_v, np.float64)
    dp = C.POINTER(C.c_double)
    Wd = [np.ascontiguousarray(w, np.float64) for w in W]
    Wp = (dp * 4)(*[w.ctypes.data_as(dp) for w in Wd])
    dW = [np.zeros_like(w) for w in Wd]
    dWp = (dp * 4)(*[d.ctypes.data_as(dp) for d in dW])
    loss, dgl, dtau = C.c_double(0), C.c_double(0), C.c_double(0)
    if lib().oracle_window_along(C.byref(params), Wp, g_leak, tau_decay_e, C.byref(tr), C.byref(sp), C.byref(loss),
                                 dWp, C.byref(dgl), C.byref(dtau)):
        raise RuntimeError("oracle_window_along failed")
    return {"loss": loss.value, "dW": dW, "dg_leak": dgl.value, "dtau_decay_e": dtau.value}
