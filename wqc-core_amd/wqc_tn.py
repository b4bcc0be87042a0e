"""Python binding of libwqc_tn.so, the MI355X MPS engine, through its C ABI (include/wqc_tn.h).

``MpsState`` mirrors the reference's ``MpsState`` (src/tn/mps.rs:55-781): same method names,
argument meaning and error behaviour (a non-OK status raises ``WqcTnError`` carrying the
status code and the engine's message). The engine runs on the GPU only: there is no CPU
fallback, and loading fails loudly when the library or the device is missing.
"""

from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwqc_tn.so")

OK, INVALID_ARG, OUT_OF_MEMORY, EXEC_FAILED, CANCELLED, BOND_EXCEEDED, NO_DEVICE = range(7)
PRECISION_FP64, PRECISION_MIXED_FP32, PRECISION_MIXED_BF16 = 0, 1, 2
SAMPLE_EXACT, SAMPLE_MPS, SAMPLE_AUTO = 0, 1, 2

EXPORTED = (
    "wqc_tn_last_error", "wqc_tn_backend_from_env", "wqc_tn_engine_status", "wqc_tn_device_open",
    "wqc_tn_device_close", "wqc_tn_state_new", "wqc_tn_state_free", "wqc_tn_qubit_count",
    "wqc_tn_set_max_bond_dim", "wqc_tn_set_reject_truncation", "wqc_tn_apply_gate",
    "wqc_tn_apply_circuit", "wqc_tn_apply_adjacent_disjoint_gates", "wqc_tn_apply_unary_at",
    "wqc_tn_apply_tdg_at", "wqc_tn_apply_swap_at", "wqc_tn_apply_pair_at", "wqc_tn_apply_cnot_at",
    "wqc_tn_request_cancel", "wqc_tn_clear_cancel", "wqc_tn_amplitude_at_compact_zero",
    "wqc_tn_site_amplitudes", "wqc_tn_control_probability", "wqc_tn_contract_to_statevector",
    "wqc_tn_bond_dims", "wqc_tn_download_site", "wqc_tn_upload_site", "wqc_tn_backend_label",
    "wqc_tn_peak_vram_bytes", "wqc_tn_get_stats", "wqc_tn_kernel_times", "wqc_tn_enable_timing",
    "wqc_tn_kernel_launches", "wqc_tn_enable_split_log", "wqc_tn_split_log",
    "wqc_tn_pauli_expectations", "wqc_tn_observable_expectation", "wqc_tn_sample_counts",
    "wqc_tn_run_slices", "wqc_tn_backward", "wqc_tn_overlap", "wqc_tn_contract_time_split",
    "wqc_tn_site_amplitudes_batch", "wqc_tn_outcome_probabilities", "wqc_tn_copy_sites",
    "wqc_tn_amplitude_at_compact_zero_noalloc", "wqc_tn_split_discarded", "wqc_tn_split_probe_residuals",
)


class Gate(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("q0", ct.c_int32), ("q1", ct.c_int32), ("q2", ct.c_int32),
                ("angle", ct.c_double)]


class Slice(ct.Structure):
    _fields_ = [("qubit_count", ct.c_size_t), ("max_bond_dim", ct.c_size_t),
                ("gate_offset", ct.c_size_t), ("gate_count", ct.c_size_t)]


class Stats(ct.Structure):
    _fields_ = [("two_qubit_units", ct.c_uint64), ("one_qubit_gates", ct.c_uint64),
                ("exact_splits", ct.c_uint64), ("rsvd_kept", ct.c_uint64),
                ("rsvd_fallbacks", ct.c_uint64), ("svd_check_failures", ct.c_uint64),
                ("jacobi_sweeps", ct.c_uint64), ("truncated_splits", ct.c_uint64),
                ("peak_vram_bytes", ct.c_uint64), ("max_discarded_weight", ct.c_double),
                ("sum_discarded_weight", ct.c_double), ("jacobi_inner_sweeps", ct.c_uint64),
                ("jacobi_pair_solves", ct.c_uint64), ("jacobi_sweeps_lo", ct.c_uint64),
                ("jacobi_stalled", ct.c_uint64), ("jacobi_unconverged", ct.c_uint64)]


class WqcTnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"wqc_tn status {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libwqc_tn.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
    L = ct.CDLL(LIB_PATH)
    vp, sz, dp, i32 = ct.c_void_p, ct.c_size_t, ct.POINTER(ct.c_double), ct.c_int
    gp = ct.POINTER(Gate)
    L.wqc_tn_last_error.restype = ct.c_char_p
    L.wqc_tn_backend_from_env.restype = ct.c_char_p
    L.wqc_tn_engine_status.argtypes = [ct.c_char_p, sz, ct.POINTER(ct.c_uint64)]
    L.wqc_tn_device_open.argtypes = [i32, ct.POINTER(vp)]
    L.wqc_tn_device_close.argtypes = [vp]
    L.wqc_tn_state_new.argtypes = [vp, sz, sz, i32, ct.POINTER(vp)]
    L.wqc_tn_state_free.argtypes = [vp]
    L.wqc_tn_qubit_count.restype = sz
    L.wqc_tn_qubit_count.argtypes = [vp]
    L.wqc_tn_set_max_bond_dim.argtypes = [vp, sz]
    L.wqc_tn_set_reject_truncation.argtypes = [vp, i32]
    L.wqc_tn_apply_gate.argtypes = [vp, gp]
    L.wqc_tn_apply_circuit.argtypes = [vp, gp, sz]
    L.wqc_tn_apply_adjacent_disjoint_gates.argtypes = [vp, gp, sz]
    L.wqc_tn_apply_unary_at.argtypes = [vp, sz, gp]
    L.wqc_tn_apply_tdg_at.argtypes = [vp, sz]
    L.wqc_tn_apply_swap_at.argtypes = [vp, sz]
    L.wqc_tn_apply_pair_at.argtypes = [vp, sz, gp, i32]
    L.wqc_tn_apply_cnot_at.argtypes = [vp, sz, i32]
    L.wqc_tn_request_cancel.argtypes = [vp]
    L.wqc_tn_clear_cancel.argtypes = [vp]
    L.wqc_tn_amplitude_at_compact_zero.argtypes = [vp, dp]
    L.wqc_tn_amplitude_at_compact_zero_noalloc.argtypes = [vp, dp]
    L.wqc_tn_site_amplitudes.argtypes = [vp, sz, dp]
    L.wqc_tn_control_probability.argtypes = [vp, sz, dp]
    L.wqc_tn_contract_to_statevector.argtypes = [vp, dp, sz]
    L.wqc_tn_pauli_expectations.argtypes = [vp, ct.c_char_p, sz, dp, dp]
    L.wqc_tn_observable_expectation.argtypes = [vp, ct.c_char_p, dp, sz, dp]
    i32p, u64p = ct.POINTER(ct.c_int32), ct.POINTER(ct.c_uint64)
    L.wqc_tn_backward.argtypes = [vp, sz, sz, ct.c_int, ct.c_void_p, sz, ct.c_char_p, dp, sz, sz, sz,
                                  dp, dp]
    L.wqc_tn_run_slices.argtypes = [vp, ct.POINTER(Slice), sz, ct.c_void_p, ct.c_int, ct.c_int, dp,
                                    ct.POINTER(ct.c_int)]
    L.wqc_tn_sample_counts.argtypes = [vp, i32p, i32p, sz, sz, ct.c_uint64, ct.c_uint64, ct.c_int,
                                       u64p, u64p, sz, ct.POINTER(ct.c_size_t)]
    L.wqc_tn_overlap.argtypes = [vp, vp, dp]
    L.wqc_tn_site_amplitudes_batch.argtypes = [vp, ct.POINTER(ct.c_int32), sz, dp]
    L.wqc_tn_contract_time_split.argtypes = [vp, sz, ct.c_void_p, sz, sz, sz, ct.c_int, dp]
    L.wqc_tn_outcome_probabilities.argtypes = [vp, i32p, i32p, sz, sz, u64p, dp, sz, ct.POINTER(ct.c_size_t)]
    L.wqc_tn_bond_dims.argtypes = [vp, ct.POINTER(sz), sz]
    L.wqc_tn_download_site.argtypes = [vp, sz, dp, sz, ct.POINTER(sz), ct.POINTER(sz)]
    L.wqc_tn_upload_site.argtypes = [vp, sz, dp, sz, sz]
    L.wqc_tn_copy_sites.argtypes = [vp, vp]
    L.wqc_tn_backend_label.restype = ct.c_char_p
    L.wqc_tn_backend_label.argtypes = [vp]
    L.wqc_tn_peak_vram_bytes.restype = ct.c_uint64
    L.wqc_tn_peak_vram_bytes.argtypes = [vp]
    L.wqc_tn_get_stats.argtypes = [vp, ct.POINTER(Stats)]
    L.wqc_tn_kernel_times.argtypes = [vp, dp, sz]
    L.wqc_tn_enable_timing.argtypes = [vp, i32]
    L.wqc_tn_kernel_launches.argtypes = [vp, ct.POINTER(ct.c_uint64), sz]
    L.wqc_tn_enable_split_log.argtypes = [vp, i32]
    L.wqc_tn_split_log.restype = sz
    L.wqc_tn_split_log.argtypes = [vp, ct.POINTER(ct.c_int32), sz]
    L.wqc_tn_split_discarded.restype = sz
    L.wqc_tn_split_discarded.argtypes = [vp, dp, sz]
    L.wqc_tn_split_probe_residuals.restype = sz
    L.wqc_tn_split_probe_residuals.argtypes = [vp, dp, sz]
    _lib = L
    return L


def _check(rc: int):
    if rc != OK:
        raise WqcTnError(rc, load().wqc_tn_last_error().decode())


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def gate_array(gates):
    arr = (Gate * max(1, len(gates)))()
    for i, g in enumerate(gates):
        arr[i].kind, arr[i].q0, arr[i].q1, arr[i].q2, arr[i].angle = (
            int(g[0]), int(g[1]), int(g[2]), int(g[3]), float(g[4]))
    return arr


class Device:
    """An opened gfx950 device (GpuMpsDevice::try_new, src/tn/gpu/device.rs:68)."""

    def __init__(self, ordinal: int = 0):
        L = load()
        h = ct.c_void_p()
        _check(L.wqc_tn_device_open(ordinal, ct.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().wqc_tn_device_close(self._h)
            self._h = None


_default_device = None


def default_device() -> Device:
    global _default_device
    if _default_device is None:
        _default_device = Device(0)
    return _default_device


def run_slices(slices, precision: int = PRECISION_FP64, workers: int = 0, device=None):
    """<0...0|psi> of independent slices [(qubit_count, max_bond_dim, gates), ...]
    (contract_slice per slice, src/tn/contract.rs:10-30); returns (amplitudes, status codes).
    workers = 0: one batch, every layer step of all slices in shared launches; workers > 0:
    that many host threads, one state per slice at a time."""
    L = load()
    dev = device or default_device()
    descs = (Slice * max(1, len(slices)))()
    allg = []
    for i, (n, chi, gates) in enumerate(slices):
        descs[i].qubit_count, descs[i].max_bond_dim = n, chi
        descs[i].gate_offset, descs[i].gate_count = len(allg), len(gates)
        allg.extend(gates)
    garr = gate_array(allg)
    amps = np.zeros(2 * max(1, len(slices)))
    status = (ct.c_int * max(1, len(slices)))()
    rc = L.wqc_tn_run_slices(dev._h, descs, len(slices), ct.cast(garr, ct.c_void_p), precision,
                             workers, _dp(amps), status)
    out = amps[: 2 * len(slices)].view(np.complex128)
    codes = [status[i] for i in range(len(slices))]
    if rc != OK and all(c == OK for c in codes):
        _check(rc)
    return out, codes


def contract_time_split(qubit_count, gates, split, max_bond_dim, precision: int = PRECISION_FP64, device=None):
    """<0...0| U |0...0> as the overlap of |0...0> through gates[:split] with |0...0> through
    the adjoints of gates[split:] in reverse, both untruncated at max_bond_dim
    (contract_time_split, src/tn/mps.rs:1241-1249)."""
    L = load()
    dev = device or default_device()
    garr = gate_array(gates)
    out = np.zeros(2)
    _check(L.wqc_tn_contract_time_split(dev._h, qubit_count, ct.cast(garr, ct.c_void_p), len(gates), split,
                                        max_bond_dim, precision, _dp(out)))
    return complex(out[0], out[1])


def backward(qubit_count, max_bond_dim, gates, terms, checkpoint_every=2, adjoint_bond=0,
             precision: int = PRECISION_FP64, device=None):
    """Adjoint gradient of sum_k c_k <psi|P_k|psi> over (label, real coeff) terms with respect
    to every RX/RY/RZ angle, truncated splits included (reverse mode over the site tensors of the
    plain FP64 exact-split forward pass); returns (value, grad) with grad[i] for gates[i]."""
    L = load()
    dev = device or default_device()
    garr = gate_array(gates)
    labels = "".join(t[0] for t in terms).encode("ascii")
    co = np.ascontiguousarray([float(t[1]) for t in terms], dtype=np.float64)
    grad = np.zeros(max(1, len(gates)))
    val = np.zeros(2)
    _check(L.wqc_tn_backward(dev._h, qubit_count, max_bond_dim, precision, ct.cast(garr, ct.c_void_p),
                             len(gates), labels, _dp(co), len(terms), checkpoint_every, adjoint_bond,
                             _dp(grad), _dp(val)))
    return complex(val[0], val[1]), grad[: len(gates)]


class MpsState:
    """Device-resident MPS (src/tn/mps.rs:55-781)."""

    def __init__(self, qubit_count: int, max_bond_dim: int, precision: int = PRECISION_FP64,
                 device: Device | None = None):
        L = load()
        dev = device or default_device()
        h = ct.c_void_p()
        _check(L.wqc_tn_state_new(dev._h, qubit_count, max_bond_dim, precision, ct.byref(h)))
        self._h = h
        self._dev = dev
        self.qubit_count = qubit_count
        self.max_bond_dim = max_bond_dim

    try_new_with_bond = classmethod(lambda cls, n, chi, **kw: cls(n, chi, **kw))

    def free(self):
        if getattr(self, "_h", None):
            load().wqc_tn_state_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def set_max_bond_dim(self, chi: int):
        _check(load().wqc_tn_set_max_bond_dim(self._h, chi))
        self.max_bond_dim = chi

    def set_reject_truncation(self, on: bool):
        _check(load().wqc_tn_set_reject_truncation(self._h, 1 if on else 0))

    def apply_gate(self, g):
        _check(load().wqc_tn_apply_gate(self._h, gate_array([g])))

    def apply_gates(self, gates):
        for g in gates:
            self.apply_gate(g)

    def apply_circuit(self, gates):
        _check(load().wqc_tn_apply_circuit(self._h, gate_array(gates), len(gates)))

    def apply_adjacent_disjoint_gates(self, gates):
        _check(load().wqc_tn_apply_adjacent_disjoint_gates(self._h, gate_array(gates), len(gates)))

    def apply_unary_at(self, site, g):
        _check(load().wqc_tn_apply_unary_at(self._h, site, gate_array([g])))

    def apply_tdg_at(self, site):
        _check(load().wqc_tn_apply_tdg_at(self._h, site))

    def apply_swap_at(self, site):
        _check(load().wqc_tn_apply_swap_at(self._h, site))

    def apply_pair_at(self, site, g, ctrl_on_left=True):
        _check(load().wqc_tn_apply_pair_at(self._h, site, gate_array([g]), 1 if ctrl_on_left else 0))

    def apply_cnot_at(self, site, ctrl_on_left=True):
        _check(load().wqc_tn_apply_cnot_at(self._h, site, 1 if ctrl_on_left else 0))

    def request_cancel(self):
        load().wqc_tn_request_cancel(self._h)

    def clear_cancel(self):
        load().wqc_tn_clear_cancel(self._h)

    def amplitude_at_compact_zero(self) -> complex:
        out = np.zeros(2)
        _check(load().wqc_tn_amplitude_at_compact_zero(self._h, _dp(out)))
        return complex(out[0], out[1])

    def amplitude_at_compact_zero_noalloc(self) -> complex:
        out = np.zeros(2)
        _check(load().wqc_tn_amplitude_at_compact_zero_noalloc(self._h, _dp(out)))
        return complex(out[0], out[1])

    def overlap(self, bra: "MpsState") -> complex:
        """<bra|self> (overlap, src/tn/mps.rs:1421-1474); environments above 256 entries fail."""
        out = np.zeros(2)
        _check(load().wqc_tn_overlap(self._h, bra._h, _dp(out)))
        return complex(out[0], out[1])

    def site_amplitudes_batch(self, targets):
        """[(a0, a1)] of site_amplitudes for every target, from one sweep of environments."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        out = np.zeros(4 * max(1, len(t)))
        _check(load().wqc_tn_site_amplitudes_batch(self._h, t.ctypes.data_as(ct.POINTER(ct.c_int32)), len(t),
                                                   _dp(out)))
        return [(complex(out[4 * k], out[4 * k + 1]), complex(out[4 * k + 2], out[4 * k + 3]))
                for k in range(len(t))]

    def site_amplitudes(self, target: int):
        out = np.zeros(4)
        _check(load().wqc_tn_site_amplitudes(self._h, target, _dp(out)))
        return complex(out[0], out[1]), complex(out[2], out[3])

    def control_probability(self, control: int) -> float:
        out = np.zeros(1)
        _check(load().wqc_tn_control_probability(self._h, control, _dp(out)))
        return float(out[0])

    def contract_to_statevector(self):
        n = self.qubit_count
        out = np.zeros(2 ** (n + 1)) if n <= 20 else np.zeros(2)
        _check(load().wqc_tn_contract_to_statevector(self._h, _dp(out), out.size // 2))
        return out.view(np.complex128)

    def pauli_expectations(self, labels, with_imag=False):
        """<psi|P|psi> per Pauli label (Qiskit order, un-normalized), src/expectation.rs:140-167."""
        labels = list(labels)
        buf = "".join(labels).encode("ascii")
        re = np.zeros(max(1, len(labels)))
        im = np.zeros(max(1, len(labels)))
        _check(load().wqc_tn_pauli_expectations(self._h, buf, len(labels), _dp(re), _dp(im)))
        if with_imag:
            return re[:len(labels)] + 1j * im[:len(labels)]
        return re[:len(labels)]

    def sample_counts(self, measures, classical_bits, shots, seed, method=None):
        """{bitstring: count} with the reference's Qiskit keys (cbit 0 rightmost),
        src/sample.rs:192-211; measures = [(qubit, cbit), ...]."""
        if method is None:
            method = SAMPLE_AUTO
        mq = np.array([m[0] for m in measures] or [0], dtype=np.int32)
        mc = np.array([m[1] for m in measures] or [0], dtype=np.int32)
        cap = int(max(1, min(shots, 1 << min(classical_bits, 62))))
        keys = np.zeros(cap, dtype=np.uint64)
        cnt = np.zeros(cap, dtype=np.uint64)
        nout = ct.c_size_t(0)
        i32p, u64p = ct.POINTER(ct.c_int32), ct.POINTER(ct.c_uint64)
        _check(load().wqc_tn_sample_counts(self._h, mq.ctypes.data_as(i32p), mc.ctypes.data_as(i32p),
                                           len(measures), classical_bits, shots, seed, method,
                                           keys.ctypes.data_as(u64p), cnt.ctypes.data_as(u64p), cap,
                                           ct.byref(nout)))
        out = {}
        for i in range(nout.value):
            out[format(int(keys[i]), "0%db" % classical_bits)] = int(cnt[i])
        return out

    def outcome_probabilities(self, measures, classical_bits):
        """{bitstring: probability} of compute_outcome_probabilities (src/sample.rs:85-131), the
        snapped Born table of the projected statevector (qubit_count <= 20)."""
        mq = np.array([m[0] for m in measures] or [0], dtype=np.int32)
        mc = np.array([m[1] for m in measures] or [0], dtype=np.int32)
        cap = int(max(1, 1 << min(max(classical_bits, 0), self.qubit_count, 20)))
        keys = np.zeros(cap, dtype=np.uint64)
        probs = np.zeros(cap)
        nout = ct.c_size_t(0)
        i32p, u64p = ct.POINTER(ct.c_int32), ct.POINTER(ct.c_uint64)
        _check(load().wqc_tn_outcome_probabilities(self._h, mq.ctypes.data_as(i32p), mc.ctypes.data_as(i32p),
                                                   len(measures), classical_bits, keys.ctypes.data_as(u64p),
                                                   _dp(probs), cap, ct.byref(nout)))
        return {format(int(keys[i]), "0%db" % classical_bits): float(probs[i]) for i in range(nout.value)}

    def observable_expectation(self, terms):
        """Snapped sum of coeff * <P> over (label, coeff) terms, src/expectation.rs:170-200."""
        labels = "".join(t[0] for t in terms).encode("ascii")
        co = np.array([[complex(t[1]).real, complex(t[1]).imag] for t in terms], dtype=np.float64).ravel()
        out = np.zeros(2)
        _check(load().wqc_tn_observable_expectation(self._h, labels, _dp(co), len(terms), _dp(out)))
        return complex(out[0], out[1])

    def bond_dims(self):
        out = (ct.c_size_t * self.qubit_count)()
        _check(load().wqc_tn_bond_dims(self._h, out, self.qubit_count))
        return [int(x) for x in out]

    def site(self, i: int) -> np.ndarray:
        l, r = ct.c_size_t(), ct.c_size_t()
        _check(load().wqc_tn_download_site(self._h, i, None, 0, ct.byref(l), ct.byref(r)))
        out = np.zeros(l.value * 2 * r.value * 2)
        _check(load().wqc_tn_download_site(self._h, i, _dp(out), out.size // 2, ct.byref(l), ct.byref(r)))
        return out.view(np.complex128).reshape(l.value, 2, r.value)

    def set_site(self, i: int, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        _check(load().wqc_tn_upload_site(self._h, i, _dp(a.view(np.float64)), a.shape[0], a.shape[2]))

    def copy_sites_from(self, src: "MpsState"):
        """self.sites = src.sites (device to device)."""
        _check(load().wqc_tn_copy_sites(self._h, src._h))

    def backend_label(self) -> str:
        return load().wqc_tn_backend_label(self._h).decode()

    def peak_vram_bytes(self) -> int:
        return int(load().wqc_tn_peak_vram_bytes(self._h))

    def stats(self) -> dict:
        s = Stats()
        _check(load().wqc_tn_get_stats(self._h, ct.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def enable_timing(self, on=True):
        _check(load().wqc_tn_enable_timing(self._h, 1 if on else 0))

    FAMILIES = ("one_qubit", "merge", "range_finder", "jacobi", "factors", "readout", "lq", "basis")

    def kernel_times(self):
        """Device milliseconds per kernel family (timing must be enabled)."""
        out = np.zeros(len(self.FAMILIES))
        k = load().wqc_tn_kernel_times(self._h, _dp(out), len(self.FAMILIES))
        return dict(zip(self.FAMILIES[:k], out[:k].tolist()))

    def kernel_launches(self):
        out = (ct.c_uint64 * len(self.FAMILIES))()
        k = load().wqc_tn_kernel_launches(self._h, out, len(self.FAMILIES))
        return dict(zip(self.FAMILIES[:k], [int(x) for x in out[:k]]))

    def enable_split_log(self, on=True):
        _check(load().wqc_tn_enable_split_log(self._h, 1 if on else 0))

    def split_log(self):
        """[(rows, cols, bond, path, ell, chi_kept, fp64_sweeps, low_sweeps)] per two-qubit
        unit since enabling."""
        m = load().wqc_tn_split_log(self._h, None, 0)
        buf = (ct.c_int32 * max(8, 8 * m))()
        load().wqc_tn_split_log(self._h, buf, m)
        a = np.frombuffer(buf, dtype=np.int32)[: 8 * m].reshape(m, 8)
        return [tuple(int(x) for x in row) for row in a]

    def split_discarded(self):
        """Discarded weight of each unit in split-log order."""
        m = load().wqc_tn_split_discarded(self._h, None, 0)
        out = np.zeros(max(1, m))
        load().wqc_tn_split_discarded(self._h, _dp(out), m)
        return out[:m].tolist()

    def split_probe_residuals(self):
        """Relative probe residual of each unit's kept factorization, in split-log order."""
        m = load().wqc_tn_split_probe_residuals(self._h, None, 0)
        out = np.zeros(max(1, m))
        load().wqc_tn_split_probe_residuals(self._h, _dp(out), m)
        return out[:m].tolist()


def backend_from_env() -> str:
    return load().wqc_tn_backend_from_env().decode()
