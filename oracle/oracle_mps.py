"""ctypes binding of the CPU FP64 oracle (liboracle_mps.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg use it as the checker. The product path never imports this module.
"""

from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_mps.so")


class OGate(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("q0", ct.c_int32), ("q1", ct.c_int32),
                ("q2", ct.c_int32), ("angle", ct.c_double)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        vp, sz, dp, i32 = ct.c_void_p, ct.c_size_t, ct.POINTER(ct.c_double), ct.c_int
        szp = ct.POINTER(ct.c_size_t)
        gp = ct.POINTER(OGate)
        L.oracle_last_error.restype = ct.c_char_p
        L.oracle_new.restype = vp
        L.oracle_new.argtypes = [sz, sz]
        L.oracle_free.argtypes = [vp]
        L.oracle_set_reject_truncation.argtypes = [vp, i32]
        L.oracle_set_max_bond.argtypes = [vp, sz]
        L.oracle_set_record.argtypes = [vp, i32]
        for f in ("oracle_apply_gates", "oracle_apply_circuit", "oracle_apply_disjoint"):
            getattr(L, f).argtypes = [vp, gp, sz]
        L.oracle_apply_gate.argtypes = [vp, gp]
        L.oracle_apply_unary_at.argtypes = [vp, sz, gp]
        L.oracle_apply_pair_at.argtypes = [vp, sz, i32, i32]
        L.oracle_amplitude_zero.argtypes = [vp, dp]
        L.oracle_qubits.restype = sz
        L.oracle_qubits.argtypes = [vp]
        L.oracle_bond_dims.argtypes = [vp, szp, sz]
        L.oracle_get_site.argtypes = [vp, sz, dp, sz, szp, szp]
        L.oracle_set_site.argtypes = [vp, sz, dp, sz, sz]
        L.oracle_statevector.argtypes = [vp, dp, sz]
        L.oracle_dense_state.argtypes = [vp, dp, sz]
        L.oracle_site_amplitudes.argtypes = [vp, sz, dp]
        L.oracle_counters.argtypes = [vp, ct.POINTER(ct.c_uint64), sz]
        L.oracle_max_discarded.restype = ct.c_double
        L.oracle_max_discarded.argtypes = [vp]
        L.oracle_split_log.restype = sz
        L.oracle_split_log.argtypes = [vp, ct.POINTER(ct.c_int32), szp, sz]
        L.oracle_split_theta.argtypes = [dp, sz, sz, sz, i32, dp, dp, szp,
                                         ct.POINTER(ct.c_int), dp]
        L.oracle_rsvd_sites.argtypes = [dp, sz, sz, sz, sz, sz, sz, dp, dp, szp]
        L.oracle_checked_svd.argtypes = [dp, sz, sz, dp, dp, dp]
        L.oracle_retained_chi.restype = sz
        L.oracle_retained_chi.argtypes = [dp, sz, sz, sz, sz]
        L.oracle_needs_f64_fallback.argtypes = [dp, sz, sz]
        L.oracle_mix_seed.restype = ct.c_uint64
        L.oracle_mix_seed.argtypes = [sz, sz, sz]
        L.oracle_rademacher.argtypes = [sz, sz, ct.c_uint64, dp]
        L.oracle_thin_q.argtypes = [dp, sz, sz, dp]
        L.oracle_thin_q_reflectors.argtypes = [dp, sz, sz, ct.c_double, dp, dp, ct.POINTER(ct.c_int)]
        L.oracle_pauli_expectation.argtypes = [dp, sz, ct.c_char_p, dp]
        L.oracle_sample_counts.argtypes = [dp, sz, ct.POINTER(ct.c_int32),
                                           ct.POINTER(ct.c_int32), sz, sz, ct.c_uint64,
                                           ct.c_uint64, ct.POINTER(ct.c_uint64)]
        L.oracle_chacha_block.argtypes = [ct.POINTER(ct.c_uint32), ct.c_uint64, ct.c_uint64,
                                          i32, ct.POINTER(ct.c_uint32)]
        L.oracle_stdrng_f64.argtypes = [ct.c_uint64, sz, dp]
        L.oracle_set_qr_rank_tol.argtypes = [ct.c_double]
        L.oracle_num_threads.restype = ct.c_int
        L.oracle_set_threads.argtypes = [ct.c_int]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def gate_array(gates):
    arr = (OGate * max(1, len(gates)))()
    for i, g in enumerate(gates):
        arr[i].kind, arr[i].q0, arr[i].q1, arr[i].q2, arr[i].angle = (
            int(g[0]), int(g[1]), int(g[2]), int(g[3]), float(g[4]))
    return arr


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle status {code}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise OracleError(rc, lib().oracle_last_error().decode())


class OracleMps:
    """The oracle's MpsState (src/tn/mps.rs:55-781)."""

    def __init__(self, n: int, chi: int, reject_truncation: bool = False, record: bool = False):
        L = lib()
        self._p = L.oracle_new(n, chi)
        if not self._p:
            raise OracleError(3, L.oracle_last_error().decode())
        self.n = n
        if reject_truncation:
            L.oracle_set_reject_truncation(self._p, 1)
        if record:
            L.oracle_set_record(self._p, 1)

    def __del__(self):
        if getattr(self, "_p", None):
            lib().oracle_free(self._p)
            self._p = None

    def apply_gate(self, g):
        arr = gate_array([g])
        _check(lib().oracle_apply_gate(self._p, arr))

    def apply_gates(self, gates):
        arr = gate_array(gates)
        _check(lib().oracle_apply_gates(self._p, arr, len(gates)))

    def apply_circuit(self, gates):
        arr = gate_array(gates)
        _check(lib().oracle_apply_circuit(self._p, arr, len(gates)))

    def apply_adjacent_disjoint_gates(self, gates):
        arr = gate_array(gates)
        _check(lib().oracle_apply_disjoint(self._p, arr, len(gates)))

    def apply_pair_at(self, site, kind, ctrl_on_left=True):
        _check(lib().oracle_apply_pair_at(self._p, site, kind, 1 if ctrl_on_left else 0))

    def amplitude_at_compact_zero(self) -> complex:
        out = np.zeros(2)
        lib().oracle_amplitude_zero(self._p, _dp(out))
        return complex(out[0], out[1])

    def bond_dims(self):
        out = (ct.c_size_t * self.n)()
        lib().oracle_bond_dims(self._p, out, self.n)
        return [int(x) for x in out]

    def site(self, i):
        l, r = ct.c_size_t(), ct.c_size_t()
        lib().oracle_get_site(self._p, i, None, 0, ct.byref(l), ct.byref(r))
        out = np.zeros(l.value * 2 * r.value * 2)
        _check(lib().oracle_get_site(self._p, i, _dp(out), out.size // 2, ct.byref(l), ct.byref(r)))
        return out.view(np.complex128).reshape(l.value, 2, r.value)

    def set_site(self, i, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        _check(lib().oracle_set_site(self._p, i, _dp(a.view(np.float64)), a.shape[0], a.shape[2]))

    def contract_to_statevector(self):
        out = np.zeros(2 ** (self.n + 1)) if self.n <= 20 else np.zeros(2)
        _check(lib().oracle_statevector(self._p, _dp(out), out.size // 2))
        return out.view(np.complex128)

    def to_dense_state(self):
        out = np.zeros(2 ** (self.n + 1))
        _check(lib().oracle_dense_state(self._p, _dp(out), out.size // 2))
        return out.view(np.complex128)

    def site_amplitudes(self, t):
        out = np.zeros(4)
        _check(lib().oracle_site_amplitudes(self._p, t, _dp(out)))
        return complex(out[0], out[1]), complex(out[2], out[3])

    def control_probability(self, c):
        _, a1 = self.site_amplitudes(c)
        return a1.real * a1.real + a1.imag * a1.imag

    def counters(self):
        out = (ct.c_uint64 * 6)()
        lib().oracle_counters(self._p, out, 6)
        keys = ("units", "exact", "rsvd_kept", "rsvd_fallbacks", "svd_failures", "truncated")
        return dict(zip(keys, [int(x) for x in out]))

    def max_discarded(self):
        return lib().oracle_max_discarded(self._p)

    def split_log(self):
        m = lib().oracle_split_log(self._p, None, None, 0)
        paths = (ct.c_int32 * max(1, m))()
        chis = (ct.c_size_t * max(1, m))()
        lib().oracle_split_log(self._p, paths, chis, m)
        return [int(x) for x in paths[:m]], [int(x) for x in chis[:m]]


def split_theta(theta: np.ndarray, max_bond: int, reject: bool = False):
    """split_theta on theta of shape (dl, 2, 2, dr) -> (left, right, path, discarded)."""
    th = np.ascontiguousarray(theta, dtype=np.complex128)
    dl, dr = th.shape[0], th.shape[3]
    k = min(2 * dl, 2 * dr)
    left = np.zeros(dl * 2 * k, dtype=np.complex128)
    right = np.zeros(k * 2 * dr, dtype=np.complex128)
    chi = ct.c_size_t()
    path = ct.c_int()
    disc = ct.c_double()
    _check(lib().oracle_split_theta(_dp(th.view(np.float64)), dl, dr, max_bond, 1 if reject else 0,
                                    _dp(left.view(np.float64)), _dp(right.view(np.float64)),
                                    ct.byref(chi), ct.byref(path), ct.byref(disc)))
    c = chi.value
    return (left[: dl * 2 * c].reshape(dl, 2, c), right[: c * 2 * dr].reshape(c, 2, dr),
            path.value, disc.value)


def rsvd_sites(mat: np.ndarray, ell: int, dl: int, dr: int, max_bond: int):
    m = np.ascontiguousarray(mat, dtype=np.complex128)
    rows, cols = m.shape
    left = np.zeros(dl * 2 * ell, dtype=np.complex128)
    right = np.zeros(ell * 2 * dr, dtype=np.complex128)
    chi = ct.c_size_t()
    ok = lib().oracle_rsvd_sites(_dp(m.view(np.float64)), rows, cols, ell, dl, dr, max_bond,
                                 _dp(left.view(np.float64)), _dp(right.view(np.float64)),
                                 ct.byref(chi))
    if not ok:
        return None
    c = chi.value
    return left[: dl * 2 * c].reshape(dl, 2, c), right[: c * 2 * dr].reshape(c, 2, dr)


def checked_svd(a: np.ndarray):
    m = np.ascontiguousarray(a, dtype=np.complex128)
    rows, cols = m.shape
    k = min(rows, cols)
    u = np.zeros((rows, k), dtype=np.complex128)
    s = np.zeros(k)
    vt = np.zeros((k, cols), dtype=np.complex128)
    _check(lib().oracle_checked_svd(_dp(m.view(np.float64)), rows, cols, _dp(u.view(np.float64)),
                                    _dp(s), _dp(vt.view(np.float64))))
    return u, s, vt


def retained_chi(values, max_bond, rows, cols):
    v = np.ascontiguousarray(values, dtype=np.float64)
    return int(lib().oracle_retained_chi(_dp(v), v.size, max_bond, rows, cols))


def needs_f64_fallback(values, chi):
    v = np.ascontiguousarray(values, dtype=np.float64)
    return bool(lib().oracle_needs_f64_fallback(_dp(v), v.size, chi))


def mix_seed(rows, cols, ell):
    return int(lib().oracle_mix_seed(rows, cols, ell))


def rademacher(cols, ell, salt):
    out = np.zeros(cols * ell * 2)
    lib().oracle_rademacher(cols, ell, salt, _dp(out))
    return out.view(np.complex128).reshape(cols, ell)


def thin_q(y: np.ndarray):
    m = np.ascontiguousarray(y, dtype=np.complex128)
    rows, ell = m.shape
    q = np.zeros((rows, ell), dtype=np.complex128)
    lib().oracle_thin_q(_dp(m.view(np.float64)), rows, ell, _dp(q.view(np.float64)))
    return q


def thin_q_reflectors(y: np.ndarray, tau: float):
    """thin_q at rank threshold tau: (Q, V, mask), V's column i the unit reflector of column i
    (zero above row i; all zero where mask[i] == 0, the columns that took no reflector)."""
    m = np.ascontiguousarray(y, dtype=np.complex128)
    rows, ell = m.shape
    q = np.zeros((rows, ell), dtype=np.complex128)
    v = np.zeros((rows, ell), dtype=np.complex128)
    mask = (ct.c_int * ell)()
    lib().oracle_thin_q_reflectors(_dp(m.view(np.float64)), rows, ell, float(tau), _dp(q.view(np.float64)),
                                   _dp(v.view(np.float64)), mask)
    return q, v, np.array(list(mask), dtype=bool)


def pauli_expectation(sv: np.ndarray, label: str) -> complex:
    v = np.ascontiguousarray(sv, dtype=np.complex128)
    n = len(label)
    out = np.zeros(2)
    _check(lib().oracle_pauli_expectation(_dp(v.view(np.float64)), n, label.encode(), _dp(out)))
    return complex(out[0], out[1])


def sample_counts(sv: np.ndarray, n: int, measures, cbits: int, shots: int, seed: int):
    """Counts keyed by Qiskit bitstrings (cbit 0 rightmost), as src/sample.rs returns."""
    v = np.ascontiguousarray(sv, dtype=np.complex128)
    mq = (ct.c_int32 * max(1, len(measures)))(*[m[0] for m in measures])
    mc = (ct.c_int32 * max(1, len(measures)))(*[m[1] for m in measures])
    counts = (ct.c_uint64 * (1 << cbits))()
    _check(lib().oracle_sample_counts(_dp(v.view(np.float64)), n, mq, mc, len(measures), cbits,
                                      shots, seed, counts))
    return {format(k, f"0{cbits}b"): int(c) for k, c in enumerate(counts) if c}


def chacha_block(key_words, counter, nonce, rounds):
    key = (ct.c_uint32 * 8)(*key_words)
    out = (ct.c_uint32 * 16)()
    lib().oracle_chacha_block(key, counter, nonce, rounds, out)
    return list(out)


def stdrng_f64(seed, count):
    out = np.zeros(count)
    lib().oracle_stdrng_f64(seed, count, _dp(out))
    return out


def set_qr_rank_tol(tau: float):
    """0.0: literal floating-point range finder; tau > 0: exact-arithmetic rank rule."""
    lib().oracle_set_qr_rank_tol(float(tau))


def set_threads(n: int):
    lib().oracle_set_threads(n)


def num_threads() -> int:
    return int(lib().oracle_num_threads())
