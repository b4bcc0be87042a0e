"""numpy dense statevector simulator: the independent check on the MPS oracle.

Restates the semantics of the reference's DenseTnState (src/tn/dense.rs:53-99) and its
gate tensors (src/tn/gates.rs:10-104): basis bit q is qubit q; CNOT(c, t) flips t when
c is 1; CZ puts -1 on |11>; CCNOT is the Toffoli. numpy is a library independent of the
oracle's C++ code, so agreement pins gate tensors, qubit order and routing.
"""

from __future__ import annotations

import math

import numpy as np

X, Y, Z, H, S, T, CNOT, CZ, CCNOT, RX, RY, RZ, MEASURE, RESET, IF = range(1, 16)


def unary(kind: int, angle: float = 0.0) -> np.ndarray:
    r2 = 1.0 / math.sqrt(2.0)
    if kind == H:
        return np.array([[r2, r2], [r2, -r2]], dtype=complex)
    if kind == X:
        return np.array([[0, 1], [1, 0]], dtype=complex)
    if kind == Y:
        return np.array([[0, -1j], [1j, 0]], dtype=complex)
    if kind == Z:
        return np.array([[1, 0], [0, -1]], dtype=complex)
    if kind == S:
        return np.array([[1, 0], [0, 1j]], dtype=complex)
    if kind == T:
        return np.array([[1, 0], [0, complex(r2, r2)]], dtype=complex)
    s, c = math.sin(angle / 2), math.cos(angle / 2)
    if kind == RX:
        return np.array([[c, -1j * s], [-1j * s, c]], dtype=complex)
    if kind == RY:
        return np.array([[c, -s], [s, c]], dtype=complex)
    if kind == RZ:
        return np.array([[complex(c, -s), 0], [0, complex(c, s)]], dtype=complex)
    raise ValueError(kind)


def _apply_1q(psi: np.ndarray, n: int, q: int, u: np.ndarray) -> np.ndarray:
    v = psi.reshape((2,) * n)  # axis k <-> bit (n-1-k)
    ax = n - 1 - q
    v = np.moveaxis(np.tensordot(u, v, axes=([1], [ax])), 0, ax)
    return v.reshape(-1)


def apply_gate(psi: np.ndarray, n: int, g) -> np.ndarray:
    kind = g[0]
    if kind in (MEASURE, RESET, IF):
        return psi
    if kind in (H, X, Y, Z, S, T, RX, RY, RZ):
        return _apply_1q(psi, n, g[1], unary(kind, g[4]))
    idx = np.arange(psi.size)
    if kind == CNOT:
        c, t = g[1], g[2]
        m = ((idx >> c) & 1) == 1
        out = psi.copy()
        out[idx[m]] = psi[idx[m] ^ (1 << t)]
        return out
    if kind == CZ:
        c, t = g[1], g[2]
        m = (((idx >> c) & 1) == 1) & (((idx >> t) & 1) == 1)
        out = psi.copy()
        out[m] = -out[m]
        return out
    if kind == CCNOT:
        c1, c2, t = g[1], g[2], g[3]
        m = (((idx >> c1) & 1) == 1) & (((idx >> c2) & 1) == 1)
        out = psi.copy()
        out[idx[m]] = psi[idx[m] ^ (1 << t)]
        return out
    raise ValueError(kind)


def run(n: int, gates) -> np.ndarray:
    psi = np.zeros(2 ** n, dtype=complex)
    psi[0] = 1.0
    for g in gates:
        psi = apply_gate(psi, n, g)
    return psi
