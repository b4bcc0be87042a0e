"""CPU pins of the adjoint sweep's formulas (tests/adjoint_ref.py) against central differences.

- one truncated split, through a scalar that depends on its factors only up to unitary bond
  rotations (as any later computation does): L = Re tr(G^H A A^H) + Re tr(H^H B^H B)
  + Re tr(K^H A B); wide, tall, rank-deficient (exact zero singular values, at most chi kept)
  and exactly degenerate kept values, with the Jacobi rows in a shuffled order;
- whole truncating circuits with long-range CNOTs (SWAP units), H/S/T gates and rotations,
  including one whose splits have exactly equal kept singular values.
A dropped term of the coefficient formulas, a swapped role of the factors for tall splits or a
wrong limit for equal values moves these by 1e-2 or more; the bars are 1e-6.
"""

import math

import numpy as np
import pytest

import adjoint_ref as A
import circuits as C


def _haar(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _scalar(M, chi, G, H, K):
    a, b, _ = A.factors(M, chi)
    return (np.trace(G.conj().T @ a @ a.conj().T) + np.trace(H.conj().T @ b.conj().T @ b)
            + np.trace(K.conj().T @ a @ b)).real


@pytest.mark.parametrize("rows,cols,chi,svals", [
    (8, 12, 3, None), (12, 8, 3, None), (9, 5, 5, None),
    (8, 8, 3, [0.9, 0.6, 0.5, 0.3, 0.2, 0, 0, 0]), (10, 6, 2, [0.9, 0.6, 0.5, 0, 0, 0]),
    (6, 8, 6, [0.9, 0.9, 0.5, 0.3, 0.3, 0.1]), (8, 6, 4, [0.7, 0.7, 0.7, 0.2, 0.1, 0.05]),
    (4, 2, 4, [0.78, 0.78])])
def test_split_adjoint_vs_central_differences(rows, cols, chi, svals):
    rng = np.random.default_rng(rows * 100 + cols * 10 + chi)
    p = min(rows, cols)
    if svals is None:
        M = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
    else:
        M = _haar(rng, rows)[:, :p] @ np.diag(svals) @ _haar(rng, cols)[:p, :]
    G, H = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)) for d in (rows, cols))
    K = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
    a, b, svd = A.factors(M, chi)
    abar = (G + G.conj().T) @ a + K @ b.conj().T
    bbar = b @ (H + H.conj().T) + a.conj().T @ K
    mbar = A.split_adjoint(svd, abar, bbar, row_order=rng.permutation(p))
    h = 1e-6
    worst = 0.0
    for _ in range(6):
        d = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
        fd = (_scalar(M + h * d, chi, G, H, K) - _scalar(M - h * d, chi, G, H, K)) / (2 * h)
        worst = max(worst, abs(fd - np.real(np.vdot(mbar, d))))
    assert worst <= 1e-6, worst


def _circuit(rng, n, depth):
    gates = []
    for _ in range(depth):
        for q in range(n):
            gates.append(C.one((C.RX, C.RY, C.RZ)[rng.below(3)], q, 2 * math.pi * rng.uniform()))
            if rng.below(3) == 0:
                gates.append(C.one((C.H, C.S, C.T)[rng.below(3)], q))
        for q in range(rng.below(2), n - 1, 2):
            gates.append(C.two((C.CZ, C.CNOT)[rng.below(2)], q, q + 1))
        if rng.below(2):
            a, b = rng.below(n), rng.below(n)
            if a != b:
                gates.append(C.two(C.CNOT, a, b))
    return gates


def _terms(rng, n, count):
    terms = []
    for _ in range(count):
        lab = ["I"] * n
        for _ in range(1 + rng.below(3)):
            lab[rng.below(n)] = "XYZ"[rng.below(3)]
        terms.append(("".join(lab), 2 * rng.uniform() - 1))
    return terms


@pytest.mark.parametrize("n,chi,depth,seed", [(6, 3, 5, 1), (7, 4, 5, 2), (8, 4, 6, 3)])
def test_circuit_gradient_vs_central_differences(n, chi, depth, seed):
    rng = C.SplitMix64(0xADD + seed)
    gates = _circuit(rng, n, depth)
    terms = _terms(rng, n, 5)
    _, grad = A.gradient(n, chi, gates, terms)
    truncated = sum(1 for rec in A.forward(n, chi, gates, record=True)[1]
                    if rec[0] == "u" and np.sum(rec[5][2] > 1e-12) > rec[5][3])
    assert truncated > 0
    h = 1e-5
    worst = 0.0
    for i, g in enumerate(gates):
        if g[0] not in A.GEN:
            assert grad[i] == 0.0
            continue
        gp, gm = list(gates), list(gates)
        gp[i] = g[:4] + (g[4] + h,)
        gm[i] = g[:4] + (g[4] - h,)
        fd = (A.energy(A.forward(n, chi, gp)[0], terms) - A.energy(A.forward(n, chi, gm)[0], terms)) / (2 * h)
        worst = max(worst, abs(fd - grad[i]))
    assert worst <= 1e-6, worst
