"""Pins the oracle's rank-thresholded thin_q (oracle/mps_oracle.cpp, restating thin_q of
src/tn/mps.rs:1091-1101 inside rsvd_sites, src/tn/mps.rs:983-1053), the QR every golden
range-finder split is built on.

The reference's Omega has rank 48 at the saturated shapes (test_oracle.py::
test_reference_omega_has_rank_48), so Y = A (A^H (A Omega)) has 48 independent columns and the
other ell - 48 columns of Q are completion columns. In exact arithmetic a dependent column has
a zero Householder residual and takes no reflector, so a completion column is
H_p0 ... H_p47 e_i. The oracle's tau-mode (residual <= 1e-10 of the column norm counts as zero)
computes that. These checks would each fail on a plausible mistake:

* Q unitary and range(Y) inside range(Q): a lost or doubled reflector;
* Q rebuilt in numpy as H_p0 H_p1 ... H_p47 I[:, :ell] from the exported reflectors: the
  product taken in the wrong order;
* H_p47 ... H_p0 Y upper triangular with the diagonal entry at a reflector column equal to
  -(x0/|x0|) |x| (x the column's residual below the diagonal): the reflection axis
  x + (x0/|x0|)|x| e_0 of nalgebra's Householder (the opposite sign gives +(x0/|x0|)|x|);
* the reflector positions equal the independent columns of Omega, found by numpy ranks;
* an mpmath replay at 50 digits of the same Householder QR on a small, exactly rank-deficient
  complex-integer Y, where dependent residuals are truly zero (below 1e-40), matches the
  tau-mode oracle to 1e-12, while the literal floating-point oracle (tau = 0, reflectors
  from rounding noise) and a replay with the opposite reflection sign both differ from it.

Parity unpinned: nalgebra's own QR convention (reflection axis, and the phase it applies in
QR::q) is restated from its documented algorithm; nalgebra 0.33.3 is not in the image to run.
"""

import math

import mpmath
import numpy as np
import pytest

from oracle import oracle_mps as om
from test_oracle import py_rademacher


def householder_product(v, mask, ell):
    """H_p0 H_p1 ... H_pk applied to I[:, :ell] (H = I - 2 v v^H, v the exported columns)."""
    rows = v.shape[0]
    q = np.eye(rows, ell, dtype=complex)
    for i in reversed(np.flatnonzero(mask)):
        w = v[:, i:i + 1]
        q = q - 2.0 * w @ (w.conj().T @ q)
    return q


def apply_reflectors(v, mask, y):
    """H_pk ... H_p0 Y, the triangularized Y."""
    r = y.copy()
    for i in np.flatnonzero(mask):
        w = v[:, i:i + 1]
        r = r - 2.0 * w @ (w.conj().T @ r)
    return r


def independent_columns(mat, rel=1e-10):
    """Positions of the columns that raise the numerical rank, left to right."""
    pos, basis = [], np.zeros((mat.shape[0], 0), dtype=complex)
    for e in range(mat.shape[1]):
        cand = np.hstack([basis, mat[:, e:e + 1]])
        s = np.linalg.svd(cand, compute_uv=False)
        if s[-1] > rel * s[0]:
            pos.append(e)
            basis = cand
    return pos


@pytest.mark.parametrize("rows,cols,ell,decay", [(256, 256, 144, 0.03), (320, 200, 112, 0.05)])
def test_tau_mode_thin_q_on_reference_sketch(rows, cols, ell, decay):
    rng = np.random.default_rng(rows + ell)
    u = np.linalg.qr(rng.standard_normal((rows, rows)) + 1j * rng.standard_normal((rows, rows)))[0]
    w = np.linalg.qr(rng.standard_normal((cols, cols)) + 1j * rng.standard_normal((cols, cols)))[0]
    k = min(rows, cols)
    a = (u[:, :k] * np.exp(-decay * np.arange(k))) @ w[:, :k].conj().T
    omega = py_rademacher(cols, ell, om.mix_seed(rows, cols, ell))
    y = a @ (a.conj().T @ (a @ omega))
    q, v, mask = om.thin_q_reflectors(y, 1e-10)
    ny = np.linalg.norm(y)
    # Reflectors exactly at Omega's independent columns (48 of them at these shapes).
    pos = independent_columns(omega)
    assert len(pos) == 48
    assert list(np.flatnonzero(mask)) == pos
    # Unit reflectors, Q unitary, range(Y) inside range(Q).
    assert np.allclose(np.linalg.norm(v[:, mask], axis=0), 1.0, atol=1e-14)
    assert np.max(np.abs(q.conj().T @ q - np.eye(ell))) <= 1e-13
    assert np.linalg.norm(y - q @ (q.conj().T @ y)) <= 1e-12 * ny
    # Q is the product of the reflectors in order, completion columns included.
    assert np.max(np.abs(householder_product(v, mask, ell) - q)) <= 1e-13
    # The reflectors triangularize Y with nalgebra's sign: R[p, p] = -(x0/|x0|)|x|.
    r = apply_reflectors(v, mask, y)
    assert np.max(np.abs(np.tril(r, -1))) <= 1e-12 * ny
    part = y.copy()
    for i in range(ell):
        x = part[i:, i]
        if mask[i]:
            ph = x[0] / abs(x[0])
            want = -ph * np.linalg.norm(x)
            assert abs(r[i, i] - want) <= 1e-10 * abs(want), i
            wv = v[i:, i:i + 1]
            part[i:, :] = part[i:, :] - 2.0 * wv @ (wv.conj().T @ part[i:, :])
        else:
            # a skipped column's residual is below tau of its norm
            assert np.linalg.norm(x) <= 1e-10 * np.linalg.norm(y[:, i])


def mp_thin_q(y, flip=False, dps=50):
    """Householder QR of y at dps digits with the oracle's rule in exact arithmetic: a column
    whose residual is zero (below 1e-40 of its norm) takes no reflector; the reflection axis is
    x + (x0/|x0|)|x| e_0 (flip: x - (x0/|x0|)|x| e_0). Returns Q (rows x ell) as complex128."""
    with mpmath.workdps(dps):
        rows, ell = y.shape
        a = [[mpmath.mpc(complex(y[r, c])) for c in range(ell)] for r in range(rows)]
        cn = [mpmath.sqrt(sum(abs(a[r][c]) ** 2 for r in range(rows))) for c in range(ell)]
        vs = []
        for i in range(min(ell, rows)):
            alpha = mpmath.sqrt(sum(abs(a[r][i]) ** 2 for r in range(i, rows)))
            if alpha <= mpmath.mpf("1e-40") * cn[i]:
                vs.append(None)
                continue
            x0 = a[i][i]
            ph = x0 / abs(x0) if abs(x0) > 0 else mpmath.mpc(1)
            v = [a[r][i] for r in range(i, rows)]
            v[0] = v[0] - ph * alpha if flip else v[0] + ph * alpha
            vn = mpmath.sqrt(sum(abs(e) ** 2 for e in v))
            v = [e / vn for e in v]
            for c in range(i, ell):
                dot = 2 * sum(mpmath.conj(v[r - i]) * a[r][c] for r in range(i, rows))
                for r in range(i, rows):
                    a[r][c] -= v[r - i] * dot
            vs.append(v)
        q = [[mpmath.mpc(1 if r == c else 0) for c in range(ell)] for r in range(rows)]
        for i in reversed(range(len(vs))):
            v = vs[i]
            if v is None:
                continue
            for c in range(ell):
                dot = 2 * sum(mpmath.conj(v[r - i]) * q[r][c] for r in range(i, rows))
                for r in range(i, rows):
                    q[r][c] -= v[r - i] * dot
        return np.array([[complex(q[r][c]) for c in range(ell)] for r in range(rows)])


def test_tau_mode_matches_exact_arithmetic_replay():
    rng = np.random.default_rng(2024)
    rows, ell = 12, 8
    base = rng.integers(-4, 5, size=(rows, 3)) + 1j * rng.integers(-4, 5, size=(rows, 3))
    # columns 0, 1, 3, 6 independent; 2, 4, 5, 7 exact integer combinations of earlier ones
    extra = rng.integers(-4, 5, size=(rows, 1)) + 1j * rng.integers(-4, 5, size=(rows, 1))
    y = np.hstack([
        base[:, :2], base[:, :1] * (2 - 1j) + base[:, 1:2] * 3, base[:, 2:3],
        base[:, 2:3] * (1 + 1j) - base[:, :1], base[:, 1:2] * 2, extra,
        extra - base[:, 2:3] * (3 + 2j)]).astype(complex)
    assert np.linalg.matrix_rank(y) == 4
    exact = mp_thin_q(y)
    q, _, mask = om.thin_q_reflectors(y, 1e-10)
    assert list(np.flatnonzero(mask)) == [0, 1, 3, 6]
    assert np.max(np.abs(q - exact)) <= 1e-12
    # The checks discriminate: reflectors from rounding noise (tau = 0) and the opposite
    # reflection sign both give different completion columns.
    q_literal, _, mask_literal = om.thin_q_reflectors(y, 0.0)
    assert mask_literal.sum() > 4
    assert np.max(np.abs(q_literal - exact)) > 1e-6
    assert np.max(np.abs(mp_thin_q(y, flip=True) - q)) > 1e-6
    # and the independent columns' span agrees whatever the completion
    assert np.linalg.norm(y - q @ (q.conj().T @ y)) <= 1e-12 * np.linalg.norm(y)


def test_tau_mode_rsvd_matches_exact_replay():
    """The whole range finder of the oracle in tau-mode against a numpy pipeline that takes its
    Q from the 50-digit replay: B = Q^H A, LAPACK SVD of B, the reference's truncation and
    fallback rules. A is a complex-integer 80 x 64 matrix and Omega's scale is 1/sqrt(64) = 1/8,
    so Y = A (A^H (A Omega)) is exact in FP64 and exactly rank deficient (ell = chi + 16 = 32,
    Omega of rank 24: 8 of B's rows come from completion columns); the gauge-invariant
    products agree to rounding."""
    rows, cols, chi = 80, 64, 16
    ell = chi + 16
    rng = np.random.default_rng(77)
    a = (rng.integers(-3, 4, size=(rows, cols)) + 1j * rng.integers(-3, 4, size=(rows, cols))).astype(complex)
    omega = py_rademacher(cols, ell, om.mix_seed(rows, cols, ell))
    assert np.linalg.matrix_rank(omega, tol=1e-10 * np.linalg.norm(omega, 2)) == 24
    y = a @ (a.conj().T @ (a @ omega))
    assert np.array_equal(y * 8, np.round(y.real * 8) + 1j * np.round(y.imag * 8))  # exact
    q_exact = mp_thin_q(y)
    b = q_exact.conj().T @ a
    ub, s, vh = np.linalg.svd(b, full_matrices=False)
    vals = np.where(s < 1e-14, 0.0, s)
    k = om.retained_chi(vals, chi, rows, cols)
    assert k == chi and not om.needs_f64_fallback(vals, k)
    want = (q_exact @ ub[:, :k] * s[:k]) @ vh[:k]
    om.set_qr_rank_tol(1e-10)
    try:
        res = om.rsvd_sites(a, ell, rows // 2, cols // 2, chi)
    finally:
        om.set_qr_rank_tol(0.0)
    assert res is not None
    left, right = res
    assert left.shape[2] == k
    got = left.reshape(rows, k) @ right.reshape(k, cols)
    assert np.max(np.abs(got - want)) <= 1e-12 * s[0]
    # The literal floating-point oracle completes Q from rounding noise: a different product.
    lit = om.rsvd_sites(a, ell, rows // 2, cols // 2, chi)
    if lit is not None:
        gl = lit[0].reshape(rows, -1) @ lit[1].reshape(-1, cols)
        assert np.max(np.abs(gl - want)) > 1e-6 * s[0]
