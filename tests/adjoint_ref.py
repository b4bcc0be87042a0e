"""numpy restatement of the adjoint sweep of wqc_tn_backward (test infrastructure).

The forward follows the engine's operations: one-qubit updates site <- U site, two-qubit gates
as units on adjacent sites (long-range gates bubbled with SWAP units, the 4x4 resolved for the
control's side), each unit a merge theta = L R, the gate, and the truncated split of
W = theta' (or theta'^H when tall) into U_k sqrt(S), sqrt(S) V_k^H with the reference's rule
(keep sigma > 1e-12, at most chi). The backward is reverse mode over the site tensors; a unit's
split is passed back with the same coefficient formulas as the device kernel k_svd_adj_coef
(the rotated rows X = J W of a Jacobi with J = U^H, the kept rows in the factors' column order,
pairs of equal singular values at their limit). Everything here is checked against central
differences, so the formulas are pinned independently of the device code.
"""

from __future__ import annotations

import numpy as np

import circuits as C
import dense_ref as D

NULL_ROWS_SQ = 1e-26
PAULI = {"X": np.array([[0, 1], [1, 0]], complex), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1.0 + 0j, -1.0])}
GEN = {C.RX: PAULI["X"], C.RY: PAULI["Y"], C.RZ: PAULI["Z"]}


def _m4(co, to, ci, ti):
    return (co * 2 + to) * 4 + ci * 2 + ti


def two_qubit(kind):
    u = np.zeros(16, complex)
    if kind == C.CNOT:
        for a, b in ((0, 0), (1, 1)):
            u[_m4(0, a, 0, b)] = 1
        u[_m4(1, 0, 1, 1)] = u[_m4(1, 1, 1, 0)] = 1
    elif kind == C.CZ:
        u[_m4(0, 0, 0, 0)] = u[_m4(0, 1, 0, 1)] = u[_m4(1, 0, 1, 0)] = 1
        u[_m4(1, 1, 1, 1)] = -1
    else:  # SWAP
        u[_m4(0, 0, 0, 0)] = u[_m4(0, 1, 1, 0)] = u[_m4(1, 0, 0, 1)] = u[_m4(1, 1, 1, 1)] = 1
    return u


def resolve(u, ctrl_first):
    out = np.zeros(16, complex)
    for so in range(2):
        for to in range(2):
            for si in range(2):
                for ti in range(2):
                    co, tco = (so, to) if ctrl_first else (to, so)
                    ci, tci = (si, ti) if ctrl_first else (ti, si)
                    out[(so * 2 + to) * 4 + si * 2 + ti] = u[_m4(co, tco, ci, tci)]
    return out.reshape(2, 2, 2, 2)


def operations(gates):
    ops = []
    swap = resolve(two_qubit(None), True)
    for gi, g in enumerate(gates):
        kind = g[0]
        if kind in (C.CZ, C.CNOT):
            ctrl, tgt = g[1], g[2]
            lo, hi = min(ctrl, tgt), max(ctrl, tgt)
            ops += [("u", pos, swap, gi) for pos in range(hi - 1, lo, -1)]
            ops.append(("u", lo, resolve(two_qubit(kind), ctrl < tgt), gi))
            ops += [("u", pos, swap, gi) for pos in range(lo + 1, hi)]
        elif kind not in (C.MEASURE, C.RESET, C.IF):
            ops.append(("1", g[1], D.unary(kind, g[4]), gi))
    return ops


def factors(M, chi):
    """(A, B, svd of W) with W = M or M^H, A = U_k sqrt(S) (or its role swapped when tall)."""
    tall = M.shape[0] > M.shape[1]
    W = M.conj().T if tall else M
    u, s, vh = np.linalg.svd(W, full_matrices=False)
    k = max(1, min(int(np.sum(s > 1e-12)), chi))
    r = np.sqrt(s[:k])
    aw, bw = u[:, :k] * r, r[:, None] * vh[:k]
    a, b = (bw.conj().T, aw.conj().T) if tall else (aw, bw)
    return a, b, (W, u, s, k, tall)


def coefficients(ju, xv, sig, inv, c):
    """C (p x p) and E (p x c) of k_svd_adj_coef; ju, xv unscaled (J Abar_W, X Bbar_W^H)."""
    p = len(sig)
    krow = np.full(c, -1)
    for row, i in enumerate(inv):
        if i >= 0:
            krow[i] = row
    rs = np.sqrt(sig[krow])
    JU = ju * rs[None, :]
    XV = xv * rs[None, :]
    Cm = np.zeros((p, p), complex)
    E = np.zeros((p, c), complex)
    floor2 = NULL_ROWS_SQ / p
    for x in range(p):
        ix = inv[x]
        if ix >= 0:
            E[x, ix] = 1 / np.sqrt(sig[x])
        for y in range(p):
            iy = inv[y]
            sx, sy = sig[x], sig[y]
            if x == y:
                if ix >= 0:
                    ju1, xv1 = JU[x, ix], XV[x, ix]
                    sbar = (ju1.real + xv1.real / sx) / (2 * sx)
                    pim = 0.5 * (ju1.imag - xv1.imag / sx) / sx
                    Cm[x, x] = complex(sbar / sx - xv1.real / sx ** 3, pim / sx + xv1.imag / sx ** 3)
            elif ix >= 0 and iy >= 0:
                jux, juy, xvx, xvy = JU[x, iy], JU[y, ix], XV[x, iy], XV[y, ix]
                jvx, jvy = xvx / sx, xvy / sy
                a, b = np.conj(jux) - juy, np.conj(jvx) - jvy
                dm, dp = sy - sx, sy + sx
                al = (a - b) / (2 * dp)
                if abs(dm) > 1e-8 * max(sx, sy):
                    al += (a + b) / (2 * dm)
                else:
                    al += (np.conj(jux) + jvy + juy + np.conj(jvx)) / (2 * dp)
                Cm[x, y] = np.conj(al) / sy - np.conj(xvy) / (sx * sy * sy)
            else:
                v = 0
                if iy >= 0:
                    d = (sy - sx) * (sy + sx)
                    if d != 0:
                        v += (JU[x, iy] * sy + XV[x, iy]) / (d * sy)
                if ix >= 0:
                    d = (sx - sy) * (sx + sy)
                    if d != 0:
                        v += np.conj(JU[y, ix]) / d
                        if sy * sy > floor2:
                            v += np.conj(XV[y, ix]) / (sx * d)
                Cm[x, y] = v
    return Cm, E


def split_adjoint(svd, abar, bbar, row_order=None):
    """Mbar of a truncated split from the cotangents of its two factors. row_order permutes the
    Jacobi rows (row r holds singular pair row_order[r]) to exercise the kept-row map."""
    W, u, s, k, tall = svd
    p = len(s)
    perm = np.arange(p) if row_order is None else np.asarray(row_order)
    J = u.conj().T[perm]
    X = J @ W
    sig = s[perm]
    inv = np.full(p, -1)
    for i in range(k):
        inv[int(np.flatnonzero(perm == i)[0])] = i
    aw, bw = (bbar.conj().T, abar.conj().T) if tall else (abar, bbar)
    Cm, E = coefficients(J @ aw, X @ bw.conj().T, sig, inv, k)
    wbar = J.conj().T @ (Cm @ X + E @ bw)
    return wbar.conj().T if tall else wbar


def forward(n, chi, gates, record=False):
    sites = [np.zeros((1, 2, 1), complex) for _ in range(n)]
    for s in sites:
        s[0, 0, 0] = 1
    tape = []
    for op in operations(gates):
        if op[0] == "u":
            _, p, u4, gi = op
            L, R = sites[p], sites[p + 1]
            dl, dr = L.shape[0], R.shape[2]
            th = np.einsum("stxy,axyb->astb", u4, np.einsum("asc,ctb->astb", L, R))
            a, b, svd = factors(th.reshape(2 * dl, 2 * dr), chi)
            if record:
                tape.append(("u", p, L, R, u4, svd))
            sites[p] = a.reshape(dl, 2, -1)
            sites[p + 1] = b.reshape(-1, 2, dr)
        else:
            _, q, u, gi = op
            if record:
                tape.append(("1", q, sites[q], u, gi))
            sites[q] = np.einsum("st,atb->asb", u, sites[q])
    return sites, tape


def energy(sites, terms, cotangents=False):
    """sum_k c_k <psi|P_k|psi> over (label, coeff) terms in Qiskit order; with cotangents, also
    2 dE/d conj(site) for every site."""
    n = len(sites)
    e = 0.0
    cot = [np.zeros_like(s) for s in sites]
    for lab, c in terms:
        ops = [PAULI.get(lab[n - 1 - q], np.eye(2, dtype=complex)) for q in range(n)]
        left = [np.ones((1, 1), complex)]
        for q in range(n):
            pa = np.einsum("st,atb->asb", ops[q], sites[q])
            left.append(np.einsum("xa,xsy,asb->yb", left[-1], sites[q].conj(), pa))
        e += c * left[-1][0, 0].real
        if not cotangents:
            continue
        right = [None] * (n + 1)
        right[n] = np.ones((1, 1), complex)
        for q in range(n - 1, -1, -1):
            pa = np.einsum("st,atb->asb", ops[q], sites[q])
            right[q] = np.einsum("xsy,asb,yb->xa", sites[q].conj(), pa, right[q + 1])
        for q in range(n):
            w = np.einsum("xa,xsy,st,yb->atb", left[q], sites[q].conj(), ops[q], right[q + 1]).conj()
            cot[q] += 2 * c * w
    return (e, cot) if cotangents else e


def gradient(n, chi, gates, terms):
    sites, tape = forward(n, chi, gates, record=True)
    e, cot = energy(sites, terms, cotangents=True)
    grad = np.zeros(len(gates))
    for rec in reversed(tape):
        if rec[0] == "1":
            _, q, a_in, u, gi = rec
            a_out = np.einsum("st,atb->asb", u, a_in)
            if gates[gi][0] in GEN:
                grad[gi] += np.real(np.vdot(cot[q], np.einsum("st,atb->asb", -0.5j * GEN[gates[gi][0]], a_out)))
            cot[q] = np.einsum("ts,atb->asb", u.conj(), cot[q])
        else:
            _, p, L, R, u4, svd = rec
            dl, dr, bond = L.shape[0], R.shape[2], L.shape[2]
            k = svd[3]
            mbar = split_adjoint(svd, cot[p].reshape(2 * dl, k), cot[p + 1].reshape(k, 2 * dr))
            tb = np.einsum("stxy,astb->axyb", u4.conj(), mbar.reshape(dl, 2, 2, dr)).reshape(2 * dl, 2 * dr)
            cot[p] = (tb @ R.reshape(bond, 2 * dr).conj().T).reshape(dl, 2, bond)
            cot[p + 1] = (L.reshape(2 * dl, bond).conj().T @ tb).reshape(bond, 2, dr)
    return e, grad
