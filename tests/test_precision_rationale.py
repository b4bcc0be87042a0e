"""Why the mixed forms keep the theta merge in FP64 (CPU, numpy).

The reference truncates at sigma > 1e-12 (src/tn/svd_guard.rs:12, 25: SINGULAR_CUTOFF, retained_chi). Rounding theta
to complex64 before an exact FP64 SVD, the effect of any FP32-accurate theta GEMM (FP32 complex or
split-BF16 with FP32 accumulation), leaves noise singular values near 1e-7 |theta|, far above that
cutoff, so they are kept: the bond dimensions inflate (here about twice the FP64 total on a
24-qubit, chi 64 RX/RY/RZ + CZ brickwork) and the state departs from the reference's kept spectrum.
The mixed forms therefore lower only the Jacobi rounds, which the FP64 finish corrects."""

import numpy as np

import circuits as C


def _one(kind, t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    if kind == C.RX:
        return np.array([[c, -1j * s], [-1j * s, c]])
    if kind == C.RY:
        return np.array([[c, -s], [s, c]], complex)
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


def _bonds(n, chi, gates, round32):
    sites = [np.zeros((1, 2, 1), complex) for _ in range(n)]
    for s in sites:
        s[0, 0, 0] = 1
    for kind, a, b, _, angle in gates:
        if kind == C.CZ:
            p = min(a, b)
            dl, dr = sites[p].shape[0], sites[p + 1].shape[2]
            th = np.tensordot(sites[p], sites[p + 1], axes=(2, 0)).copy()
            th[:, 1, 1, :] *= -1
            th = th.reshape(2 * dl, 2 * dr)
            if round32:
                th = th.astype(np.complex64).astype(complex)
            u, s, vh = np.linalg.svd(th, full_matrices=False)
            k = max(1, min(int(np.sum(s > 1e-12)), chi))
            r = np.sqrt(s[:k])
            sites[p] = (u[:, :k] * r).reshape(dl, 2, k)
            sites[p + 1] = (r[:, None] * vh[:k]).reshape(k, 2, dr)
        else:
            sites[a] = np.einsum("st,atb->asb", _one(kind, angle), sites[a])
    return [s.shape[2] for s in sites[:-1]]


def test_fp32_theta_inflates_bonds():
    n, chi = 24, 64
    gates = C.rotation_cz_brickwork(n, 16, 11)
    b64 = _bonds(n, chi, gates, False)
    b32 = _bonds(n, chi, gates, True)
    assert sum(b32) >= 1.5 * sum(b64), (sum(b64), sum(b32))
    assert all(x >= y for x, y in zip(b32, b64))
