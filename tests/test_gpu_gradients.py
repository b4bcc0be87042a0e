"""Adjoint gradients (wqc_tn_backward) of expectation-mode observables with respect to the
RX/RY/RZ angles, through the C ABI. The sweep is reverse mode over the site tensors of the
forward pass, truncated splits included (the adjoint of the truncated SVD with the sqrt(S)
gauge), so it is the derivative of the truncated forward map itself.

Bars:
- untruncated circuits (n <= 8, chi above the exact bond): against the dense-statevector
  parameter shift (exact for exp(-i theta P / 2) gates), 1e-10;
- config 5 at chi = 256 (no truncation occurs): against the GPU forward's parameter shift on a
  spread of 8 angles, 1e-8;
- truncating random circuits (long-range CNOTs routed by SWAP units, H/S/T between the
  rotations, chi 4 to 8): against central finite differences of the GPU forward (h = 1e-5) on
  every angle, 1e-6;
- config 5 at chi = 64 (38 truncating splits): against central finite differences (h = 1e-4,
  FP64) on 32 of the 512 angles, 1e-6. The parameter shift is not the derivative of a truncated
  map (it is exact only for a + b cos + c sin in each angle) and differs from both by about 2e-4
  there; it is printed, not asserted.
"""

import math
import sys
import time

import numpy as np
import pytest

import circuits as C
import dense_ref as D

pytestmark = pytest.mark.gpu

PAULI = {
    "I": np.eye(2, dtype=complex),
    "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=complex),
    "Z": np.array([[1, 0], [0, -1]], dtype=complex),
}


@pytest.fixture(scope="module")
def wt():
    import wqc_tn

    wqc_tn.load()
    return wqc_tn


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dense_energy(n, gates, terms):
    psi = D.run(n, gates)
    e = 0.0
    for lab, c in terms:
        v = psi.copy()
        for q in range(n):
            ch = lab[n - 1 - q]
            if ch != "I":
                v = D._apply_1q(v, n, q, PAULI[ch])
        e += c * np.vdot(psi, v).real
    return e


def shifted(gates, i, delta):
    g = list(gates)
    k, q0, q1, q2, a = g[i]
    g[i] = (k, q0, q1, q2, a + delta)
    return g


def gpu_energy(wt, n, chi, gates, terms):
    st = wt.MpsState(n, chi)
    st.apply_circuit(gates)
    v = st.pauli_expectations([lab for lab, _ in terms])
    st.free()
    return float(sum(c * x for (_, c), x in zip(terms, v)))


def random_rotation_circuit(rng, n, depth):
    gates = []
    for _ in range(depth):
        for q in range(n):
            kind = (C.RX, C.RY, C.RZ)[rng.below(3)]
            gates.append(C.one(kind, q, 2 * math.pi * rng.uniform()))
            if rng.below(3) == 0:
                gates.append(C.one((C.H, C.S, C.T)[rng.below(3)], q))
        for q in range(rng.below(2), n - 1, 2):
            gates.append(C.two((C.CZ, C.CNOT)[rng.below(2)], q, q + 1))
        if n > 3 and rng.below(2):
            a, b = rng.below(n), rng.below(n)
            if a != b:
                gates.append(C.two(C.CNOT, a, b))
    return gates


def random_terms(rng, n, count):
    terms = []
    for _ in range(count):
        lab = ["I"] * n
        for _ in range(1 + rng.below(3)):
            lab[rng.below(n)] = "XYZ"[rng.below(3)]
        terms.append(("".join(lab), 2 * rng.uniform() - 1))
    return terms


@pytest.mark.parametrize("n,depth,every", [(4, 4, 1), (6, 5, 2), (8, 4, 3)])
def test_adjoint_matches_dense_parameter_shift(wt, n, depth, every):
    rng = C.SplitMix64(0xAD + 17 * n + depth)
    gates = random_rotation_circuit(rng, n, depth)
    terms = random_terms(rng, n, 6)
    val, grad = wt.backward(n, 64, gates, terms, checkpoint_every=every)
    assert abs(val.real - dense_energy(n, gates, terms)) <= 1e-10
    worst = 0.0
    for i, g in enumerate(gates):
        if g[0] not in (C.RX, C.RY, C.RZ):
            assert grad[i] == 0.0
            continue
        ps = 0.5 * (dense_energy(n, shifted(gates, i, math.pi / 2), terms)
                    - dense_energy(n, shifted(gates, i, -math.pi / 2), terms))
        worst = max(worst, abs(grad[i] - ps))
        assert abs(grad[i] - ps) <= 1e-10, (i, grad[i], ps)
    say(f"n={n}: worst |adjoint - dense shift| = {worst:.2e}")


def test_config5_untruncated_vs_parameter_shift(wt):
    n, _, gates, _, tf = C.config5()
    terms = [(lab, c) for c, lab in tf]
    chi = 256
    val, grad = wt.backward(n, chi, gates, terms, checkpoint_every=2)
    e0 = gpu_energy(wt, n, chi, gates, terms)
    assert abs(val.real - e0) <= 1e-10
    rot = [i for i, g in enumerate(gates) if g[0] in (C.RY, C.RZ)]
    assert len(rot) == 512
    worst = 0.0
    for i in rot[:: len(rot) // 8]:
        ps = 0.5 * (gpu_energy(wt, n, chi, shifted(gates, i, math.pi / 2), terms)
                    - gpu_energy(wt, n, chi, shifted(gates, i, -math.pi / 2), terms))
        worst = max(worst, abs(grad[i] - ps))
        assert abs(grad[i] - ps) <= 1e-8, (i, grad[i], ps)
    say(f"config5 chi=256: worst |adjoint - shift| = {worst:.2e}")


@pytest.mark.parametrize("n,chi,depth,seed,every", [(8, 4, 6, 1, 2), (10, 6, 6, 2, 1), (12, 5, 8, 3, 3),
                                                    (12, 8, 12, 3, 0)])
def test_truncating_circuits_vs_finite_differences(wt, n, chi, depth, seed, every):
    rng = C.SplitMix64(0x7A + seed)
    gates = random_rotation_circuit(rng, n, depth)
    terms = random_terms(rng, n, 6)
    val, grad = wt.backward(n, chi, gates, terms, checkpoint_every=every)
    st = wt.MpsState(n, chi)
    st.apply_circuit(gates)
    trunc = st.stats()["truncated_splits"]
    st.free()
    assert trunc > 0
    assert abs(val.real - gpu_energy(wt, n, chi, gates, terms)) <= 1e-10
    h = 1e-5
    worst = 0.0
    for i, g in enumerate(gates):
        if g[0] not in (C.RX, C.RY, C.RZ):
            assert grad[i] == 0.0
            continue
        fd = (gpu_energy(wt, n, chi, shifted(gates, i, h), terms)
              - gpu_energy(wt, n, chi, shifted(gates, i, -h), terms)) / (2 * h)
        worst = max(worst, abs(grad[i] - fd))
        assert abs(grad[i] - fd) <= 1e-6, (i, grad[i], fd)
    say(f"n={n} chi={chi}: {trunc} truncating splits, worst |adjoint - FD| = {worst:.2e}")


def test_config5_chi64_vs_finite_differences_and_shift(wt):
    n, chi, gates, _, tf = C.config5()
    terms = [(lab, c) for c, lab in tf]
    t = time.time()
    val, grad = wt.backward(n, chi, gates, terms, checkpoint_every=2)
    say(f"config5 chi=64: backward (512 angles, checkpoints every 2 runs) {time.time() - t:.2f} s")
    e0 = gpu_energy(wt, n, chi, gates, terms)
    assert abs(val.real - e0) <= 1e-9
    rot = [i for i, g in enumerate(gates) if g[0] in (C.RY, C.RZ)]
    h = 1e-4
    worst_fd = worst_ps = worst_fd_ps = 0.0
    for i in rot[:: len(rot) // 32]:
        fd = (gpu_energy(wt, n, chi, shifted(gates, i, h), terms)
              - gpu_energy(wt, n, chi, shifted(gates, i, -h), terms)) / (2 * h)
        ps = 0.5 * (gpu_energy(wt, n, chi, shifted(gates, i, math.pi / 2), terms)
                    - gpu_energy(wt, n, chi, shifted(gates, i, -math.pi / 2), terms))
        worst_fd = max(worst_fd, abs(grad[i] - fd))
        worst_ps = max(worst_ps, abs(grad[i] - ps))
        worst_fd_ps = max(worst_fd_ps, abs(fd - ps))
    say(f"config5 chi=64: worst |adjoint - FD| = {worst_fd:.2e}, |adjoint - shift| = {worst_ps:.2e}, "
        f"|FD - shift| = {worst_fd_ps:.2e}")
    assert worst_fd <= 1e-6
