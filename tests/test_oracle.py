"""Pins the CPU FP64 oracle (oracle/mps_oracle.cpp) before it is trusted as the checker.

Each oracle function and what pins it:

* gate tensors, 1q update, merge, U contraction, SWAP routing, CCNOT decomposition:
  numpy dense statevector (tests/dense_ref.py) on seeded circuits of every gate type,
  the reference's own MPS == dense cases (src/tn/mps.rs:1482-1540, 1922-1994, 2069-2113);
* exact split (truncated_svd_split): LAPACK (numpy.linalg.svd) singular values and the
  gauge-invariant product left.right; the nalgebra-failure KAT (src/tn/mps.rs:1999-2028);
* Jacobi SVD: tall, wide, rank-deficient rebuild <= 1e-12 (src/tn/mps.rs:2032-2065);
* truncation rule: the reference's vectors (src/tn/svd_guard.rs:289-342) and edge cases;
* range finder (rsvd_sites): the gapped rank-4 KAT (src/tn/mps.rs:1795-1854), Omega against
  an independent Python restatement of the integer LCG, and the whole pipeline against a
  numpy/LAPACK restatement (gauge-invariant product to 1e-10);
* thin_q: orthonormality and Q.R = Y with R upper triangular (LAPACK-free closed checks);
* amplitude_at_compact_zero: dense amplitude; H on 50/100 qubits = 1/sqrt(2)
  (src/tn/contract.rs:216-230); the 20q chi=128 circuit of src/tn/contract.rs:172-186;
* contract_to_statevector: dense state; the n > 20 refusal (src/tn/mps.rs:1864-1868);
* site_amplitudes / control_probability: brute force over the dense state (sum over
  all other qubits of the amplitude, no conjugate, as the all-ones environments give);
* sampler: KAT {"0":118,"1":138} (src/sample.rs:276-290), RFC 8439 ChaCha20 block
  vector (same block function at 20 rounds), seeded reproducibility;
* Pauli expectation: the reference's convention KATs (src/expectation.rs:291-466).

Parity unpinned (only an import or build of the reference could pin them): the exact
phase of each singular vector nalgebra returns (gauge; does not reach amplitudes), and
rand_chacha's stream beyond the one KAT and the block-function vector.
"""

import math

import numpy as np
import pytest

import circuits as C
import dense_ref as D
from oracle import oracle_mps as om


def seeded_random_circuit(rng, n, depth):
    gates = []
    for _ in range(depth):
        k = rng.uniform()
        q = rng.below(n)

        def other(excl):
            while True:
                w = rng.below(n)
                if w not in excl:
                    return w

        if n >= 3 and k < 0.08:
            c2 = other([q])
            t = other([q, c2])
            gates.append((C.CCNOT, q, c2, t, 0.0))
        elif n >= 2 and k < 0.35:
            t = other([q])
            gates.append(C.two(C.CNOT if rng.below(2) else C.CZ, q, t))
        elif k < 0.65:
            theta = -7.0 + 14.0 * rng.uniform()
            gates.append(C.one((C.RX, C.RY, C.RZ)[rng.below(3)], q, theta))
        else:
            gates.append(C.one((C.H, C.X, C.Y, C.Z, C.S, C.T)[rng.below(6)], q))
    return gates


def test_random_circuits_match_dense():
    # The reference's random-circuit case (src/tn/mps.rs:1922-1972), own seeded stream.
    rng = C.SplitMix64(0x5EED_2026_1128)
    for case in range(300):
        n = 1 + rng.below(6)
        depth = 1 + rng.below(24)
        gates = seeded_random_circuit(rng, n, depth)
        m = om.OracleMps(n, 128)
        m.apply_gates(gates)
        got = m.contract_to_statevector()
        want = D.run(n, gates)
        d = np.max(np.abs(got - want))
        assert d < 1e-9, (case, n, gates, d)


def test_ccnot_every_order():
    # src/tn/mps.rs:1976-1994
    def spread(n, phase):
        g = []
        for q in range(n):
            g.append(C.one(C.RY, q, 0.3 + 0.4 * q + phase))
            g.append(C.one(C.RZ, q, 0.7 - 0.2 * q + phase))
        return g

    for n in range(3, 6):
        for c1 in range(n):
            for c2 in range(n):
                for t in range(n):
                    if len({c1, c2, t}) < 3:
                        continue
                    gates = spread(n, 0.1) + [(C.CCNOT, c1, c2, t, 0.0)] + spread(n, -0.25)
                    m = om.OracleMps(n, 128)
                    m.apply_gates(gates)
                    d = np.max(np.abs(m.contract_to_statevector() - D.run(n, gates)))
                    assert d < 1e-10, (n, c1, c2, t, d)


def test_nonadjacent_and_rank_one_zero():
    gates = [C.one(C.H, 0), C.one(C.H, 1), C.two(C.CNOT, 0, 2), C.two(C.CZ, 3, 1),
             C.two(C.CNOT, 2, 0)]
    m = om.OracleMps(4, 8)
    m.apply_gates(gates)
    assert np.max(np.abs(m.contract_to_statevector() - D.run(4, gates))) < 1e-8
    # src/tn/mps.rs:2102-2113: qubit 0 stays |1>, <00|psi> is an exact zero.
    gates = [C.one(C.X, 0), C.two(C.CZ, 1, 0), C.one(C.RX, 1, -0.736), C.two(C.CZ, 1, 0),
             C.two(C.CNOT, 0, 1), C.two(C.CZ, 1, 0)]
    m = om.OracleMps(2, 128)
    m.apply_gates(gates)
    assert np.max(np.abs(m.contract_to_statevector() - D.run(2, gates))) < 1e-12


@pytest.mark.parametrize("n,depth", [(6, 6), (8, 8)])
def test_brickwork_matches_dense(n, depth):
    # src/tn/mps.rs:2069-2099
    gates = []
    for layer in range(depth):
        for q in range(n):
            gates.append(C.one(C.RY, q, 0.17 + 0.03 * (q + layer)))
            gates.append(C.one(C.RZ, q, 0.4 - 0.05 * q))
        for q in range(layer % 2, n - 1, 2):
            gates.append(C.two(C.CNOT, q, q + 1))
    m = om.OracleMps(n, 128)
    m.apply_gates(gates)
    assert np.max(np.abs(m.contract_to_statevector() - D.run(n, gates))) < 1e-9
    assert m.counters()["svd_failures"] == 0


def test_disjoint_layer_matches_sequential():
    # src/tn/mps.rs:1772-1793
    prep = [C.one(C.H, 0), C.one(C.H, 2), C.one(C.RY, 1, 0.3), C.one(C.RY, 3, 0.7)]
    layer = [C.two(C.CNOT, 0, 1), C.two(C.CZ, 3, 2)]
    seq = om.OracleMps(4, 8)
    seq.apply_gates(prep + layer)
    bat = om.OracleMps(4, 8)
    bat.apply_gates(prep)
    bat.apply_adjacent_disjoint_gates(layer)
    assert np.array_equal(seq.contract_to_statevector(), bat.contract_to_statevector())


def test_product_amplitudes_50_100():
    # src/tn/contract.rs:216-230
    for n in (50, 100):
        m = om.OracleMps(n, 128)
        m.apply_circuit([C.one(C.H, 0)])
        a = m.amplitude_at_compact_zero()
        assert abs(a.real - 1 / math.sqrt(2)) < 1e-8 and abs(a.imag) < 1e-8


def test_doc_bond128_circuit_matches_dense():
    # The 20-qubit chi=128 circuit of src/tn/contract.rs:172-186 (8 layers of
    # RY(1.1 + 0.05 q + 0.07 layer) and a CNOT brickwork). doc/tn-engine.md:58 quotes the
    # snapped key (16861, 0) for "the same circuit", but the dense statevector of this
    # generator gives <0|psi> = -0.010626... (key (-10626, 0)); the doc's value belongs to
    # an earlier generator. The dense amplitude pins the oracle here.
    n, layers = 20, 8
    gates = []
    for layer in range(layers):
        for q in range(n):
            gates.append(C.one(C.RY, q, 1.1 + 0.05 * q + 0.07 * layer))
        for q in range(layer % 2, n - 1, 2):
            gates.append(C.two(C.CNOT, q, q + 1))
    m = om.OracleMps(n, 128)
    m.apply_circuit(gates)
    a = m.amplitude_at_compact_zero()
    want = D.run(n, gates)[0]
    assert abs(a - want) <= 1e-10 * abs(want)
    assert (round(a.real / 1e-6), round(a.imag / 1e-6)) == (-10626, 0)
    assert max(m.bond_dims()) <= 16


def test_statevector_refuses_above_20():
    m = om.OracleMps(21, 4)
    with pytest.raises(om.OracleError) as e:
        m.contract_to_statevector()
    assert "qubit_count <= 20" in str(e.value)


def test_site_amplitudes_brute_force():
    rng = C.SplitMix64(99)
    n = 6
    gates = C.brickwork(n, 5, 1234)
    m = om.OracleMps(n, 64)
    m.apply_gates(gates)
    psi = D.run(n, gates)
    for t in range(n):
        a0, a1 = m.site_amplitudes(t)
        idx = np.arange(2 ** n)
        w0 = psi[((idx >> t) & 1) == 0].sum()
        w1 = psi[((idx >> t) & 1) == 1].sum()
        assert abs(a0 - w0) < 1e-12 and abs(a1 - w1) < 1e-12
        assert abs(m.control_probability(t) - abs(w1) ** 2) < 1e-12
    del rng


# ---------------------------------------------------------------- truncation rule

def test_truncation_rule_vectors():
    # src/tn/svd_guard.rs:289-342
    vals = [1.0 - 0.4 * i / 7.0 for i in range(8)]
    assert om.retained_chi(vals, 8, 8, 8) == 8
    assert not om.needs_f64_fallback(vals, 8)
    vals = [1.0, 0.5, 1e-12 * 1.5]
    assert om.needs_f64_fallback(vals, om.retained_chi(vals, 8, 8, 8))
    vals = [1.0, 1.0 - 1e-8, 0.2]
    chi = om.retained_chi(vals, 8, 8, 8)
    assert chi == 3 and not om.needs_f64_fallback(vals, chi)
    assert om.needs_f64_fallback([1.0, 0.5, 0.5 * (1.0 - 1e-4)], 2)
    # edge cases of retained_chi (src/tn/svd_guard.rs:25-35)
    assert om.retained_chi([], 8, 8, 8) == 0
    assert om.retained_chi([1e-13, 0.0], 8, 8, 8) == 1
    assert om.retained_chi([3, 2, 1, 1e-11, 1e-13], 8, 8, 8) == 4
    assert om.retained_chi([3, 2, 1, 1], 2, 8, 8) == 2
    assert om.retained_chi([3, 2, 1, 1], 8, 3, 8) == 3
    assert om.needs_f64_fallback([1.0, float("nan")], 1)
    assert om.needs_f64_fallback([1.0, -1e-3], 1)


# ---------------------------------------------------------------- SVD

def test_checked_svd_nalgebra_failure_matrix():
    # src/tn/mps.rs:1999-2028
    m = np.array([[4.1435686496782016e-16, -1.597611177955782e-16j],
                  [-0.35975005528623j, 0.9330487113337411]])
    u, s, vt = om.checked_svd(m)
    assert abs(s[0] - 1.0) < 1e-12
    assert np.max(np.abs(u @ np.diag(s) @ vt - m)) < 1e-12


def test_jacobi_tall_wide_rank_deficient():
    # src/tn/mps.rs:2032-2065
    def entry(r, c, k):
        t = (r * 7 + c * 3) * 0.37 + k
        return complex(math.sin(t), math.cos(1.3 * t))

    for rows, cols in [(4, 4), (6, 3), (3, 6), (8, 8)]:
        full = np.array([[entry(r, c, 0.0) for c in range(cols)] for r in range(rows)])
        rank1 = np.array([[entry(r, 0, 0.2) * entry(0, c, 0.9) for c in range(cols)]
                          for r in range(rows)])
        for m in (full, rank1):
            u, s, vt = om.checked_svd(m)
            assert np.all(np.diff(s) <= 0)
            assert np.max(np.abs(u @ np.diag(s) @ vt - m)) < 1e-12


@pytest.mark.parametrize("rows,cols", [(64, 64), (96, 40), (40, 96), (128, 128)])
def test_svd_matches_lapack(rows, cols):
    rng = np.random.default_rng(rows * 1000 + cols)
    a = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
    a = a * np.geomspace(1.0, 1e-9, cols)[None, :]
    u, s, vt = om.checked_svd(a)
    s_ref = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(s - s_ref)) <= 1e-13 * s_ref[0]
    assert np.max(np.abs(u.conj().T @ u - np.eye(len(s)))) < 1e-12
    assert np.max(np.abs(u @ np.diag(s) @ vt - a)) < 1e-12 * s_ref[0]


def truncated_product(theta, chi):
    dl, dr = theta.shape[0], theta.shape[3]
    m = theta.reshape(2 * dl, 2 * dr)
    u, s, vh = np.linalg.svd(m)
    return (u[:, :chi] * s[:chi]) @ vh[:chi], s


@pytest.mark.parametrize("dl,dr,chi", [(8, 8, 32), (16, 12, 20), (32, 32, 48), (40, 24, 16)])
def test_exact_split_matches_lapack(dl, dr, chi):
    rng = np.random.default_rng(dl * 100 + dr)
    th = (rng.standard_normal((dl, 2, 2, dr)) + 1j * rng.standard_normal((dl, 2, 2, dr)))
    left, right, path, disc = om.split_theta(th, chi)
    assert path == 0
    k = left.shape[2]
    prod = left.reshape(2 * dl, k) @ right.reshape(k, 2 * dr)
    want, s = truncated_product(th, k)
    assert k == min(chi, 2 * dl, 2 * dr)
    assert np.max(np.abs(prod - want)) <= 1e-12 * s[0]
    # sqrt(sigma) gauge: both factors carry the same column/row norms sqrt(s_k)
    ln = np.linalg.norm(left.reshape(2 * dl, k), axis=0)
    rn = np.linalg.norm(right.reshape(k, 2 * dr), axis=1)
    assert np.allclose(ln ** 2, s[:k], rtol=1e-12) and np.allclose(rn ** 2, s[:k], rtol=1e-12)
    assert abs(disc - (s[k:] ** 2).sum() / (s ** 2).sum()) <= 1e-12


def test_reject_truncation_raises_bond_exceeded():
    # truncated_svd_split with reject_truncation (src/tn/mps.rs:1137-1143)
    rng = np.random.default_rng(5)
    th = rng.standard_normal((8, 2, 2, 8)) + 0j
    with pytest.raises(om.OracleError) as e:
        om.split_theta(th, 4, reject=True)
    assert e.value.code == 5 and "bond exceeded" in str(e.value)


# ---------------------------------------------------------------- range finder

def py_rademacher(cols, ell, salt):
    """Independent restatement of src/tn/mps.rs:1103-1114 with Python integers."""
    state = salt | 1
    scale = 1.0 / math.sqrt(cols)
    out = np.zeros(cols * ell, dtype=complex)
    for i in range(cols * ell):
        state = (state * 6364136223846793005 + 1) % (1 << 64)
        out[i] = complex(scale if state & 0x100 == 0 else -scale,
                         scale if state & 0x200 == 0 else -scale)
    return out.reshape(cols, ell)


def test_omega_and_mix_seed():
    for rows, cols, ell in [(256, 256, 144), (1024, 1024, 528), (32, 24, 8)]:
        want = ((rows * 0x9E37 + cols) * 0x9E37 + ell) % (1 << 64)
        assert om.mix_seed(rows, cols, ell) == want
    salt = om.mix_seed(256, 256, 144)
    assert np.array_equal(om.rademacher(256, 144, salt), py_rademacher(256, 144, salt))


def test_thin_q_closed_checks():
    rng = np.random.default_rng(11)
    y = rng.standard_normal((60, 20)) + 1j * rng.standard_normal((60, 20))
    q = om.thin_q(y)
    assert np.max(np.abs(q.conj().T @ q - np.eye(20))) < 1e-13
    r = q.conj().T @ y
    assert np.max(np.abs(np.tril(r, -1))) < 1e-12
    assert np.max(np.abs(q @ r - y)) < 1e-12


def test_rsvd_gapped_rank4():
    # src/tn/mps.rs:1795-1854 (orthonormal factors from numpy QR of the same raw matrices)
    rows, cols, ell, dl, dr = 32, 24, 8, 16, 12
    sig = [8.0, 4.0, 2.0, 1.0]
    u_raw = np.array([[complex(math.sin(r * 3 + c * 5), math.cos((r + c) * 0.3)) for c in range(4)]
                      for r in range(rows)])
    v_raw = np.array([[complex(math.cos(r * 7 + c * 11), math.sin((r * 2 + c) * 0.2))
                       for c in range(4)] for r in range(cols)])
    uq = np.linalg.qr(u_raw)[0]
    vq = np.linalg.qr(v_raw)[0]
    mat = (uq * sig) @ vq.conj().T
    res = om.rsvd_sites(mat, ell, dl, dr, 4)
    assert res is not None
    left, right = res
    assert left.shape[2] == 4
    prod = left.reshape(rows, 4) @ right.reshape(4, cols)
    assert np.max(np.abs(prod - mat)) < 1e-6


def numpy_rsvd(mat, max_bond):
    """numpy/LAPACK restatement of rsvd_sites (src/tn/mps.rs:983-1053); returns the
    gauge-invariant product and kept chi, or None for the reference's None."""
    rows, cols = mat.shape
    ell = min(max_bond + 16, rows, cols)
    om_ = py_rademacher(cols, ell, ((rows * 0x9E37 + cols) * 0x9E37 + ell) % (1 << 64))
    y = mat @ (mat.conj().T @ (mat @ om_))
    q = np.linalg.qr(y)[0]
    b = q.conj().T @ mat
    u, s, vh = np.linalg.svd(b, full_matrices=False)
    vals = np.where(s < 1e-14, 0.0, s)
    chi = om.retained_chi(vals, max_bond, rows, cols)
    if chi == 0 or om.needs_f64_fallback(vals, chi):
        return None
    return (q @ u[:, :chi] * s[:chi]) @ vh[:chi], chi


@pytest.mark.parametrize("side,chi,decay", [(320, 32, 0.02), (300, 30, 0.05), (400, 32, 0.01)])
def test_rsvd_matches_numpy_restatement(side, chi, decay):
    """Range finder against LAPACK where the sketch has full column rank (ell <= 48): there
    Q is fixed by range(Y) and the product is convention-free. With ell > 48 the sketch has
    rank 48 (test_reference_omega_has_rank_48) and Q's completion follows rounding, so
    any two QR implementations differ at O(sigma_48) and only the exact-rank rule compares."""
    rng = np.random.default_rng(side + chi)
    dl = dr = side // 2
    ell = min(chi + 16, side)
    omega = py_rademacher(side, ell, ((side * 0x9E37 + side) * 0x9E37 + ell) % (1 << 64))
    assert np.linalg.matrix_rank(omega) == ell
    u = np.linalg.qr(rng.standard_normal((side, side)) + 1j * rng.standard_normal((side, side)))[0]
    v = np.linalg.qr(rng.standard_normal((side, side)) + 1j * rng.standard_normal((side, side)))[0]
    s = np.exp(-decay * np.arange(side))
    mat = (u * s) @ v.conj().T
    th = mat.reshape(dl, 2, 2, dr)
    left, right, path, _ = om.split_theta(th, chi)
    assert path == 1
    want, kchi = numpy_rsvd(mat, chi)
    k = left.shape[2]
    assert k == kchi
    prod = left.reshape(side, k) @ right.reshape(k, side)
    assert np.max(np.abs(prod - want)) <= 1e-10 * s[0]


# ---------------------------------------------------------------- sampling / expectation

def test_sampler_kat_and_chacha_vector():
    m = om.OracleMps(1, 1)
    m.apply_circuit([C.one(C.H, 0)])
    sv = m.contract_to_statevector()
    assert om.sample_counts(sv, 1, [(0, 0)], 1, 256, 99) == {"0": 118, "1": 138}
    assert om.sample_counts(sv, 1, [(0, 0)], 1, 256, 99) == om.sample_counts(sv, 1, [(0, 0)], 1, 256, 99)
    # RFC 8439 section 2.3.2 block function vector (20 rounds)
    key = [int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)]
    out = om.chacha_block(key, 1 | (0x09000000 << 32), 0x4A000000, 20)
    assert out[0] == 0xE4E7F110 and out[15] == 0x4E3C50A2
    # snapped CPU and f32 GPU statevectors share counts (src/sample.rs:257-274)
    gpu = np.array([0.4999999701976776 - 0.4999999701976776j, 0.4999999701976776 + 0.4999999701976776j])
    cpu = np.array([0.5000000000000001 - 0.5000000000000001j, 0.5000000000000001 + 0.5000000000000001j])
    assert om.sample_counts(gpu, 1, [(0, 0)], 1, 256, 99) == om.sample_counts(cpu, 1, [(0, 0)], 1, 256, 99)


def _expect(n, gates, label):
    m = om.OracleMps(n, 8)
    m.apply_circuit(gates)
    return om.pauli_expectation(m.contract_to_statevector(), label)


def test_pauli_convention_kats():
    # src/expectation.rs:291-466
    assert abs(_expect(1, [], "Z") - 1.0) < 1e-12
    assert abs(_expect(1, [C.one(C.H, 0), C.one(C.S, 0)], "Y") - 1.0) < 1e-9
    assert abs(_expect(1, [C.one(C.H, 0), C.one(C.S, 0), C.one(C.Z, 0)], "Y") + 1.0) < 1e-9
    cb = [C.one(C.H, 0), C.one(C.S, 0), C.two(C.CNOT, 0, 1)]
    for lab in ("XY", "YX"):
        v = _expect(2, cb, lab)
        assert abs(v.real - 1.0) < 1e-9 and abs(v.imag) < 1e-12
    assert abs(_expect(2, [C.one(C.H, 0), C.two(C.CNOT, 0, 1)], "YY") + 1.0) < 1e-9
    assert abs(_expect(1, [C.one(C.H, 0)], "Z")) < 1e-12
    assert abs(_expect(2, [C.one(C.H, 0), C.two(C.CNOT, 0, 1)], "ZZ") - 1.0) < 1e-12


def test_range_finder_amplifies_rounding():
    """The reference's range finder (src/tn/mps.rs:983-1053) is ill-conditioned in its
    input: Y = A A^H A Omega is not re-orthonormalized and Omega has rank 48, so the
    Householder completion of Q follows the direction of nearly collinear columns. A 1e-14
    relative perturbation of the state before the first range-finder layer changes
    <0|psi> by far more than 1e-10 three range-finder layers later. Amplitudes of
    workloads with range-finder splits are therefore reproducible only to this
    amplification, by any implementation; per-split parity on identical inputs is not
    affected."""
    from test_gpu_parity import layers_of

    om.set_qr_rank_tol(1e-10)
    try:
        n, chi = 40, 128
        L = layers_of(C.rotation_cz_brickwork(40, 26, 0x5EED), n)
        o = om.OracleMps(n, chi)
        for lay in L[:20]:
            o.apply_circuit(lay)
        sites = [o.site(i) for i in range(n)]
        amps = []
        for eps in (0.0, 1e-14):
            rng = np.random.default_rng(7)
            p = om.OracleMps(n, chi)
            for i, s in enumerate(sites):
                d = (rng.standard_normal(s.shape) + 1j * rng.standard_normal(s.shape)) * eps * np.abs(s).max()
                p.set_site(i, s + d)
            for lay in L[20:23]:
                p.apply_circuit(lay)
            amps.append(p.amplitude_at_compact_zero())
        assert abs(amps[1] - amps[0]) / abs(amps[0]) > 1e-7
    finally:
        om.set_qr_rank_tol(0.0)


def test_reference_omega_has_rank_48():
    """rademacher (src/tn/mps.rs:1103-1114) reads bits 8 and 9 of a 2^64 LCG, which
    depend only on the low 10 state bits: entries repeat every 1024 draws, so Omega has at
    most 64 distinct rows and numerical rank 48 at the saturated shapes."""
    def fast_rademacher(cols, ell, salt):
        # bits 8 and 9 of the state depend only on its low 10 bits (period 1024)
        a, x = 6364136223846793005 % 1024, (salt | 1) % 1024
        low = []
        for _ in range(1024):
            x = (a * x + 1) % 1024
            low.append(x)
        low = np.array(low)[np.arange(cols * ell) % 1024]
        sc = 1.0 / math.sqrt(cols)
        re = np.where(low & 0x100, -sc, sc)
        im = np.where(low & 0x200, -sc, sc)
        return (re + 1j * im).reshape(cols, ell)

    O = fast_rademacher(256, 144, om.mix_seed(256, 256, 144))
    assert np.array_equal(O, py_rademacher(256, 144, om.mix_seed(256, 256, 144)))
    for rows, cols, ell in [(256, 256, 144), (512, 512, 272), (1024, 1024, 528)]:
        O = fast_rademacher(cols, ell, om.mix_seed(rows, cols, ell))
        s = np.linalg.svd(O, compute_uv=False)
        assert int((s > 1e-10 * s[0]).sum()) == 48
