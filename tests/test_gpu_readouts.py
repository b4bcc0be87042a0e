"""GPU readouts beyond <0...0|psi>: Pauli expectations (transfer matrices on the device) and
sampling, through the C ABI.

Checks:
- expectations against the oracle's dense pauli_expectation on the same circuit
  (src/expectation.rs:140-167 semantics) at n <= 14, 1e-10 absolute (states have norm 1);
- the reference's convention known-answer tests (src/expectation.rs:291-466) on the GPU;
- at n = 32 and n = 40, where the reference refuses (n > 20), against a numpy einsum
  transfer-matrix contraction of the same sites (GPU sites, and the oracle's own MPS).
"""

import math
import sys

import numpy as np
import pytest

import circuits as C
from oracle import oracle_mps as om

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


def einsum_expectation(sites, label):
    """<psi|P|psi> of an MPS given as (l, 2, r) arrays, site q = qubit q = label[n-1-q]."""
    n = len(sites)
    E = np.ones((1, 1), dtype=complex)
    for q, A in enumerate(sites):
        P = PAULI[label[n - 1 - q]]
        E = np.einsum("xsy,st,xa,atb->yb", A.conj(), P, E, A, optimize=True)
    return E[0, 0]


def random_labels(rng, n, count):
    out = ["I" * n]
    for _ in range(count):
        lab = ["I"] * n
        k = 1 + rng.below(min(n, 4))
        for _ in range(k):
            lab[rng.below(n)] = "IXYZ"[1 + rng.below(3)]
        if rng.below(4) == 0:
            lab = ["IXYZ"[rng.below(4)] for _ in range(n)]
        out.append("".join(lab))
    return out


def test_expectations_vs_dense_oracle(wt):
    from test_oracle import seeded_random_circuit

    rng = C.SplitMix64(0xE4_2026_1203)
    worst = 0.0
    for case in range(12):
        n = 4 + rng.below(11)
        gates = seeded_random_circuit(rng, n, 6 + rng.below(20))
        g = wt.MpsState(n, 64)
        g.apply_circuit(gates)
        o = om.OracleMps(n, 64)
        o.apply_circuit(gates)
        sv = o.contract_to_statevector()
        labels = random_labels(rng, n, 24)
        got = g.pauli_expectations(labels, with_imag=True)
        for lab, v in zip(labels, got):
            want = om.pauli_expectation(sv, lab)
            worst = max(worst, abs(v.real - want.real))
            assert abs(v.real - want.real) <= 1e-10, (case, n, lab, v, want)
            assert abs(v.imag) <= 1e-10, (case, lab, v)
    say("worst |gpu - dense| =", worst)


def _gpu_expect(wt, n, gates, label):
    g = wt.MpsState(n, 8)
    g.apply_circuit(gates)
    return g.pauli_expectations([label], with_imag=True)[0]


def test_pauli_convention_kats(wt):
    # src/expectation.rs:291-466 on the device path
    assert abs(_gpu_expect(wt, 1, [], "Z") - 1.0) < 1e-12
    assert abs(_gpu_expect(wt, 1, [C.one(C.H, 0), C.one(C.S, 0)], "Y") - 1.0) < 1e-9
    assert abs(_gpu_expect(wt, 1, [C.one(C.H, 0), C.one(C.S, 0), C.one(C.Z, 0)], "Y") + 1.0) < 1e-9
    cb = [C.one(C.H, 0), C.one(C.S, 0), C.two(C.CNOT, 0, 1)]
    for lab in ("XY", "YX"):
        v = _gpu_expect(wt, 2, cb, lab)
        assert abs(v.real - 1.0) < 1e-9 and abs(v.imag) < 1e-12
    assert abs(_gpu_expect(wt, 2, [C.one(C.H, 0), C.two(C.CNOT, 0, 1)], "YY") + 1.0) < 1e-9
    assert abs(_gpu_expect(wt, 1, [C.one(C.H, 0)], "Z")) < 1e-12
    assert abs(_gpu_expect(wt, 2, [C.one(C.H, 0), C.two(C.CNOT, 0, 1)], "ZZ") - 1.0) < 1e-12
    # Qiskit order: "IX" acts on qubit 0 only; X on |0> of qubit 1 gives 0
    assert abs(_gpu_expect(wt, 2, [C.one(C.H, 0)], "IX") - 1.0) < 1e-12
    assert abs(_gpu_expect(wt, 2, [C.one(C.H, 0)], "XI")) < 1e-12


def test_config5_observable_vs_einsum_and_oracle_mps(wt):
    n, chi, gates, _, terms = C.config5()
    g = wt.MpsState(n, chi)
    g.apply_circuit(gates)
    labels = [lab for _, lab in terms]
    got = g.pauli_expectations(labels, with_imag=True)
    gsites = [g.site(q) for q in range(n)]
    o = om.OracleMps(n, chi)
    o.apply_circuit(gates)
    osites = [o.site(q) for q in range(n)]
    assert [s.shape for s in gsites] == [s.shape for s in osites]
    worst_self = worst_oracle = 0.0
    for lab, v in zip(labels, got):
        e_self = einsum_expectation(gsites, lab)
        e_orc = einsum_expectation(osites, lab)
        worst_self = max(worst_self, abs(v.real - e_self.real))
        worst_oracle = max(worst_oracle, abs(v.real - e_orc.real))
        assert abs(v.imag) < 1e-10
    say("config5: worst vs einsum on GPU sites", worst_self, "vs oracle MPS", worst_oracle)
    assert worst_self <= 1e-11
    assert worst_oracle <= 1e-9
    # Observable: snapped sum (src/expectation.rs:170-200, src/engine.rs:30-40)
    obs = g.observable_expectation([(lab, c) for c, lab in terms])
    raw = sum(c * einsum_expectation(gsites, lab).real for c, lab in terms)
    assert obs == complex(round(raw * 1e6) / 1e6, 0.0) or abs(obs.real - raw) <= 5e-7 + 1e-12
    assert obs.imag == 0.0


def test_expectation_n40_and_errors(wt):
    n, chi, gates = C.config2()
    g = wt.MpsState(n, chi)
    g.apply_circuit(gates)
    rng = C.SplitMix64(0x40)
    labels = random_labels(rng, n, 8)
    got = g.pauli_expectations(labels)
    sites = [g.site(q) for q in range(n)]
    for lab, v in zip(labels, got):
        assert abs(v - einsum_expectation(sites, lab).real) <= 1e-11, lab
    # norm term equals <psi|psi>
    assert abs(got[0] - einsum_expectation(sites, "I" * n).real) < 1e-12
    with pytest.raises(wt.WqcTnError) as e:
        g.pauli_expectations(["Q" * n])
    assert e.value.code == wt.INVALID_ARG and "invalid pauli character" in str(e.value)


# ---------------------------------------------------------------- sampling

def test_sample_kat_and_exact_semantics_vs_oracle(wt):
    # src/sample.rs:276-290: H, 256 shots, seed 99
    g = wt.MpsState(1, 4)
    g.apply_circuit([C.one(C.H, 0)])
    assert g.sample_counts([(0, 0)], 1, 256, 99, method=wt.SAMPLE_EXACT) == {"0": 118, "1": 138}
    # reproducible (src/sample.rs:292-307)
    assert g.sample_counts([(0, 0)], 1, 256, 99) == g.sample_counts([(0, 0)], 1, 256, 99)
    from test_oracle import seeded_random_circuit

    rng = C.SplitMix64(0x5A_2026_1203)
    for case in range(10):
        n = 2 + rng.below(11)
        gates = seeded_random_circuit(rng, n, 8 + rng.below(16))
        g = wt.MpsState(n, 64)
        g.apply_circuit(gates)
        o = om.OracleMps(n, 64)
        o.apply_circuit(gates)
        sv = o.contract_to_statevector()
        cbits = 1 + rng.below(n + 2)
        measures = [(rng.below(n), rng.below(cbits)) for _ in range(1 + rng.below(n))]
        shots, seed = 1 + rng.below(3000), rng.next()
        got = g.sample_counts(measures, cbits, shots, seed, method=wt.SAMPLE_EXACT)
        want = om.sample_counts(sv, n, measures, cbits, shots, seed)
        assert got == want, (case, n, measures, cbits)
        assert sum(got.values()) == shots


def test_sample_errors(wt):
    g = wt.MpsState(3, 4)
    with pytest.raises(wt.WqcTnError) as e:
        g.sample_counts([(0, 0)], 0, 10, 1)
    assert e.value.code == wt.EXEC_FAILED and "classical_bit_count must be > 0" in str(e.value)
    with pytest.raises(wt.WqcTnError) as e:
        g.sample_counts([(3, 0)], 1, 10, 1)
    assert e.value.code == wt.INVALID_ARG
    with pytest.raises(wt.WqcTnError) as e:
        g.sample_counts([(0, 2)], 2, 10, 1)
    assert e.value.code == wt.EXEC_FAILED and "classical bit index 2 out of bounds" in str(e.value)
    big = wt.MpsState(24, 4)
    with pytest.raises(wt.WqcTnError) as e:
        big.sample_counts([(0, 0)], 1, 10, 1, method=wt.SAMPLE_EXACT)
    assert e.value.code == wt.EXEC_FAILED and "qubit_count <= 20" in str(e.value)


def test_mps_sampler_matches_born_distribution(wt):
    """The MPS sampler draws from |psi|^2 / <psi|psi>: at n = 8 the empirical distribution of
    40000 shots is within total variation 0.05 of the exact one (the expected TV of a
    perfect sampler here is about 0.5 * sqrt(2 * 256 / (pi * 40000)) = 0.032)."""
    from test_oracle import seeded_random_circuit

    rng = C.SplitMix64(0x3A3A)
    n = 8
    gates = seeded_random_circuit(rng, n, 30)
    g = wt.MpsState(n, 16)
    g.apply_circuit(gates)
    sv = g.contract_to_statevector()
    p = np.abs(sv) ** 2
    p /= p.sum()
    shots = 40000
    measures = [(q, q) for q in range(n)]
    counts = g.sample_counts(measures, n, shots, 7, method=wt.SAMPLE_MPS)
    assert sum(counts.values()) == shots
    emp = np.zeros(1 << n)
    for k, c in counts.items():
        emp[int(k, 2)] = c / shots
    tv = 0.5 * np.abs(emp - p).sum()
    say("MPS sampler TV at n=8, 40000 shots:", tv)
    assert tv <= 0.05
    assert g.sample_counts(measures, n, 500, 11, method=wt.SAMPLE_MPS) == g.sample_counts(measures, n, 500, 11, method=wt.SAMPLE_MPS)


def test_mps_sampler_config2_marginals(wt):
    """Config 2 (40 qubits, chi = 128, 4096 shots): single-qubit <Z> marginals of the counts
    agree with the transfer-matrix expectations of the same MPS within 5 standard errors."""
    n, chi, gates = C.config2()
    g = wt.MpsState(n, chi)
    g.apply_circuit(gates)
    shots = 4096
    counts = g.sample_counts([(q, q) for q in range(n)], n, shots, 0xC0F2, method=wt.SAMPLE_AUTO)
    assert sum(counts.values()) == shots
    labels = ["I" * n] + ["".join("Z" if c == n - 1 - q else "I" for c in range(n)) for q in range(n)]
    ev = g.pauli_expectations(labels)
    norm = ev[0]
    worst = 0.0
    for q in range(n):
        z = ev[1 + q] / norm
        emp = sum(c * (1 - 2 * int(k[n - 1 - q])) for k, c in counts.items()) / shots
        se = math.sqrt(max(1 - z * z, 1e-6) / shots)
        worst = max(worst, abs(emp - z) / se)
        assert abs(emp - z) <= 5 * se + 1e-3, (q, emp, z)
    say("config2 sampler: worst marginal deviation in standard errors", worst)


def test_outcome_probabilities_kat_and_oracle(wt):
    """compute_outcome_probabilities (src/sample.rs:85-131) on the device statevector: the
    Bell-state table of src/distribution_proof.rs:242-265 (two entries, 0.707107^2 each after
    the 1e-6 snap), then random circuits and measure maps against the snapped table of the
    oracle's own statevector, and the reference's errors."""
    g = wt.MpsState(2, 4)
    g.apply_circuit([C.one(C.H, 0), C.two(C.CNOT, 0, 1)])
    table = g.outcome_probabilities([(0, 0), (1, 1)], 2)
    assert sorted(table) == ["00", "11"]
    assert table["00"] == 0.707107 * 0.707107 and table["11"] == 0.707107 * 0.707107
    from test_oracle import seeded_random_circuit

    rng = C.SplitMix64(0x0B_2026)
    for case in range(8):
        n = 2 + rng.below(11)
        gates = seeded_random_circuit(rng, n, 8 + rng.below(16))
        g = wt.MpsState(n, 64)
        g.apply_circuit(gates)
        o = om.OracleMps(n, 64)
        o.apply_circuit(gates)
        sv = o.contract_to_statevector()
        cbits = 1 + rng.below(n + 2)
        measures = [(rng.below(n), rng.below(cbits)) for _ in range(1 + rng.below(n))]
        want = {}
        for basis, amp in enumerate(sv):
            re, im = round(amp.real * 1e6) / 1e6, round(amp.imag * 1e6) / 1e6
            p = re * re + im * im
            if p == 0.0:
                continue
            key = 0
            for q, c in measures:
                if (basis >> q) & 1:
                    key |= 1 << c
                else:
                    key &= ~(1 << c)
            k = format(key, "0%db" % cbits)
            want[k] = want.get(k, 0.0) + p
        got = g.outcome_probabilities(measures, cbits)
        assert sorted(got) == sorted(want), (case, n, measures)
        for k in want:
            assert abs(got[k] - want[k]) <= 1e-12, (case, k, got[k], want[k])
    with pytest.raises(wt.WqcTnError) as e:
        wt.MpsState(3, 4).outcome_probabilities([(0, 0)], 0)
    assert e.value.code == wt.EXEC_FAILED and "classical_bit_count must be > 0" in str(e.value)
    with pytest.raises(wt.WqcTnError) as e:
        wt.MpsState(24, 4).outcome_probabilities([(0, 0)], 1)
    assert e.value.code == wt.EXEC_FAILED and "qubit_count <= 20" in str(e.value)
