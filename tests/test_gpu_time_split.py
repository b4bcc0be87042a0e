"""Time-split overlap (contract_time_split, src/tn/mps.rs:1241-1474) through the C ABI:
wqc_tn_contract_time_split and wqc_tn_overlap.

The reference's own tests (src/tn/mps.rs:1561-1697) are restated on the device: the dense
statevector amplitude <0...0|U|0...0> (tests/dense_ref.py) is the bar (1e-10 absolute, as the
reference's 1e-8 tightened), reject_truncation errors where a half needs more bond than
max_bond, the 256-entry environment limit, S/T adjoint sequences. Added: random circuits with
long-range and three-qubit gates split at every tenth gate, and the error paths.
"""

import math
import sys

import numpy as np
import pytest

import circuits as C
import dense_ref as D

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def wt():
    import wqc_tn

    wqc_tn.load()
    return wqc_tn


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ry_cnot_brickwork(n, depth):
    """src/tn/mps.rs:1542-1559: RY(q, 0.17 + 0.01 (q + layer)) on every qubit, then CNOTs on
    alternating pairs; the split is the first gate of layer depth / 2."""
    gates, split = [], 0
    for layer in range(depth):
        if layer == depth // 2:
            split = len(gates)
        for q in range(n):
            gates.append(C.one(C.RY, q, 0.17 + 0.01 * (q + layer)))
        for q in range(layer % 2, n - 1, 2):
            gates.append(C.two(C.CNOT, q, q + 1))
    return gates, split


def dense_amp(n, gates):
    return D.run(n, gates)[0]


def test_hadamard_matches_dense(wt):
    # src/tn/mps.rs:1561-1573
    gates = [C.one(C.H, 0), C.one(C.S, 0), C.one(C.RZ, 0, 0.4)]
    got = wt.contract_time_split(1, gates, 2, 128)
    assert abs(got - dense_amp(1, gates)) < TOL


def test_brickwork_matches_dense_and_rejects_truncation(wt):
    # src/tn/mps.rs:1576-1596: 16 qubits, depth 16; each half keeps bond 16, one sweep of the
    # whole depth needs 256 > 128 and must fail instead of truncating.
    n = 16
    gates, split = ry_cnot_brickwork(n, 16)
    got = wt.contract_time_split(n, gates, split, 128)
    want = dense_amp(n, gates)
    say("time split brickwork 16", got, want)
    assert abs(got - want) < TOL
    with pytest.raises(wt.WqcTnError) as e:
        wt.contract_time_split(n, gates, len(gates), 128)
    assert e.value.code == wt.BOND_EXCEEDED


def test_s_and_t_adjoint_sequences(wt):
    # src/tn/mps.rs:1682-1697: S^H = Z S, T^H = Z T T T
    gates = [C.one(C.H, 0), C.one(C.S, 0), C.one(C.T, 0)]
    got = wt.contract_time_split(1, gates, 1, 128)
    assert abs(got - dense_amp(1, gates)) < TOL


def test_long_range_gates_at_chi16(wt):
    # src/tn/mps.rs:1647-1680: a far CNOT on a width-16 half must be rejected at chi 16; on a
    # width-8 half (depth 6) the SWAP chain fits and the amplitude matches.
    n = 16
    gates, _ = ry_cnot_brickwork(n, 8)
    gates.append(C.two(C.CNOT, 0, n - 1))
    with pytest.raises(wt.WqcTnError) as e:
        wt.contract_time_split(n, gates, len(gates), 16)
    assert e.value.code == wt.BOND_EXCEEDED
    gates, _ = ry_cnot_brickwork(n, 6)
    gates.append(C.two(C.CNOT, 0, n - 1))
    got = wt.contract_time_split(n, gates, len(gates), 16)
    assert abs(got - dense_amp(n, gates)) < TOL


def forced_bond_sites(n, bond):
    """src/tn/mps.rs:1633-1645: site i of shape (left, 2, right) with bond `bond` inside,
    1/sqrt(left) at [a, 0, min(a, right - 1)]."""
    sites = []
    for i in range(n):
        left = 1 if i == 0 else bond
        right = 1 if i + 1 == n else bond
        t = np.zeros((left, 2, right), dtype=complex)
        for a in range(left):
            t[a, 0, min(a, right - 1)] = 1.0 / math.sqrt(left)
        sites.append(t)
    return sites


def test_overlap_environment_limit(wt):
    # src/tn/mps.rs:1625-1631: a 17 x 17 environment is refused, 16 x 16 is contracted.
    for bond, ok in ((17, False), (16, True)):
        ket, bra = wt.MpsState(3, 32), wt.MpsState(3, 32)
        sites = forced_bond_sites(3, bond)
        for i, s in enumerate(sites):
            ket.set_site(i, s)
            bra.set_site(i, s)
        if ok:
            want = np.einsum("xsa,xsb->ab", sites[0], sites[0].conj())
            for s in sites[1:]:
                want = np.einsum("ab,asc,bsd->cd", want, s, s.conj())
            assert abs(ket.overlap(bra) - want[0, 0]) < 1e-12
        else:
            with pytest.raises(wt.WqcTnError) as e:
                ket.overlap(bra)
            assert e.value.code == wt.EXEC_FAILED and "256" in str(e.value)
        ket.free()
        bra.free()


def random_circuit(rng, n, count):
    gates = []
    for _ in range(count):
        r = rng.below(10)
        if r < 5:
            kind = (C.H, C.X, C.Y, C.Z, C.S, C.T, C.RX, C.RY, C.RZ)[rng.below(9)]
            gates.append(C.one(kind, rng.below(n), 2 * math.pi * rng.uniform()))
        elif r < 9:
            a, b = rng.below(n), rng.below(n)
            if a != b:
                gates.append(C.two((C.CNOT, C.CZ)[rng.below(2)], a, b))
        else:
            q = [rng.below(n) for _ in range(3)]
            if len(set(q)) == 3:
                gates.append((C.CCNOT, q[0], q[1], q[2], 0.0))
    return gates


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_circuits_every_split_vs_dense(wt, seed):
    rng = C.SplitMix64(0x7153 + seed)
    n = 6 + rng.below(5)
    gates = random_circuit(rng, n, 60)
    want = dense_amp(n, gates)
    worst = 0.0
    for split in range(0, len(gates) + 1, 10):
        got = wt.contract_time_split(n, gates, split, 64)
        worst = max(worst, abs(got - want))
        assert abs(got - want) < TOL, (n, split, got, want)
    say("time split random n=%d worst |gpu - dense| = %.2e" % (n, worst))


def test_time_split_errors(wt):
    gates = [C.one(C.H, 0), (C.MEASURE, 0, 0, 0, 0.0), C.one(C.X, 1)]
    # MEASURE in the first half is a no-op there (src/tn/mps.rs:1374-1376 via apply_gate)
    got = wt.contract_time_split(2, gates, 2, 8)
    assert abs(got - dense_amp(2, gates)) < TOL
    # ... and refused in the second half (apply_gate_adjoint, src/tn/mps.rs:301-306)
    with pytest.raises(wt.WqcTnError) as e:
        wt.contract_time_split(2, gates, 1, 8)
    assert e.value.code == wt.EXEC_FAILED and "measurement or reset" in str(e.value)
    with pytest.raises(wt.WqcTnError) as e:
        wt.contract_time_split(2, gates, 4, 8)
    assert e.value.code == wt.EXEC_FAILED and "past the gate list" in str(e.value)
    a, b = wt.MpsState(3, 4), wt.MpsState(4, 4)
    with pytest.raises(wt.WqcTnError) as e:
        a.overlap(b)
    assert e.value.code == wt.INVALID_ARG
