"""Config 4: a batch of 256 independent slices (widths 12..60, depths 8..32, chi in
{16, 32, 64, 128, 256}, config-1 gate set) through wqc_tn_run_slices.

Bars:
- the batched form (workers = 0: every layer step of all slices in shared launches) and the
  threaded form (8 host threads, one state per slice) run every slice (status OK, finite
  amplitude);
- both agree with one-at-a-time runs on one state: to 1e-12 relative on every slice whose
  splits were all exact (the forms differ only in the summation order of a few batched
  reductions), and with identical bonds everywhere (checked on the one-at-a-time states);
- against the oracle: 1e-10 relative on range-finder-free slices (at least 16 of them), and
  identical bonds and split counters on at least four chi = 128 and four chi = 256 slices whose
  splits take the range finder (the range finder amplifies rounding in any implementation, see
  test_oracle.py::test_range_finder_amplifies_rounding, so amplitudes there are compared only
  as printed figures).
"""

import os
import sys
import time

import numpy as np
import pytest

import circuits as C
from oracle import oracle_mps as om

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wt():
    import wqc_tn

    wqc_tn.load()
    return wqc_tn


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def one_at_a_time(wt, n, chi, gates):
    st = wt.MpsState(n, chi)
    st.enable_split_log(True)
    st.apply_circuit(gates)
    a = st.amplitude_at_compact_zero()
    bonds = st.bond_dims()
    paths = [e[3] for e in st.split_log()]
    st.free()
    return a, bonds, paths


def test_config4_batched_and_threaded(wt):
    slices = C.config4()
    assert len(slices) == 256
    t = time.time()
    amps_b, codes_b = wt.run_slices(slices, workers=0)
    tb = time.time() - t
    t = time.time()
    amps_t, codes_t = wt.run_slices(slices, workers=8)
    tt = time.time() - t
    say("config4: 256 slices batched %.2f s, threaded (8 workers) %.2f s" % (tb, tt))
    bad = [(i, codes_b[i], codes_t[i]) for i in range(256) if codes_b[i] != wt.OK or codes_t[i] != wt.OK]
    assert not bad, bad
    assert np.all(np.isfinite(amps_b)) and np.all(np.isfinite(amps_t))
    worst, exact = 0.0, 0
    for i in range(0, 256, 4):
        n, chi, gates = slices[i]
        a, _, paths = one_at_a_time(wt, n, chi, gates)
        if any(p != 0 for p in paths):
            continue
        exact += 1
        e = max(rel(amps_b[i], a), rel(amps_t[i], a))
        worst = max(worst, e)
        assert e <= 1e-12, (i, a, amps_b[i], amps_t[i])
    say("config4: batched/threaded vs one at a time on %d range-finder-free slices: worst %.1e" % (exact, worst))
    assert exact >= 16


def test_config4_slices_vs_oracle(wt):
    slices = C.config4()
    om.set_threads(min(16, os.cpu_count() or 1))
    amps, codes = wt.run_slices(slices, workers=0)
    assert all(c == wt.OK for c in codes)
    # Range-finder-free slices cheap enough for the CPU oracle.
    checked = 0
    for i, (n, chi, gates) in enumerate(slices):
        units = C.two_qubit_units(gates, n)
        if chi > 64 or units * min(chi, 2 ** (n // 2)) ** 3 > 4e8:
            continue
        o = om.OracleMps(n, chi, record=True)
        o.apply_circuit(gates)
        want = o.amplitude_at_compact_zero()
        _, bonds, _ = one_at_a_time(wt, n, chi, gates)
        assert bonds == o.bond_dims(), (i, [(q, x, y) for q, (x, y) in enumerate(zip(bonds, o.bond_dims())) if x != y])
        assert o.counters()["rsvd_kept"] == 0
        assert rel(amps[i], want) <= 1e-10, (i, amps[i], want)
        checked += 1
        if checked >= 24:
            break
    say("config4: range-finder-free slices against the oracle:", checked)
    assert checked >= 16
    # chi = 128 and chi = 256 slices whose splits reach the range finder (theta 256^2 and
    # 512^2), the cheapest four of each, against the oracle's exact-arithmetic rank rule.
    om.set_qr_rank_tol(1e-10)
    try:
        for cap in (128, 256):
            cands = []
            for i, (n, chi, gates) in enumerate(slices):
                if chi != cap or 2 ** (n // 2) < 2 * cap:
                    continue
                cands.append((C.two_qubit_units(gates, n) * min(cap, 2 ** (n // 2)) ** 3, i))
            cands.sort()
            done = 0
            for _, i in cands:
                n, chi, gates = slices[i]
                a, bonds, paths = one_at_a_time(wt, n, chi, gates)
                if paths.count(1) == 0:
                    continue
                t = time.time()
                o = om.OracleMps(n, chi, record=True)
                o.apply_circuit(gates)
                ctr = o.counters()
                want = o.amplitude_at_compact_zero()
                say("config4 slice %d: %d qubits chi %d, range-finder splits gpu %d oracle %d, oracle %.1f s, "
                    "amplitude rel diff %.1e" % (i, n, chi, paths.count(1), ctr["rsvd_kept"], time.time() - t,
                                                 rel(a, want)))
                assert bonds == o.bond_dims(), i
                assert paths.count(1) == ctr["rsvd_kept"] and paths.count(2) == ctr["rsvd_fallbacks"], i
                done += 1
                if done >= 4:
                    break
            assert done >= 4, cap
    finally:
        om.set_qr_rank_tol(0.0)


@pytest.mark.parametrize("prec", [1, 2])
def test_config4_mixed_forms(wt, prec):
    """Config 4 in the mixed forms (FP32 or split-BF16 early Jacobi sweeps, FP64 finish, FP64
    truncation): the 256-slice batch runs every slice (status OK, finite amplitude), and on the
    slices that never reach the range finder (chi <= 64: theta at most 128 x 128, below the
    range finder's size threshold) the amplitude matches the FP64 batch to 1e-12 of the state's
    norm (the states are normalized, so this bounds the absolute error of <0...0|psi>; measured
    9.6e-15)."""
    slices = C.config4()
    amps64, codes64 = wt.run_slices(slices, precision=wt.PRECISION_FP64, workers=0)
    t = time.time()
    amps, codes = wt.run_slices(slices, precision=prec, workers=0)
    say("config4 precision %d: 256 slices batched %.2f s" % (prec, time.time() - t))
    assert all(c == wt.OK for c in codes) and all(c == wt.OK for c in codes64)
    assert np.all(np.isfinite(amps))
    worst_abs, worst_rel, checked = 0.0, 0.0, 0
    for i, (n, chi, gates) in enumerate(slices):
        if chi > 64:
            continue
        checked += 1
        worst_abs = max(worst_abs, abs(amps[i] - amps64[i]))
        worst_rel = max(worst_rel, rel(amps[i], amps64[i]))
    say("config4 precision %d vs FP64 on %d range-finder-free slices: worst |diff| %.1e, worst rel %.1e"
        % (prec, checked, worst_abs, worst_rel))
    assert checked >= 100
    assert worst_abs <= 1e-12
