"""Config 4: a batch of 256 independent slices (widths 12..60, depths 8..32, chi in
{16, 32, 64, 128, 256}, config-1 gate set) through wqc_tn_run_slices.

Bars:
- every slice runs (status OK, finite amplitude);
- the concurrent batch is bitwise equal to running the same slices one at a time on a
  single state (determinism under concurrency);
- a cost-bounded subsample against the oracle: 1e-10 relative where the oracle took no
  range-finder split; where it did, identical bond dimensions (the range finder amplifies
  rounding in any implementation, see test_oracle.py::test_range_finder_amplifies_rounding).
"""

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


def test_config4_batch(wt):
    slices = C.config4()
    assert len(slices) == 256
    t = time.time()
    amps, codes = wt.run_slices(slices, workers=8)
    say("config4: 256 slices in %.2f s" % (time.time() - t))
    assert all(c == wt.OK for c in codes)
    assert np.all(np.isfinite(amps))
    # Bitwise equality with one-at-a-time runs on a subsample (every 8th slice).
    for i in range(0, 256, 8):
        n, chi, gates = slices[i]
        st = wt.MpsState(n, chi)
        st.apply_circuit(gates)
        a = st.amplitude_at_compact_zero()
        st.free()
        assert a == amps[i], (i, a, amps[i])
    # Oracle subsample: slices whose two-qubit work is small enough for the CPU oracle.
    om.set_threads(8)
    checked = exact_checked = 0
    for i, (n, chi, gates) in enumerate(slices):
        units = C.two_qubit_units(gates, n)
        if chi > 64 or units * min(chi, 2 ** (n // 2)) ** 3 > 4e8:
            continue
        o = om.OracleMps(n, chi, record=True)
        o.apply_circuit(gates)
        want = o.amplitude_at_compact_zero()
        ctr = o.counters()
        st = wt.MpsState(n, chi)
        st.apply_circuit(gates)
        bonds = st.bond_dims()
        st.free()
        assert bonds == o.bond_dims(), i
        if ctr["rsvd_kept"] == 0:
            assert abs(amps[i] - want) <= 1e-10 * max(abs(want), 1e-300), (i, amps[i], want)
            exact_checked += 1
        checked += 1
        if checked >= 48:
            break
    say("config4: oracle-checked slices", checked, "of which range-finder-free", exact_checked)
    assert exact_checked >= 16
