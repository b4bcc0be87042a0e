"""Size limits, the device-memory gate and the checked split on the GPU, through the C ABI.

* mps_accepts_registers_above_40_qubits (src/tn/mps.rs:1856-1861): 0 qubits is an error,
  41 and 100 qubits are accepted.
* mps_workspace_over_host_budget_is_insufficient_memory (src/tn/mps.rs:1871-1889), against
  device memory: N chi^2 32 bytes above the free VRAM is OUT_OF_MEMORY, naming the bytes
  required (the reference's mps_workspace_bytes, src/tn/mps.rs:26-33).
* checked_svd_reproduces_the_matrix_nalgebra_got_wrong (src/tn/mps.rs:1999-2028) through the
  device split: the 2 x 2 matrix on which nalgebra returned a top singular value of 1.00795
  is split with top singular value 1 to 1e-12 and a rebuild error below 1e-12, the split's
  checks (norm and probe vectors of svd_reproduces, src/tn/svd_guard.rs:242-283) passing.
* the probe half of svd_reproduces (src/tn/svd_guard.rs:261-280) rejects a factorization that
  the norm half accepts: with the FP64 sweep cap set to 1 from the host (WQC_TN_JACOBI_MAX),
  the rotated rows of a 64-row split are not yet orthogonal, their norms still sum to |A|_F^2
  (the rotation is unitary), and the split must fail with EXEC_FAILED through the probe
  residuals; the same circuit at the default cap runs with no check failure.
"""

import os
import re

import numpy as np
import pytest

import circuits as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wt():
    import wqc_tn

    wqc_tn.load()
    return wqc_tn


def test_register_sizes(wt):
    with pytest.raises(wt.WqcTnError) as e:
        wt.MpsState(0, 8)
    assert e.value.code == wt.INVALID_ARG
    for n, chi in ((41, 8), (100, 128)):
        st = wt.MpsState(n, chi)
        assert st.bond_dims() == [1] * n
        assert st.amplitude_at_compact_zero() == 1.0
        st.free()


def test_workspace_over_device_budget(wt):
    n, chi = 100, 1 << 20
    required = n * chi * chi * 32
    with pytest.raises(wt.WqcTnError) as e:
        wt.MpsState(n, chi)
    assert e.value.code == wt.OUT_OF_MEMORY
    assert str(required) in str(e.value)
    # the reference's own figure for (100, 128): 52 428 800 bytes, well inside the budget
    assert 100 * 128 * 128 * 32 == 52_428_800
    wt.MpsState(100, 128).free()


def test_nalgebra_failure_matrix_through_device_split(wt):
    m = np.array([[4.1435686496782016e-16, -1.597611177955782e-16j],
                  [-0.35975005528623j, 0.9330487113337411]])
    # theta = L . R with L = I (site 0 = (1, 2, 2)) and R = site 1 (2, 2, 1); CZ negates
    # theta[1][1], so R carries -m[1][1] there and the split sees theta' = m.
    r = m.copy()
    r[1, 1] = -r[1, 1]
    st = wt.MpsState(2, 8)
    st.set_site(0, np.eye(2, dtype=complex).reshape(1, 2, 2))
    st.set_site(1, r.reshape(2, 2, 1))
    st.apply_gate(C.two(C.CZ, 0, 1))
    k = st.bond_dims()[0]
    left = st.site(0).reshape(2, k)
    right = st.site(1).reshape(k, 2)
    # sigma_2 of m is ~1e-16, below the 1e-12 cutoff: one singular value is kept
    assert k == 1
    sigma1 = np.linalg.norm(left[:, 0]) ** 2
    assert abs(sigma1 - 1.0) < 1e-12
    assert abs(np.linalg.norm(right[0]) ** 2 - 1.0) < 1e-12
    assert np.max(np.abs(left @ right - m)) < 1e-12
    s = st.stats()
    assert s["svd_check_failures"] == 0 and s["jacobi_unconverged"] == 0


def test_counters_of_a_truncating_run(wt):
    n, chi, gates = C.config2()
    st = wt.MpsState(n, chi)
    st.apply_circuit(gates)
    s = st.stats()
    assert s["svd_check_failures"] == 0
    assert s["jacobi_unconverged"] == 0
    assert s["peak_vram_bytes"] >= sum(st.site(i).nbytes for i in range(n))


def test_probe_check_rejects_an_unconverged_split(wt):
    n, chi = 10, 64
    gates = C.rotation_cz_brickwork(n, 12, seed=314)
    os.environ["WQC_TN_JACOBI_MAX"] = "1"
    try:
        st = wt.MpsState(n, chi)
        with pytest.raises(wt.WqcTnError) as e:
            st.apply_circuit(gates)
        s = st.stats()
        st.free()
    finally:
        del os.environ["WQC_TN_JACOBI_MAX"]
    assert e.value.code == wt.EXEC_FAILED
    msg = str(e.value)
    print(msg, flush=True)
    assert "failed its check" in msg
    g = re.search(r"sum s\^2 = ([0-9.eE+-]+), \|A\|_F\^2 = ([0-9.eE+-]+), probe residuals ([0-9.eE+-]+) "
                  r"([0-9.eE+-]+), job \d+ of \d+, n=(\d+) m=(\d+)", msg)
    assert g, msg
    s2, fro2, r0, r1 = (float(g.group(i)) for i in range(1, 5))
    m = int(g.group(6))
    # the norm half passes: the rows were only rotated
    assert abs(s2 - fro2) <= 1e-9 * fro2
    # the probe half fails: a residual above 1e-9 |A|_F |x| + 1e-14, |x| = sqrt(m)
    assert max(r0, r1) > 1e-9 * np.sqrt(fro2) * np.sqrt(m) + 1e-14
    assert s["svd_check_failures"] >= 1 and s["jacobi_unconverged"] >= 1
    st = wt.MpsState(n, chi)
    st.apply_circuit(gates)
    s = st.stats()
    st.free()
    assert s["svd_check_failures"] == 0 and s["jacobi_unconverged"] == 0
