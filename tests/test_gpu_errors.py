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
* the probe residuals of svd_reproduces (src/tn/svd_guard.rs:261-280), as the device computes
  them for every split, are at rounding level and nonzero on rotated splits.
"""

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



@pytest.mark.parametrize("prec", [0, 1])
def test_probe_residuals_are_genuine(wt, prec):
    """The probe half of svd_reproduces is computed from the kept factorization, not from the
    matrix itself: on every unit of a truncating 24-qubit brickwork (chi 48) the relative residual is at rounding level (<= 1e-12, the check's bar
    being 1e-9), and on every unit whose Jacobi rotated a matrix of at least 32 x 32 it is
    nonzero (a residual of exactly zero there would mean the rebuild never left the matrix)."""
    n, chi = 24, 48
    gates = C.rotation_cz_brickwork(n, 24, seed=2718)
    st = wt.MpsState(n, chi, prec)
    st.enable_split_log(True)
    st.apply_circuit(gates)
    log = st.split_log()
    res = st.split_probe_residuals()
    s = st.stats()
    st.free()
    assert len(res) == len(log) == s["two_qubit_units"]
    assert s["svd_check_failures"] == 0
    res = np.array(res)
    big = np.array([min(e[0], e[1]) >= 32 and e[6] >= 2 for e in log])
    print("prec %d: %d units, %d rotated >= 32 x 32, %d range finder; residual max %.2e median %.2e"
          % (prec, len(res), big.sum(), sum(e[3] == 1 for e in log), res.max(), np.median(res[big])), flush=True)
    assert big.sum() >= 30
    assert np.all(np.isfinite(res)) and res.max() <= 1e-12
    assert np.all(res[big] > 0.0)
