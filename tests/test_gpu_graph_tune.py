"""Launch-level options of the engine that must not change a result, through the C ABI:

* HIP graphs: the launches of a Jacobi sweep captured once and replayed per sweep, with the
  host's readback, synchronization and done-flag copies between the replays
  (WQC_TN_GRAPH=0 launches them one by one). Every replay must give what the eager launches
  give, bit for bit, in the plain and the mixed forms.
* GEMM tiles: the 64 x 64 and 32 x 32 tile forms of the MFMA GEMM accumulate every output
  element in the same order; forcing either (WQC_TN_GEMM_TILE) gives the same bits.
* The autotuner: with WQC_TN_TUNE_CACHE set, measured tile choices are written to that file and
  read back by the next device handle; results do not depend on the choices.
"""

import os

import numpy as np
import pytest

import circuits as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wt():
    import wqc_tn

    wqc_tn.load()
    return wqc_tn


def run(wt, n, chi, gates, prec, env, device=None):
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        st = wt.MpsState(n, chi, prec, device=device)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    st.apply_circuit(gates)
    out = (st.amplitude_at_compact_zero(), st.bond_dims(), [st.site(i) for i in range(n)])
    st.free()
    return out


def same(a, b):
    return a[0] == b[0] and a[1] == b[1] and all(np.array_equal(x, y) for x, y in zip(a[2], b[2]))


@pytest.mark.parametrize("prec", [0, 1])
def test_sweep_graphs_replay_bit_for_bit(wt, prec):
    for n, chi, gates in (C.config2(), (40, 128, C.rotation_cz_brickwork(40, 24, 0x5EED))):
        eager = run(wt, n, chi, gates, prec, {"WQC_TN_GRAPH": "0", "WQC_TN_GEMM_TILE": "64"})
        graph = run(wt, n, chi, gates, prec, {"WQC_TN_GRAPH": "1", "WQC_TN_GEMM_TILE": "64"})
        assert same(eager, graph), (prec, n, eager[0], graph[0])


@pytest.mark.parametrize("prec", [0, 1])
def test_gemm_tiles_give_the_same_bits(wt, prec):
    n, chi, gates = 40, 128, C.rotation_cz_brickwork(40, 24, 0x5EED)
    t64 = run(wt, n, chi, gates, prec, {"WQC_TN_GEMM_TILE": "64"})
    t32 = run(wt, n, chi, gates, prec, {"WQC_TN_GEMM_TILE": "32"})
    assert same(t64, t32), (prec, t64[0], t32[0])


def test_autotuner_cache_file(wt, tmp_path):
    path = str(tmp_path / "tune.txt")
    n, chi, gates = 40, 128, C.rotation_cz_brickwork(40, 24, 0x5EED)
    os.environ["WQC_TN_TUNE_CACHE"] = path
    try:
        d1 = wt.Device(0)
        tuned = run(wt, n, chi, gates, 1, {}, device=d1)
        lines = open(path).read().split("\n")
        entries = [l for l in lines if l.startswith("gemm ")]
        assert entries and all(l.split()[-1] in ("32", "64") for l in entries)
        d2 = wt.Device(0)  # reads the file: nothing left to measure, no new lines
        again = run(wt, n, chi, gates, 1, {}, device=d2)
        assert open(path).read().split("\n") == lines
    finally:
        os.environ.pop("WQC_TN_TUNE_CACHE", None)
    plain = run(wt, n, chi, gates, 1, {"WQC_TN_AUTOTUNE": "0"})
    assert same(tuned, plain) and same(again, plain)
