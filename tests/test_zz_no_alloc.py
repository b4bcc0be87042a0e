"""wqc_tn_amplitude_at_compact_zero_noalloc against wqc_tn_amplitude_at_compact_zero, through the
C ABI: the same kernels on the same state give the same value bit for bit, on two calls in a
row (the second on a smaller state, whose vectors are shorter than what the shared scratch
array still holds from the first), and the allocation-free call grows no buffer of the state.
A bond beyond what the scratch array holds is refused with INVALID_ARG."""

import numpy as np
import pytest

import circuits as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wt():
    import wqc_tn

    wqc_tn.load()
    return wqc_tn


def layers_of(gates, n):
    out, i = [], 0
    while i < len(gates):
        j = i + n
        while j < len(gates) and gates[j][0] in (C.CZ, C.CNOT):
            j += 1
        out.append(gates[i:j])
        i = j
    return out


def test_noalloc_amplitude_matches_and_allocates_nothing(wt):
    n, chi, gates = C.config3()
    big = wt.MpsState(n, chi)
    for lay in layers_of(gates, n)[:22]:
        big.apply_circuit(lay)
    assert max(big.bond_dims()) >= 256
    n2, chi2, gates2 = C.config1()
    small = wt.MpsState(n2, chi2)
    small.apply_circuit(gates2)
    for st in (big, small, big):
        want = st.amplitude_at_compact_zero()
        got = st.amplitude_at_compact_zero_noalloc()
        print("bonds up to", max(st.bond_dims()), "amplitude", want, "noalloc", got, flush=True)
        assert got == want
    # A state whose readout workspace was never sized: the allocation-free call leaves the
    # footprint where it is, the ordinary one grows it.
    fresh = wt.MpsState(40, 64)
    fresh.apply_circuit([C.one(C.H, q) for q in range(40)])
    peak0 = fresh.peak_vram_bytes()
    a = fresh.amplitude_at_compact_zero_noalloc()
    assert fresh.peak_vram_bytes() == peak0
    assert abs(a - 2.0 ** -20) <= 1e-15
    assert fresh.amplitude_at_compact_zero() == a
    assert fresh.peak_vram_bytes() > peak0


def test_noalloc_refuses_bonds_beyond_the_scratch(wt):
    st = wt.MpsState(2, 2048)
    b = 1500
    rng = np.random.default_rng(3)
    s0 = rng.standard_normal((1, 2, b)) + 1j * rng.standard_normal((1, 2, b))
    s1 = rng.standard_normal((b, 2, 1)) + 1j * rng.standard_normal((b, 2, 1))
    st.set_site(0, s0)
    st.set_site(1, s1)
    with pytest.raises(wt.WqcTnError) as e:
        st.amplitude_at_compact_zero_noalloc()
    assert e.value.code == wt.INVALID_ARG and "scratch" in str(e.value)
    want = complex(np.sum(s0[0, 0, :] * s1[:, 0, 0]))
    got = st.amplitude_at_compact_zero()
    assert abs(got - want) <= 1e-12 * abs(want)
