"""Writes the golden fixtures of tests/golden/ from the CPU FP64 oracle and numpy.

Run from the repository root: python tests/golden/make_golden.py [name ...]
Fixtures hold inputs by reference (workload generator + seed, which the tests rebuild)
and expected outputs: amplitude <0...0|psi>, bond dimensions and split counters.
Range-finder splits use the oracle's exact-arithmetic rank rule (tau = 1e-10).
"""

import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (ROOT, os.path.join(ROOT, "wqc-core_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import circuits as C  # noqa: E402
import dense_ref as D  # noqa: E402
from oracle import oracle_mps as om  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def layers_of(gates, n):
    out, i = [], 0
    while i < len(gates):
        j = i + n
        while j < len(gates) and gates[j][0] in (C.CZ, C.CNOT):
            j += 1
        out.append(gates[i:j])
        i = j
    return out


def run_oracle(n, chi, gates, checkpoints=()):
    om.set_qr_rank_tol(1e-10)
    o = om.OracleMps(n, chi, record=True)
    t0 = time.time()
    cps = {}
    if checkpoints:
        for L, lay in enumerate(layers_of(gates, n)):
            o.apply_circuit(lay)
            if L + 1 in checkpoints:
                a = o.amplitude_at_compact_zero()
                cps[str(L + 1)] = {"amplitude": [a.real, a.imag], "bonds": o.bond_dims(),
                                   "counters": o.counters()}
                print(f"  layer {L + 1}: {a} {o.counters()} {time.time() - t0:.1f}s", flush=True)
    else:
        o.apply_circuit(gates)
    a = o.amplitude_at_compact_zero()
    paths, chis = o.split_log()
    return {"amplitude": [a.real, a.imag], "bonds": o.bond_dims(), "counters": o.counters(),
            "max_discarded_weight": o.max_discarded(), "split_paths": paths, "split_chis": chis,
            "checkpoints": cps, "oracle_seconds": time.time() - t0,
            "oracle_threads": om.num_threads()}


def write(name, obj):
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(obj, f)
    print("wrote", name, flush=True)


def g_config1():
    n, chi, gates = C.config1()
    psi = D.run(n, gates)
    r = run_oracle(n, chi, gates)
    r["dense_amplitude"] = [psi[0].real, psi[0].imag]
    r["workload"] = "config1"
    write("config1", r)


def g_config2():
    n, chi, gates = C.config2()
    r = run_oracle(n, chi, gates)
    r["workload"] = "config2"
    write("config2", r)


def g_rsvd128():
    gates = C.rotation_cz_brickwork(40, 26, 0x5EED)
    r = run_oracle(40, 128, gates, checkpoints=(16, 20, 23))
    r["workload"] = "rotation_cz_brickwork(40, 26, 0x5EED) chi=128"
    write("rsvd_brickwork40_chi128", r)


def g_config3():
    n, chi, gates = C.config3()
    r = run_oracle(n, chi, gates, checkpoints=(16, 20, 24, 26, 28))
    r["workload"] = "config3"
    write("config3", r)


if __name__ == "__main__":
    names = sys.argv[1:] or ["config1", "config2", "rsvd128", "config3"]
    for nm in names:
        print("==", nm, flush=True)
        globals()["g_" + nm]()
