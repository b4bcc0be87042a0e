#!/usr/bin/env python3
"""Benchmark: two-qubit gate-merges per second at chi = 512 on one MI355X.

Workload: 100-qubit brickwork, depth 30, RX/RY/RZ (uniform axis, angle U[0, 2pi)) on every
qubit then CZ on alternating even/odd pairs, chi = 512 (1485 two-qubit units). One step is
one pass of the whole circuit from |0...0> through the C ABI (wqc_tn_apply_circuit) plus
the final <0...0|psi>; inputs (the gate list) are on the host, every site tensor lives in
HBM for the whole step.

Printed JSON line: value = merges/s of the mixed form (Householder-LQ-preconditioned Jacobi:
up to 6 FP32 sweeps, FP64 finish); the plain FP64 form's rate, their ratio, the relative amplitude error between the
forms (after the range-finder-free first 24 layers and at the end), a roofline record for
the dominant kernel and a CPU baseline ride along. `--gpus N` with N > 1 exits: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "wqc-core_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import circuits as C  # noqa: E402

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (matrix and vector)
HBM_PEAK_GBS = 8000.0


def unit_flops(rows, cols, bond, path, ell, chi):
    """Algorithmic flops of one two-qubit unit (complex MAC = 8 flops)."""
    dl, dr = rows // 2, cols // 2
    f = 8.0 * rows * bond * cols + 8.0 * 16 * dl * dr          # merge + U
    if path == 1:                                              # range finder kept
        m, n = max(ell, cols), min(ell, cols)
        f += 8.0 * (4.0 * rows * cols * ell + rows * ell * chi)
        f += 4.0 * (2.0 * rows * ell * ell - (2.0 / 3.0) * ell ** 3)
        f += 4.0 * (4.0 * m * m * n + 8.0 * m * n * n + 9.0 * n ** 3)
    else:                                                      # exact split (GVL R-SVD)
        m, n = max(rows, cols), min(rows, cols)
        f += 4.0 * (4.0 * m * m * n + 8.0 * m * n * n + 9.0 * n ** 3)
    return f


def svd_flops(rows, cols, path, ell):
    if path == 1:
        m, n = max(ell, cols), min(ell, cols)
    else:
        m, n = max(rows, cols), min(rows, cols)
    return 4.0 * (4.0 * m * m * n + 8.0 * m * n * n + 9.0 * n ** 3)


def layers_of(gates, n):
    out, i = [], 0
    while i < len(gates):
        j = i + n
        while j < len(gates) and gates[j][0] in (C.CZ, C.CNOT):
            j += 1
        out.append(gates[i:j])
        i = j
    return out


def run_gpu(wt, n, chi, gates, precision, steps, warmup, prefix_layers=24):
    amps = []
    for _ in range(warmup):
        st = wt.MpsState(n, chi, precision)
        st.apply_circuit(gates)
        amps.append(st.amplitude_at_compact_zero())
        st.free()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = wt.MpsState(n, chi, precision)
        st.apply_circuit(gates)
        amps.append(st.amplitude_at_compact_zero())
        stats = st.stats()
        st.free()
    dt = time.perf_counter() - t0
    info = {"seconds": dt, "amplitude": amps[-1], "stats": stats}
    # A further, untimed pass with HIP-event timing of every kernel family, the split log,
    # and the amplitude after the range-finder-free prefix of the circuit.
    st = wt.MpsState(n, chi, precision)
    st.enable_timing(True)
    st.enable_split_log(True)
    t1 = time.perf_counter()
    layers = layers_of(gates, n)
    for L, lay in enumerate(layers):
        st.apply_circuit(lay)
        if L + 1 == prefix_layers:
            info["prefix_amplitude"] = st.amplitude_at_compact_zero()
    amp = st.amplitude_at_compact_zero()
    info["timed_pass_seconds"] = time.perf_counter() - t1
    info["kernel_ms"] = st.kernel_times()
    info["launches"] = st.kernel_launches()
    info["split_log"] = st.split_log()
    info["timed_amplitude"] = amp
    st.free()
    return info


def profiled_traffic(form):
    """HBM bytes per Jacobi launch from the committed rocprofv3 counter summary of this form
    (profiles/*_traffic.md, FETCH_SIZE + WRITE_SIZE scaled by a copy kernel), or None."""
    import glob
    import re

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.md"))):
        text = open(path).read()
        for block in re.findall(r"```json\s*(\{.*?\})\s*```", text, flags=re.S):
            rec = json.loads(block)
            if rec.get("form") == form and rec.get("workload") == "config3":
                return rec["traffic_bytes_per_launch"], os.path.relpath(path, ROOT)
    return None, None


def rel(a, b):
    return abs(a - b) / abs(b) if b != 0 else float("inf")


def cpu_baseline(sample_units: int = 1):
    """The FP64 oracle on the host cores: saturated chi=512 units (bonds 512 on both sides,
    theta 1024 x 1024, range finder with ell = 528) on synthetic random sites."""
    import numpy as np

    from oracle import oracle_mps as om

    om.set_threads(1)
    rng = np.random.default_rng(2024)
    b = 512
    sites = []
    for l, r in [(1, b), (b, b), (b, b), (b, 1)]:
        t = rng.standard_normal((l, 2, r)) + 1j * rng.standard_normal((l, 2, r))
        t *= np.exp(-0.004 * np.arange(r))[None, None, :]
        t /= np.linalg.norm(t) / math.sqrt(r)
        sites.append(t)
    t0 = time.perf_counter()
    for _ in range(sample_units):
        o = om.OracleMps(4, 512)
        for i, s in enumerate(sites):
            o.set_site(i, s)
        o.apply_gate(C.two(C.CZ, 1, 2))
    dt = time.perf_counter() - t0
    return {"value": sample_units / dt, "unit": "merges/s", "cores": 1, "kind": "port",
            "sample": f"{sample_units} saturated chi=512 unit(s): theta 1024x1024, range finder "
                      f"ell=528, synthetic sites; FP64 oracle, 1 thread ({platform.processor() or 'x86_64'})",
            "seconds": dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--depth", type=int, default=30)
    ap.add_argument("--qubits", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-units", type=int, default=1)
    ap.add_argument("--form", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--profile-pass", action="store_true",
                    help="run exactly one pass per form (for rocprofv3 counter runs), no JSON line")
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy 256 MiB device to device three times: the reference kernel whose "
                         "FETCH_SIZE / WRITE_SIZE readings scale the counter runs")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        x = torch.ones((256 << 20) // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("# copied 268435456 bytes three times", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding design is documented, not built)", file=sys.stderr)
        sys.exit(2)

    import wqc_tn as wt

    wt.load()
    n, chi, gates = C.config3(n=args.qubits, depth=args.depth)
    units = C.two_qubit_units(gates, n)
    print(f"# workload: {n}q depth {args.depth} chi {chi}: {units} two-qubit units", file=sys.stderr,
          flush=True)
    if args.profile_pass:
        for name, prec in (("fp64", wt.PRECISION_FP64), ("mixed", wt.PRECISION_MIXED_FP32)):
            if args.form not in ("both", name):
                continue
            st = wt.MpsState(n, chi, prec)
            t0 = time.perf_counter()
            st.apply_circuit(gates)
            amp = st.amplitude_at_compact_zero()
            print(f"# {name}: one pass {time.perf_counter() - t0:.2f} s, amplitude {amp}", flush=True)
            st.free()
        return
    forms = {}
    for name, prec in (("fp64", wt.PRECISION_FP64), ("mixed", wt.PRECISION_MIXED_FP32)):
        if args.form not in ("both", name):
            continue
        print(f"# form {name}", file=sys.stderr, flush=True)
        forms[name] = run_gpu(wt, n, chi, gates, prec, args.steps, args.warmup)
    head = "mixed" if "mixed" in forms else "fp64"
    main_run = forms[head]
    ms_step = 1e3 * main_run["seconds"] / args.steps
    value = units * args.steps / main_run["seconds"]
    log = main_run["split_log"]
    svd = sum(svd_flops(e[0], e[1], e[3], e[4]) for e in log)
    alg = sum(unit_flops(*e[:6]) for e in log)
    kms = main_run["kernel_ms"]
    launches = main_run["launches"]
    jac_ms = kms.get("jacobi", 0.0)
    dom = max(kms, key=lambda k: kms[k]) if kms else "jacobi"
    achieved = svd / (jac_ms * 1e-3) / 1e12 if jac_ms > 0 else None
    traffic, traffic_src = profiled_traffic(head)
    roofline = {
        "bound": "mfma", "kernel": "k_jacobi_round" + ("_lo + k_jacobi_round" if head == "mixed" else ""),
        "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": (achieved / FP64_PEAK_TFLOPS) if achieved else None, "traffic": traffic,
        "traffic_unit": "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE)",
        "traffic_source": traffic_src,
        "launches": launches.get("jacobi"),
        "avg_launch_us": 1e3 * jac_ms / max(1, launches.get("jacobi", 1)),
        "dominant_family": dom,
    }
    form_report = {}
    for name, r in forms.items():
        form_report[name] = {
            "merges_per_s": units * args.steps / r["seconds"],
            "amplitude": [r["amplitude"].real, r["amplitude"].imag],
            "prefix24_amplitude": [r["prefix_amplitude"].real, r["prefix_amplitude"].imag],
            "alg_tflops_whole_step": sum(unit_flops(*e[:6]) for e in r["split_log"]) / r["timed_pass_seconds"] / 1e12,
            "kernel_ms": r["kernel_ms"], "launches": r["launches"], "stats": r["stats"]}
    out = {
        "metric": "2q gate-merges/s at chi=512 (mixed vs FP64 plain) + rel. amp error vs FP64",
        "value": value, "unit": "merges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"rotation+CZ brickwork {n}q depth {args.depth} chi={chi}",
                   "qubits": n, "depth": args.depth, "chi": chi, "units": units,
                   "form": "mixed (LQ-preconditioned Jacobi, FP32 early sweeps, FP64 finish)" if head == "mixed" else "plain FP64"},
        "forms": form_report,
        "roofline": roofline,
    }
    if "mixed" in forms and "fp64" in forms:
        f, m = forms["fp64"], forms["mixed"]
        out["mixed_over_fp64"] = (units / m["seconds"]) / (units / f["seconds"])
        out["rel_amp_error_vs_fp64"] = {
            "after_layer_24_exact_splits_only": rel(m["prefix_amplitude"], f["prefix_amplitude"]),
            "final_after_range_finder_layers": rel(m["amplitude"], f["amplitude"]),
        }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_units)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
