#!/usr/bin/env python3
"""Benchmark: two-qubit gate-merges per second at chi = 512 on one MI355X.

Workload: 100-qubit brickwork, depth 30, RX/RY/RZ (uniform axis, angle U[0, 2pi)) on every
qubit then CZ on alternating even/odd pairs, chi = 512 (1485 two-qubit units). One step is
one pass of the whole circuit from |0...0> through the C ABI (wqc_tn_apply_circuit) plus
the final <0...0|psi>; inputs (the gate list) are on the host, every site tensor lives in
HBM for the whole step.

Printed JSON line: value = merges/s of the faster mixed form (Householder-LQ-preconditioned
Jacobi: up to 5 low sweeps with an FP32 or a split-BF16 Gram, FP64 finish). Riding along: every
form's rate (plain FP64, mixed FP32, mixed split-BF16) and the headline's ratio to FP64;
per form the amplitude error against the complex128 oracle after the range-finder-free first
24 layers (the golden checkpoint), and for layers 25-30 the error of each layer on identical
inputs (the FP64 form's state copied into the form's state before the layer: amplitude and
per-split discarded weight against the FP64 form, whose splits match the oracle's to 1e-13);
the roofline record of the dominant kernel family (Jacobi) and of the one-qubit kernel (GB/s);
the GPU rate on a batch of saturated chi = 512 units taken from the workload's own state; and
the CPU baseline: the FP64 oracle on the same saturated units, one thread (all host cores
with --cpu-all-cores).
`--gpus N` with N > 1 exits: this build runs on one GPU.
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
FORM_LABEL = {
    "fp64": "plain FP64",
    "mixed": "mixed (LQ-preconditioned Jacobi, FP32 early sweeps, FP64 finish)",
    "mixed_bf16": "mixed split-BF16 (LQ-preconditioned Jacobi, early sweeps with split-BF16 Gram on BF16 MFMA, "
                  "FP32 rotations, FP64 finish)",
}
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


def one_qubit_bytes(bonds_before, n):
    """HBM bytes of a layer's one-qubit updates, one per site: read and write of every site
    (l, 2, r) in complex128, 2 * 16 * 2 * l * r."""
    b = [1] + list(bonds_before)
    return sum(64.0 * b[q] * b[q + 1] for q in range(n))


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
    info["one_qubit_bytes"] = 0.0
    for L, lay in enumerate(layers):
        info["one_qubit_bytes"] += one_qubit_bytes(st.bond_dims(), n)
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


def saturated_units(wt, n, chi, gates, count, layers=26):
    """(left site, right site) of `count` saturated units of the workload's own state after
    `layers` layers (plain FP64 form): positions whose left, middle and right bonds are all chi,
    so theta is 2chi x 2chi and the split takes the range finder (ell = chi + 16)."""
    st = wt.MpsState(n, chi)
    for lay in layers_of(gates, n)[:layers]:
        st.apply_circuit(lay)
    b = [1] + st.bond_dims()
    sat = [p for p in range(n - 1) if b[p] == b[p + 1] == b[p + 2] == chi]
    picks = [sat[(k * len(sat)) // count] for k in range(count)] if len(sat) >= count else \
        [sat[k % len(sat)] for k in range(count)]
    pairs = [(st.site(p), st.site(p + 1)) for p in picks]
    st.free()
    return pairs, sat


def unit_chain(pairs, chi):
    """A register holding the units side by side: boundary (1, 2, chi), A_1, B_1, ..., A_K, B_K,
    boundary (chi, 2, 1), and the disjoint layer of CZs on (A_k, B_k)."""
    import numpy as np

    rng = np.random.default_rng(2026)
    b0 = rng.standard_normal((1, 2, chi)) + 1j * rng.standard_normal((1, 2, chi))
    b1 = rng.standard_normal((chi, 2, 1)) + 1j * rng.standard_normal((chi, 2, 1))
    sites = [b0 / np.linalg.norm(b0)]
    for a, b in pairs:
        sites += [a, b]
    sites.append(b1 / np.linalg.norm(b1))
    layer = [C.two(C.CZ, 2 * k + 1, 2 * k + 2) for k in range(len(pairs))]
    return sites, layer


def gpu_saturated_rate(wt, pairs, chi, precision, repeats=3):
    """Units/s of the GPU on one batch of the saturated units (sites uploaded untimed)."""
    sites, layer = unit_chain(pairs, chi)
    best = None
    for _ in range(repeats):
        st = wt.MpsState(len(sites), chi, precision)
        for i, s in enumerate(sites):
            st.set_site(i, s)
        t0 = time.perf_counter()
        st.apply_adjacent_disjoint_gates(layer)
        dt = time.perf_counter() - t0
        st.free()
        best = dt if best is None else min(best, dt)
    return {"units": len(layer), "seconds": best, "units_per_s": len(layer) / best}


def host_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": aff, "cpu_count": os.cpu_count(), "model": model}


def cpu_baseline(pairs, chi, one_thread_units=2, all_cores=False):
    """The FP64 oracle on the host cores, on the saturated units of the workload's own state
    (theta 1024 x 1024, the reference's range finder, ell = 528): one thread on
    `one_thread_units` of them (about 13 s per unit). With all_cores, also one unit per core on
    all host cores (at most 16, the box's CPU share) in one disjoint layer, the oracle running a
    layer's jobs in parallel; that leg takes about two minutes (16 units on 16 cores in 126 s:
    its memory-bound loops scale poorly), so it is off by default."""
    from oracle import oracle_mps as om

    host = host_info()
    cores = max(1, min(16, host["nproc"]))

    def timed(sub, threads):
        om.set_threads(threads)
        sites, layer = unit_chain(sub, chi)
        o = om.OracleMps(len(sites), chi)
        for i, s in enumerate(sites):
            o.set_site(i, s)
        t0 = time.perf_counter()
        o.apply_adjacent_disjoint_gates(layer)
        return len(layer), time.perf_counter() - t0

    u1, t1 = timed(pairs[:one_thread_units], 1)
    out = {"value": u1 / t1, "unit": "merges/s", "cores": 1, "kind": "port",
           "sample": f"{u1} saturated chi=512 units of the workload's own state after layer 26 "
                     f"(theta 1024x1024, range finder ell=528) as one disjoint layer, 1 thread; "
                     f"FP64 oracle ({host['model']}, nproc {host['nproc']})",
           "seconds": t1, "host": host}
    if all_cores:
        ua, ta = timed(pairs[:cores], cores)
        out["all_cores"] = {"value": ua / ta, "units": ua, "seconds": ta, "cores": cores}
    return out


def per_layer_errors(wt, n, chi, gates, forms, first=24):
    """For layers first+1 .. end: every form applies the layer to the plain FP64 form's state
    (copied device to device), and its amplitude and per-split discarded weights are compared
    with the FP64 form's on the same input."""
    layers = layers_of(gates, n)
    f = wt.MpsState(n, chi, wt.PRECISION_FP64)
    for lay in layers[:first]:
        f.apply_circuit(lay)
    out = {name: [] for name, _ in forms}
    for k in range(first, len(layers)):
        got = {}
        for name, prec in forms:
            m = wt.MpsState(n, chi, prec)
            m.copy_sites_from(f)
            m.enable_split_log(True)
            m.apply_circuit(layers[k])
            got[name] = (m.amplitude_at_compact_zero(), m.split_discarded(), m.bond_dims())
            m.free()
        f.enable_split_log(False)
        f.enable_split_log(True)
        f.apply_circuit(layers[k])
        af, dwf, bf = f.amplitude_at_compact_zero(), f.split_discarded(), f.bond_dims()
        for name, _ in forms:
            a, dw, b = got[name]
            dmax = max([abs(x - y) for x, y in zip(dw, dwf)] or [0.0])
            rels = [abs(x - y) / y for x, y in zip(dw, dwf) if y > 1e-14]
            out[name].append({"layer": k + 1, "rel_amp_error": rel(a, af),
                              "max_abs_discarded_weight_error": dmax,
                              "max_rel_discarded_weight_error": max(rels) if rels else 0.0,
                              "bonds_differing": sum(1 for x, y in zip(b, bf) if x != y)})
    f.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--depth", type=int, default=30)
    ap.add_argument("--qubits", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-units", type=int, default=2, help="saturated units of the one-thread CPU leg")
    ap.add_argument("--sat-units", type=int, default=16, help="saturated units of the GPU batch")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the timed steps and the event-timed pass (no per-layer errors, saturated "
                         "batch or CPU leg): the command the rocprofv3 summaries profile")
    ap.add_argument("--cpu-all-cores", action="store_true",
                    help="add the all-core CPU leg (one saturated unit per core, about two minutes)")
    ap.add_argument("--form", default="all", choices=["all", "both", "fp64", "mixed", "mixed_bf16"],
                    help="both: fp64 and mixed; all: fp64, mixed and mixed_bf16")
    ap.add_argument("--profile-pass", action="store_true",
                    help="run exactly one pass per form (for rocprofv3 counter runs), no JSON line")
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy 256 MiB device to device three times: the reference kernel whose "
                         "FETCH_SIZE / WRITE_SIZE readings scale the counter runs")
    ap.add_argument("--slices-pass", action="store_true",
                    help="run the 256-slice batch (widths 12-60, depths 8-32, chi 16-256) once as one "
                         "batch and once more timed (for kernel traces) in each form --form selects, no JSON line")
    ap.add_argument("--backward-pass", action="store_true",
                    help="run the adjoint gradient of the 32-qubit, 8-layer ansatz (512 angles, "
                         "chi 64, checkpoints every 2 runs) once and once more timed, no JSON line")
    args = ap.parse_args()
    if args.copy_probe:
        # Three device-to-device copies of 256 MiB through the HIP runtime (its copy kernel).
        import ctypes as ct

        hip = ct.CDLL("libamdhip64.so")
        nbytes = 256 << 20
        a, b = ct.c_void_p(), ct.c_void_p()
        assert hip.hipMalloc(ct.byref(a), ct.c_size_t(nbytes)) == 0
        assert hip.hipMalloc(ct.byref(b), ct.c_size_t(nbytes)) == 0
        assert hip.hipMemset(a, 1, ct.c_size_t(nbytes)) == 0
        for _ in range(3):
            assert hip.hipMemcpy(b, a, ct.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(a)
        hip.hipFree(b)
        print(f"# copied {nbytes} bytes three times", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding design is documented, not built)", file=sys.stderr)
        sys.exit(2)

    import wqc_tn as wt

    wt.load()
    if args.slices_pass:
        slices = C.config4()
        forms = [("fp64", wt.PRECISION_FP64), ("mixed", wt.PRECISION_MIXED_FP32),
                 ("mixed_bf16", wt.PRECISION_MIXED_BF16)]
        for name, prec in forms:
            if args.form not in ("all", name) and not (args.form == "both" and name != "mixed_bf16"):
                continue
            for label in ("first", "timed"):
                t0 = time.perf_counter()
                amps, codes = wt.run_slices(slices, precision=prec, workers=0)
                print(f"# slice batch {name} ({label}): {len(slices)} slices in {time.perf_counter() - t0:.3f} s, "
                      f"{sum(1 for c in codes if c == wt.OK)} OK", flush=True)
        return
    if args.backward_pass:
        n5, chi5, gates5, _, tf = C.config5()
        terms = [(lab, c) for c, lab in tf]
        for label in ("first", "timed"):
            t0 = time.perf_counter()
            val, grad = wt.backward(n5, chi5, gates5, terms, checkpoint_every=2)
            print(f"# backward ({label}): {len(grad)} gates in {time.perf_counter() - t0:.3f} s, "
                  f"E = {val.real:.12f}", flush=True)
        return
    n, chi, gates = C.config3(n=args.qubits, depth=args.depth)
    units = C.two_qubit_units(gates, n)
    print(f"# workload: {n}q depth {args.depth} chi {chi}: {units} two-qubit units", file=sys.stderr,
          flush=True)
    all_forms = (("fp64", wt.PRECISION_FP64), ("mixed", wt.PRECISION_MIXED_FP32),
                 ("mixed_bf16", wt.PRECISION_MIXED_BF16))

    def wanted(name):
        return args.form == name or args.form == "all" or (args.form == "both" and name != "mixed_bf16")

    if args.profile_pass:
        for name, prec in all_forms:
            if not wanted(name):
                continue
            st = wt.MpsState(n, chi, prec)
            t0 = time.perf_counter()
            st.apply_circuit(gates)
            amp = st.amplitude_at_compact_zero()
            print(f"# {name}: one pass {time.perf_counter() - t0:.2f} s, amplitude {amp}", flush=True)
            st.free()
        return
    forms = {}
    for name, prec in all_forms:
        if not wanted(name):
            continue
        print(f"# form {name}", file=sys.stderr, flush=True)
        forms[name] = run_gpu(wt, n, chi, gates, prec, args.steps, args.warmup)
    # The headline is the faster mixed form (both finish in FP64 and meet the same error bars).
    mixed_names = [k for k in forms if k != "fp64"]
    head = max(mixed_names, key=lambda k: -forms[k]["seconds"]) if mixed_names else "fp64"
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
        "bound": "mfma", "kernel": "k_jacobi_round" + ("_lo + k_jacobi_round" if head != "fp64" else ""),
        "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": (achieved / FP64_PEAK_TFLOPS) if achieved else None, "traffic": traffic,
        "traffic_unit": "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE)",
        "traffic_source": traffic_src,
        "launches": launches.get("jacobi"),
        "avg_launch_us": 1e3 * jac_ms / max(1, launches.get("jacobi", 1)),
        "dominant_family": dom,
    }
    golden24 = None
    gpath = os.path.join(ROOT, "tests", "golden", "config3.json")
    if n == 100 and args.depth == 30 and os.path.exists(gpath):
        with open(gpath) as fh:
            golden24 = complex(*json.load(fh)["checkpoints"]["24"]["amplitude"])
    prec_of = dict(all_forms)
    low_forms = [(k, prec_of[k]) for k in forms if k != "fp64"]
    print("# per-layer errors on identical inputs", file=sys.stderr, flush=True)
    extras = not args.no_extras
    per_layer = per_layer_errors(wt, n, chi, gates, low_forms) if extras and low_forms and args.depth > 24 else {}
    print("# saturated units", file=sys.stderr, flush=True)
    pairs, sat_positions = (saturated_units(wt, n, chi, gates, max(args.sat_units, 16))
                            if extras and n >= 40 else ([], []))
    form_report = {}
    for name, r in forms.items():
        kq = r["kernel_ms"].get("one_qubit", 0.0)
        q_gbs = r["one_qubit_bytes"] / (kq * 1e-3) / 1e9 if kq > 0 else None
        form_report[name] = {
            "merges_per_s": units * args.steps / r["seconds"],
            "amplitude": [r["amplitude"].real, r["amplitude"].imag],
            "prefix24_amplitude": [r["prefix_amplitude"].real, r["prefix_amplitude"].imag],
            "rel_amp_error_vs_oracle_layer24": rel(r["prefix_amplitude"], golden24) if golden24 else None,
            "alg_tflops_whole_step": sum(unit_flops(*e[:6]) for e in r["split_log"]) / r["timed_pass_seconds"] / 1e12,
            "one_qubit": {"bound": "hbm", "achieved": q_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": q_gbs / HBM_PEAK_GBS if q_gbs else None,
                          "bytes_per_pass": r["one_qubit_bytes"], "kernel_ms": kq},
            "kernel_ms": r["kernel_ms"], "launches": r["launches"], "stats": r["stats"]}
        if name in per_layer:
            pl = per_layer[name]
            form_report[name]["identical_input_layers"] = pl
            form_report[name]["max_rel_amp_error_per_layer"] = max(e["rel_amp_error"] for e in pl)
        if pairs:
            form_report[name]["saturated_batch"] = gpu_saturated_rate(wt, pairs[:args.sat_units], chi, prec_of[name])
    out = {
        "metric": "2q gate-merges/s at chi=512 (mixed vs FP64 plain) + rel. amp error vs FP64",
        "value": value, "unit": "merges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32+f64", "mixed_bf16": "bf16x3+f32+f64"}[head],
        "data": "synthetic",
        "config": {"workload": f"rotation+CZ brickwork {n}q depth {args.depth} chi={chi}",
                   "qubits": n, "depth": args.depth, "chi": chi, "units": units,
                   "form": FORM_LABEL[head]},
        "forms": form_report,
        "roofline": roofline,
    }
    if head != "fp64" and "fp64" in forms:
        f, m = forms["fp64"], forms[head]
        out["mixed_over_fp64"] = (units / m["seconds"]) / (units / f["seconds"])
        for k in mixed_names:
            form_report[k]["over_fp64"] = forms["fp64"]["seconds"] / forms[k]["seconds"]
        out["rel_amp_error_vs_fp64"] = {
            "after_layer_24_exact_splits_only": rel(m["prefix_amplitude"], f["prefix_amplitude"]),
            "final_after_range_finder_layers": rel(m["amplitude"], f["amplitude"]),
        }
    if not args.no_cpu_baseline and pairs:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(pairs, chi, args.cpu_units, args.cpu_all_cores)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
