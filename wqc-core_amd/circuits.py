"""Seeded synthetic circuits for the MPS engine's workloads.

Gates are tuples ``(kind, q0, q1, q2, angle)`` using the reference's gate ids
(``Gate::to_stark_id``, src/engine.rs:284-300): X=1 Y=2 Z=3 H=4 S=5 T=6 CNOT=7 CZ=8
CCNOT=9 RX=10 RY=11 RZ=12, plus MEASURE=13 RESET=14 IF=15.  Two-qubit gates are
``(kind, control, target, 0, 0.0)``.

The random source is SplitMix64 (Steele, Lea & Flood 2014), so the same seed gives the
same circuit in any language.  Uniform doubles are ``(next() >> 11) * 2**-53``.

Workloads:

* ``brickwork``: per layer, every qubit gets a gate drawn uniformly from
  {H,X,Y,Z,T,S,RX,RY,RZ} (angles U[0, 2pi)), then nearest-neighbour entanglers on even
  pairs (even layers) or odd pairs (odd layers), CNOT or CZ with p = 0.5.  With
  ``long_range_p`` > 0 each layer also gets, with that probability, one CNOT on a
  uniformly drawn ordered pair of distinct qubits; the engine routes it by SWAP
  chains (src/tn/mps.rs:353-375).
* ``rotation_cz_brickwork``: RX/RY/RZ (uniform axis, angle U[0, 2pi)) on every qubit,
  then the CZ brickwork.
* ``slice_batch``: many independent brickwork slices of uneven width, depth and chi.
* ``hea_ansatz``: hardware-efficient ansatz, per layer RY, RZ on every qubit
  (angles U[-pi, pi)) then a CZ ladder (i, i+1) for all i.
"""

from __future__ import annotations

import math
from typing import List, Tuple

Gate = Tuple[int, int, int, int, float]

X, Y, Z, H, S, T, CNOT, CZ, CCNOT, RX, RY, RZ, MEASURE, RESET, IF = range(1, 16)

GATE_NAMES = {X: "X", Y: "Y", Z: "Z", H: "H", S: "S", T: "T", CNOT: "CNOT", CZ: "CZ",
              CCNOT: "CCNOT", RX: "RX", RY: "RY", RZ: "RZ", MEASURE: "MEASURE",
              RESET: "RESET", IF: "IF"}
ONE_QUBIT = (H, X, Y, Z, T, S, RX, RY, RZ)
ROTATIONS = (RX, RY, RZ)

_MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator; identical streams in every language for one seed."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def below(self, n: int) -> int:
        return self.next() % n


def one(kind: int, q: int, angle: float = 0.0) -> Gate:
    return (kind, q, 0, 0, float(angle))


def two(kind: int, c: int, t: int) -> Gate:
    return (kind, c, t, 0, 0.0)


def brickwork(n: int, depth: int, seed: int, long_range_p: float = 0.0,
              measure_all: bool = False) -> List[Gate]:
    rng = SplitMix64(seed)
    gates: List[Gate] = []
    for layer in range(depth):
        for q in range(n):
            kind = ONE_QUBIT[rng.below(9)]
            angle = rng.uniform() * 2.0 * math.pi if kind in ROTATIONS else 0.0
            gates.append(one(kind, q, angle))
        for q in range(layer % 2, n - 1, 2):
            gates.append(two(CNOT if (rng.next() & 1) == 0 else CZ, q, q + 1))
        if long_range_p > 0.0 and n >= 2 and rng.uniform() < long_range_p:
            i = rng.below(n)
            j = rng.below(n - 1)
            if j >= i:
                j += 1
            gates.append(two(CNOT, i, j))
    if measure_all:
        gates.extend(one(MEASURE, q) for q in range(n))
    return gates


def rotation_cz_brickwork(n: int, depth: int, seed: int) -> List[Gate]:
    rng = SplitMix64(seed)
    gates: List[Gate] = []
    for layer in range(depth):
        for q in range(n):
            kind = ROTATIONS[rng.below(3)]
            gates.append(one(kind, q, rng.uniform() * 2.0 * math.pi))
        for q in range(layer % 2, n - 1, 2):
            gates.append(two(CZ, q, q + 1))
    return gates


def slice_batch(count: int, seed: int, widths=(12, 60), depths=(8, 32),
                chis=(16, 32, 64, 128, 256)):
    """[(width, chi, gates)] for ``count`` independent brickwork slices."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        w = widths[0] + rng.below(widths[1] - widths[0] + 1)
        d = depths[0] + rng.below(depths[1] - depths[0] + 1)
        chi = chis[rng.below(len(chis))]
        out.append((w, chi, brickwork(w, d, rng.next())))
    return out


def hea_ansatz(n: int, layers: int, seed: int):
    """(gates, angles): RY, RZ on every qubit per layer, then CZ(i, i+1), i < n-1."""
    rng = SplitMix64(seed)
    gates: List[Gate] = []
    angles: List[float] = []
    for _ in range(layers):
        for q in range(n):
            a = (2.0 * rng.uniform() - 1.0) * math.pi
            b = (2.0 * rng.uniform() - 1.0) * math.pi
            angles.extend((a, b))
            gates.append(one(RY, q, a))
            gates.append(one(RZ, q, b))
        for q in range(n - 1):
            gates.append(two(CZ, q, q + 1))
    return gates, angles


def tfim_observable(n: int, field: float = 1.0):
    """sum_i Z_i Z_{i+1} + field * sum_i X_i as (coeff, label) terms, Qiskit order
    (rightmost character = qubit 0, src/expectation.rs:97-105)."""
    terms = []
    for i in range(n - 1):
        lab = ["I"] * n
        lab[n - 1 - i] = "Z"
        lab[n - 1 - (i + 1)] = "Z"
        terms.append((1.0, "".join(lab)))
    for i in range(n):
        lab = ["I"] * n
        lab[n - 1 - i] = "X"
        terms.append((field, "".join(lab)))
    return terms


def two_qubit_units(gates, n: int) -> int:
    """Number of apply_two_qubit_at_positions calls the gate list costs: one per
    adjacent CNOT/CZ, 2*(span-1)+1 per long-range one (src/tn/mps.rs:353-375), and
    6 CNOTs (with their routing) per CCNOT (src/tn/mps.rs:657-677)."""
    def cost(c, t):
        span = abs(c - t)
        return 2 * (span - 1) + 1
    units = 0
    for g in gates:
        k = g[0]
        if k in (CNOT, CZ):
            units += cost(g[1], g[2])
        elif k == CCNOT:
            c1, c2, t = g[1], g[2], g[3]
            units += 2 * cost(c2, t) + 2 * cost(c1, t) + 2 * cost(c1, c2)
    return units


# The five workloads with their fixed seeds.
CONFIG_SEEDS = {1: 0x0C0F_1601, 2: 0x0C0F_4002, 3: 0x0C0F_A003, 4: 0x0C0F_2564, 5: 0x0C0F_3205}


def config1():
    """16 qubits, depth 10, chi <= 32 (truncation-free)."""
    return 16, 32, brickwork(16, 10, CONFIG_SEEDS[1])


def config2():
    """40 qubits, depth 20, long-range CNOTs p = 0.1 per layer, chi = 128, MEASURE all."""
    return 40, 128, brickwork(40, 20, CONFIG_SEEDS[2], long_range_p=0.1, measure_all=True)


def config3(n: int = 100, depth: int = 30):
    """100 qubits, depth 30, RX/RY/RZ + CZ brickwork, chi = 512."""
    return n, 512, rotation_cz_brickwork(n, depth, CONFIG_SEEDS[3])


def config4(count: int = 256):
    return slice_batch(count, CONFIG_SEEDS[4])


def config5():
    """32 qubits, 8 layers RY.RZ + CZ ladder, chi = 64; (n, chi, gates, angles, terms)."""
    gates, angles = hea_ansatz(32, 8, CONFIG_SEEDS[5])
    return 32, 64, gates, angles, tfim_observable(32)
