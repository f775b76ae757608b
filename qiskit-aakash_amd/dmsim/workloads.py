"""Seeded synthetic workloads (numpy.random.default_rng(seed)).

Each generator returns a list of gates ``(name, qubits, params)`` in program
order; ``to_circuit`` turns one into a QuantumCircuit.

* ``random_u3_cx``   -- n qubits, `ngates` gates: U3 with probability p_u3
  (theta, phi, lam ~ U[0, 2pi) on a uniform qubit), else CX on a uniform
  ordered pair.  (10 qubits / 200 gates / 0.7 is the CPU reference case.)
* ``brickwork_qft``  -- `layers` brickwork layers (U3 on every qubit, CX on
  even then odd neighbour pairs) followed by a QFT (h, cu1(pi/2^k), swaps).
* ``matching_layers`` -- `layers` layers of U3 on every qubit plus CX on a
  random perfect matching.
* ``local_u3_cx``    -- U3 with probability p_u3, else CX on (q, q +- d) with
  d ~ Geometric(0.5) >= 1 (resampled when out of range), random direction.
* ``batch``          -- many independent random_u3_cx circuits of uneven
  width / depth with per-circuit noise rates.
"""
import math

import numpy as np

from .circuit import QuantumCircuit

TWO_PI = 2 * math.pi


def random_u3_cx(n, ngates, seed, p_u3=0.7):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(ngates):
        if rng.random() < p_u3:
            th, ph, la = rng.uniform(0, TWO_PI, 3)
            gates.append(('u3', (int(rng.integers(n)),), (float(th), float(ph), float(la))))
        else:
            c, t = rng.choice(n, 2, replace=False)
            gates.append(('cx', (int(c), int(t)), ()))
    return gates


def brickwork_qft(n, layers, seed):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(layers):
        for q in range(n):
            th, ph, la = rng.uniform(0, TWO_PI, 3)
            gates.append(('u3', (q,), (float(th), float(ph), float(la))))
        for q in range(0, n - 1, 2):
            gates.append(('cx', (q, q + 1), ()))
        for q in range(1, n - 1, 2):
            gates.append(('cx', (q, q + 1), ()))
    for j in range(n):
        gates.append(('h', (j,), ()))
        for k in range(j + 1, n):
            gates.append(('cu1', (k, j), (math.pi / 2 ** (k - j),)))
    for i in range(n // 2):
        gates.append(('swap', (i, n - 1 - i), ()))
    return gates


def matching_layers(n, layers, seed):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(layers):
        for q in range(n):
            th, ph, la = rng.uniform(0, TWO_PI, 3)
            gates.append(('u3', (q,), (float(th), float(ph), float(la))))
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            gates.append(('cx', (int(perm[i]), int(perm[i + 1])), ()))
    return gates


def local_u3_cx(n, ngates, seed, p_u3=0.6):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(ngates):
        if rng.random() < p_u3:
            th, ph, la = rng.uniform(0, TWO_PI, 3)
            gates.append(('u3', (int(rng.integers(n)),), (float(th), float(ph), float(la))))
            continue
        while True:
            q = int(rng.integers(n))
            d = int(rng.geometric(0.5))
            t = q + d if rng.random() < 0.5 else q - d
            if 0 <= t < n:
                break
        if rng.random() < 0.5:
            gates.append(('cx', (q, t), ()))
        else:
            gates.append(('cx', (t, q), ()))
    return gates


def batch(count, seed, n_range=(6, 12), depth_range=(50, 300), rate_range=(1e-4, 1e-2)):
    """(n, gates, decay_rate, depol_rate) for `count` circuits; circuit i uses seed+i."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(seed + i)
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        lo, hi = math.log(rate_range[0]), math.log(rate_range[1])
        decay = math.exp(rng.uniform(lo, hi))
        depol = math.exp(rng.uniform(lo, hi))
        out.append((n, random_u3_cx(n, depth, seed + 100000 + i), decay, depol))
    return out


def to_circuit(n, gates, barrier_after_each=False, measure=None, n_clbits=None):
    """QuantumCircuit from a gate list.  ``measure``: None, ('ensemble', basis)
    or ('partial', qubits, basis)."""
    qc = QuantumCircuit(n, n_clbits if n_clbits is not None else n)
    for name, qs, prm in gates:
        getattr(qc, name)(*prm, *qs)
        if barrier_after_each:
            qc.barrier()
    if measure is not None:
        if measure[0] == 'ensemble':
            qc.measure(0, 0, basis='Ensemble', add_param=measure[1])
        elif measure[0] == 'partial':
            for q in measure[1]:
                qc.measure(q, q, basis=measure[2])
    return qc


def thermal_factors(t1, t2, dt):
    """(decay_factor g, decoherence_factor f) for relaxation T1, dephasing T2
    and gate time dt.  The reference contracts coherences by sqrt(g) * f, so
    f = exp(-dt/T2) / sqrt(g) makes the total coherence decay exp(-dt/T2)."""
    g = math.exp(-dt / t1)
    f = math.exp(-dt / t2) / math.sqrt(g)
    return g, f
