"""Many independent noisy circuits in one GPU submission.

Each circuit follows the reference's semantics for a circuit with a barrier
after every gate: the gate, then memory noise on every qubit (one level per
gate, dm_simulator.py:1172-1177), and finally an ensemble Z measurement of
all qubits with polarisation retention ``depolarization_factor``.  Small
registers (<= ``lds_max_qubits``) run whole inside one workgroup's LDS; the
others keep their state in HBM and run on parallel streams, as fused tile
programs from 6 qubits up; the batch is captured once into a hipGraph and
replayed (``dm_batch_*`` in dmsim.h).

``autotune`` picks the LDS cutoff width and the workgroup size by timing the
candidates on the batch itself and caches the choice per (precision, CU
count, circuit count, width range).
"""
import ctypes
import itertools
import json
import os

import numpy as np

from . import native, ptm
from .program import Program

_TUNE_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'batch_tuning.json')

# circuits above the LDS cutoff keep their state in HBM; a graph-captured run
# holds at most 16384 launches (dm_batch_run), and large HBM-tier batches stay
# off until the failure of a 512-circuit run is understood
MAX_HBM_CIRCUITS = 12


def circuit_program(n, gates, decay_factor=1.0, decoherence_factor=1.0, thermal_factor=1.0,
                    gate_depolarization=0.0):
    """Program of one circuit: each gate then one memory-noise level."""
    prog = Program(n)
    for name, qs, prm in gates:
        if name in ('u3', 'u1'):
            R = ptm.gate_ptm(name, prm)
            if gate_depolarization:
                R = ptm.depolarizing1(gate_depolarization) @ R
            prog.one_qubit(qs[0], R)
        elif name == 'cx':
            prog.cx(qs[0], qs[1], ptm.depolarizing_pair_scale(gate_depolarization) if gate_depolarization else None)
        else:
            raise ValueError('batch circuits hold u1/u3/cx gates, got %s' % name)
        prog.memory_noise(decoherence_factor, thermal_factor, decay_factor)
    return prog


def fused_circuit_program(n, gates, decay_factor=1.0, decoherence_factor=1.0, thermal_factor=1.0,
                          gate_depolarization=0.0):
    """The same circuit as ``circuit_program`` as fused tile passes: the
    memory noise of every level enters the gate stream as one 1-qubit
    channel per qubit, and the greedy planner groups everything into blocks
    of at most 4 qubits (registers of at least 6 qubits)."""
    from . import fusion
    noisy = []
    mem = ptm.memory_noise(decoherence_factor, thermal_factor, decay_factor)
    identity = decay_factor == 1.0 and decoherence_factor == 1.0
    for gate in gates:
        noisy += fusion.noisy_gates([gate], gate_depolarization)
        if not identity:
            noisy += [fusion.NoisyGate((q,), mem, 'u3') for q in range(n)]
    prog, _ = fusion.fused_program(noisy, n, fusion.KMAX['greedy'], 'greedy', None, 'seq')
    return prog


class Batch:
    """A compiled batch on the device."""

    def __init__(self, circuits, precision='fp64', lds_max_qubits=7, lds_threads=256):
        """``circuits``: list of (n, Program, meas_err)."""
        self.lib = native.load()
        self.ns = [int(c[0]) for c in circuits]
        ops, params, starts = [], [], [0]
        off = 0
        for n, prog, _ in circuits:
            arr, pv = prog.compile()
            for i in range(len(prog)):
                o = native.DmOp()
                ctypes.pointer(o)[0] = arr[i]
                o.param = arr[i].param + off
                ops.append(o)
            params.append(pv)
            off += pv.size
            starts.append(len(ops))
        self._ops = (native.DmOp * max(1, len(ops)))(*ops)
        self._params = np.ascontiguousarray(np.concatenate(params) if params else np.zeros(1))
        self._starts = np.asarray(starts, dtype=np.int64)
        self._n = np.asarray(self.ns, dtype=np.int32)
        self._err = np.asarray([float(c[2]) for c in circuits], dtype=np.float64)
        h = ctypes.c_void_p()
        native.check(self.lib.dm_batch_create(
            native.FP32 if precision == 'fp32' else native.FP64, len(circuits),
            self._n.ctypes.data_as(native.c_int_p), self._starts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            ctypes.cast(self._ops, ctypes.c_void_p), self._params.ctypes.data_as(native.c_double_p),
            int(off), self._err.ctypes.data_as(native.c_double_p), int(lds_max_qubits), int(lds_threads),
            ctypes.byref(h)), 'dm_batch_create')
        self.h = h
        self.nprobs = sum(2 ** n for n in self.ns)

    def run(self, use_graph=True):
        """Returns (list of probability vectors, per-circuit GPU microseconds,
        total milliseconds)."""
        probs = np.empty(self.nprobs)
        us = np.empty(len(self.ns))
        ms = ctypes.c_double(0)
        native.check(self.lib.dm_batch_run(self.h, 1 if use_graph else 0, probs.ctypes.data_as(native.c_double_p),
                                           us.ctypes.data_as(native.c_double_p), ctypes.byref(ms)), 'dm_batch_run')
        out, at = [], 0
        for n in self.ns:
            out.append(probs[at:at + 2 ** n])
            at += 2 ** n
        return out, us, ms.value

    def close(self):
        if getattr(self, 'h', None):
            self.lib.dm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def from_workload(specs, precision='fp64', lds_max_qubits=7, **kw):
    """Batch from dmsim.workloads.batch specs (n, gates, decay_rate, depol_rate):
    decay_factor = 1 - decay_rate, measurement retention = 1 - depol_rate.
    Circuits above the LDS cutoff (and of at least 6 qubits) run as fused
    tile passes."""
    circuits = []
    for n, gates, dr, pr in specs:
        build = fused_circuit_program if (n > lds_max_qubits and n >= 6) else circuit_program
        circuits.append((n, build(n, gates, decay_factor=1.0 - dr), 1.0 - pr))
    return Batch(circuits, precision, lds_max_qubits=lds_max_qubits, **kw)


def autotune(specs, precision='fp64', reps=2, use_cache=True):
    """Best (lds_max_qubits, lds_threads) for this batch by measured time."""
    ns = [sp[0] for sp in specs]
    key = '%s:%s:%d:%d-%d' % (precision, native.device_info(0)['cu_count'], len(specs), min(ns), max(ns))
    cache = {}
    if use_cache and os.path.exists(_TUNE_CACHE):
        try:
            cache = json.load(open(_TUNE_CACHE))
        except (ValueError, OSError):
            cache = {}
        if key in cache:
            return tuple(cache[key]['choice']), cache[key]
    trials = {}
    ns = [sp[0] for sp in specs]
    # a cutoff that leaves more than MAX_HBM_CIRCUITS circuits in HBM is not
    # a candidate (their per-operation launches exceed the graph limit)
    lmaxes = [l for l in (5, 6, 7) if sum(1 for n in ns if n > l) <= MAX_HBM_CIRCUITS] or [7]
    for lmax, thr in itertools.product(lmaxes, (256, 512, 1024)):
        b = from_workload(specs, precision, lds_max_qubits=lmax, lds_threads=thr)
        b.run(True)
        t = min(b.run(True)[2] for _ in range(reps))
        b.close()
        trials['%d/%d' % (lmax, thr)] = t
    best = min(trials, key=trials.get)
    choice = tuple(int(x) for x in best.split('/'))
    rec = {'choice': list(choice), 'trials_ms': trials}
    if use_cache:
        cache[key] = rec
        try:
            json.dump(cache, open(_TUNE_CACHE, 'w'), indent=1)
        except OSError:
            pass
    return choice, rec
