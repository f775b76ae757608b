"""Many independent noisy circuits in one GPU submission.

Each circuit follows the reference's semantics for a circuit with a barrier
after every gate: the gate, then memory noise on every qubit (one level per
gate, dm_simulator.py:1172-1177), and finally an ensemble Z measurement of
all qubits with polarisation retention ``depolarization_factor``.  Small
registers (<= ``lds_max_qubits``) run whole inside one workgroup's LDS; the
others keep their state in one HBM arena and run as fused 6-digit tile
passes (7-digit in FP32 where 6 digits would leave contiguous runs of 4^2
or 4^3 coefficients that 7 digits lengthen); round r (pass r of every
circuit) is up to three launches side by side: the initialisation and
probability work, split into at most ``items_per_pass`` work items per
pass, and one launch per tile size with a workgroup per tile; the batch is
captured once into a hipGraph and replayed (``dm_batch_*`` in dmsim.h).

``autotune`` picks the LDS cutoff width, the LDS workgroup size and the
HBM tier's items per pass by timing the candidates on the batch itself; the
choice is kept per (precision, CU count, circuit count, width range) for the
process, and in the JSON file named by DMSIM_TUNE_CACHE when that is set.
"""
import ctypes
import itertools
import json
import os

import numpy as np

from . import native, ptm
from .program import Program

_TUNED = {}

# tile digits of the HBM tier's passes (dm_batch_create takes 6-digit fused
# tiles, and 7-digit ones in FP32)
QUEUE_TILE = 6
QUEUE_TILE_FP32_LONG = 7
# autotuner candidates
LDS_MAX_CHOICES = (5, 6, 7)
LDS_THREAD_CHOICES = (256, 512, 1024)
ITEMS_PER_PASS_CHOICES = (64, 256, 1024)


def native_program(n, gates, decay_factor=1.0, decoherence_factor=1.0, thermal_factor=1.0,
                   gate_depolarization=0.0, fused=False, precision='fp64'):
    """The program of ``circuit_program`` (``fused`` false) or
    ``fused_circuit_program`` (``fused`` true), built by the library on the
    host (dm_batch_circuit_program); the two Python builders below are its
    cross-check."""
    ops, nops, vals = native.batch_circuit_program(n, gates, decay_factor, decoherence_factor, thermal_factor,
                                                   gate_depolarization, fused, precision)
    return Program.from_compiled(n, ops, nops, vals)


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


def batch_tile(n, qs, precision):
    """Tile digits of a batched pass over block ``qs``: 6, or in FP32 7 when
    the 6-digit tile would leave contiguous runs of only 4^2 coefficients
    and 7 digits give longer ones, or runs of 4^3 that 7 digits make 4^4
    (registers of at least 7 qubits; the single-state rule, measured faster
    on the batch too)."""
    from . import fusion
    if precision == 'fp32' and n >= QUEUE_TILE_FP32_LONG and len(qs) > 1:
        pos = [n - 1 - q for q in qs]
        w6 = fusion.run_digits(n, pos, QUEUE_TILE)
        w7 = fusion.run_digits(n, pos, QUEUE_TILE_FP32_LONG)
        if (w6 < 3 and w7 > w6) or (w6 == 3 and w7 == 4):
            return QUEUE_TILE_FP32_LONG
    return QUEUE_TILE


def fused_circuit_program(n, gates, decay_factor=1.0, decoherence_factor=1.0, thermal_factor=1.0,
                          gate_depolarization=0.0, precision='fp64'):
    """The same circuit as ``circuit_program`` as fused tile passes: the
    memory noise of every level enters the gate stream as one 1-qubit
    channel per qubit, and the greedy planner groups everything into blocks
    of at most 4 qubits, each pass a tile of ``batch_tile`` digits
    (registers of at least 6 qubits)."""
    from . import fusion
    noisy = []
    mem = ptm.memory_noise(decoherence_factor, thermal_factor, decay_factor)
    identity = decay_factor == 1.0 and decoherence_factor == 1.0
    for gate in gates:
        noisy += fusion.noisy_gates([gate], gate_depolarization)
        if not identity:
            noisy += [fusion.NoisyGate((q,), mem, 'u3') for q in range(n)]
    blocks = fusion.plan(noisy, n, fusion.KMAX['greedy'], 'greedy', None, tiles=False, precision=precision)
    prog = Program(n)
    for qs, mem_ids in blocks:
        fusion.add_block(prog, noisy, qs, mem_ids, 'seq', batch_tile(n, qs, precision), precision)
    return prog


def concat_programs(programs):
    """The programs' op records and values back to back, as dm_batch_create
    takes them: (records, values, op starts, value count).  Program c's ops
    are records [starts[c], starts[c+1]); their value offsets are shifted
    past the values of the programs before it."""
    recs, params, starts = [], [], [0]
    off = 0
    for prog in programs:
        arr, pv = prog.compile()
        rec = np.frombuffer(arr, dtype=native.DMOP_DTYPE, count=len(prog)).copy()
        rec['param'] += off
        recs.append(rec)
        params.append(pv)
        off += pv.size
        starts.append(starts[-1] + len(prog))
    ops = np.concatenate(recs) if starts[-1] else np.zeros(1, dtype=native.DMOP_DTYPE)
    vals = np.ascontiguousarray(np.concatenate(params) if off else np.zeros(1))
    return ops, vals, np.asarray(starts, dtype=np.int64), off


class Batch:
    """A compiled batch on the device."""

    def __init__(self, circuits, precision='fp64', lds_max_qubits=7, lds_threads=256, items_per_pass=256):
        """``circuits``: list of (n, Program, meas_err)."""
        self.lib = native.load()
        self.ns = [int(c[0]) for c in circuits]
        self._ops, self._params, self._starts, off = concat_programs([c[1] for c in circuits])
        self._n = np.asarray(self.ns, dtype=np.int32)
        self._err = np.asarray([float(c[2]) for c in circuits], dtype=np.float64)
        h = ctypes.c_void_p()
        native.check(self.lib.dm_batch_create(
            native.FP32 if precision == 'fp32' else native.FP64, len(circuits),
            self._n.ctypes.data_as(native.c_int_p), self._starts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            ctypes.c_void_p(self._ops.ctypes.data), self._params.ctypes.data_as(native.c_double_p),
            int(off), self._err.ctypes.data_as(native.c_double_p), int(lds_max_qubits), int(lds_threads),
            int(items_per_pass), ctypes.byref(h)), 'dm_batch_create')
        self.h = h
        self.nprobs = sum(2 ** n for n in self.ns)

    def info(self):
        """HBM-tier circuits, passes and work items, kernel launches per run."""
        v = [ctypes.c_int64(0) for _ in range(4)]
        native.check(self.lib.dm_batch_info(self.h, *[ctypes.byref(x) for x in v]), 'dm_batch_info')
        return dict(zip(('hbm_circuits', 'passes', 'items', 'launches'), (x.value for x in v)))

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


def from_workload(specs, precision='fp64', lds_max_qubits=7, programs=None, **kw):
    """Batch from dmsim.workloads.batch specs (n, gates, decay_rate, depol_rate):
    decay_factor = 1 - decay_rate, measurement retention = 1 - depol_rate.
    Circuits above the LDS cutoff run as fused tile passes.  The programs
    are built by the library (``native_program``).  ``programs``, a dict,
    keeps them per (circuit, form) for later calls on the same specs."""
    return Batch(workload_programs(specs, precision, lds_max_qubits, programs), precision,
                 lds_max_qubits=lds_max_qubits, **kw)


def workload_programs(specs, precision='fp64', lds_max_qubits=7, programs=None):
    """(n, Program, measurement retention) of each spec, as ``from_workload``
    hands them to the batch."""
    circuits = []
    for i, (n, gates, dr, pr) in enumerate(specs):
        fused = n > lds_max_qubits
        key = (i, fused, precision if fused else None)
        prog = programs.get(key) if programs is not None else None
        if prog is None:
            prog = native_program(n, gates, decay_factor=1.0 - dr, fused=fused, precision=precision)
            if programs is not None:
                programs[key] = prog
        circuits.append((n, prog, 1.0 - pr))
    return circuits


def _tune_key(specs, precision):
    ns = [sp[0] for sp in specs]
    return '%s:%s:%d:%d-%d' % (precision, native.device_info(0)['cu_count'], len(specs), min(ns), max(ns))


def autotune(specs, precision='fp64', reps=2, use_cache=True, programs=None):
    """Best (lds_max_qubits, lds_threads, items_per_pass) for this batch by
    measured time: the LDS cutoff and workgroup size first (items per pass
    256), then the items per pass (which splits the initialisation) at that
    cutoff."""
    key = _tune_key(specs, precision)
    path = os.environ.get('DMSIM_TUNE_CACHE')
    if use_cache:
        if key in _TUNED:
            return tuple(_TUNED[key]['choice']), _TUNED[key]
        if path and os.path.exists(path):
            try:
                with open(path) as f:
                    disk = json.load(f)
                if key in disk:
                    _TUNED[key] = disk[key]
                    return tuple(disk[key]['choice']), disk[key]
            except (ValueError, OSError):
                pass
    trials = {}

    programs = {} if programs is None else programs

    def trial(lmax, thr, ipp):
        b = from_workload(specs, precision, lds_max_qubits=lmax, lds_threads=thr, items_per_pass=ipp,
                          programs=programs)
        b.run(True)
        t = min(b.run(True)[2] for _ in range(reps))
        b.close()
        trials['%d/%d/%d' % (lmax, thr, ipp)] = t
        return t

    best = min(((lmax, thr) for lmax, thr in itertools.product(LDS_MAX_CHOICES, LDS_THREAD_CHOICES)),
               key=lambda c: trial(c[0], c[1], 256))
    ipp = min(ITEMS_PER_PASS_CHOICES, key=lambda i: trials.get('%d/%d/%d' % (best + (i,))) or trial(*best, i))
    choice = best + (ipp,)
    rec = {'choice': list(choice), 'trials_ms': trials}
    if use_cache:
        _TUNED[key] = rec
        if path:
            try:
                disk = {}
                if os.path.exists(path):
                    with open(path) as f:
                        disk = json.load(f)
                disk[key] = rec
                with open(path, 'w') as f:
                    json.dump(disk, f, indent=1)
            except (ValueError, OSError):
                pass
    return choice, rec
