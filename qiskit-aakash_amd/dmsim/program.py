"""Op lists for ``dm_run_program``: build once, run many times."""
import numpy as np

from . import native, ptm


class Program:
    """Accumulates operations and their parameters for one register size."""

    def __init__(self, n):
        self.n = n
        self._ops = []
        self._vals = []
        self._size = 0
        self._compiled = None

    def __len__(self):
        return len(self._ops)

    @classmethod
    def from_compiled(cls, n, ops, nops, params):
        """A program from ``nops`` DmOp records (a ctypes array) and their
        values, built elsewhere (dm_batch_circuit_program)."""
        prog = cls(n)
        rec = np.frombuffer(ops, dtype=native.DMOP_DTYPE, count=nops)
        prog._ops = list(zip(rec['kind'].tolist(), rec['k'].tolist(), rec['qubits'].tolist(), rec['param'].tolist()))
        vals = np.ascontiguousarray(params, dtype=np.float64)
        prog._vals = [vals]
        prog._size = vals.size
        prog._compiled = (ops, vals)
        return prog

    def _add(self, kind, qubits, values=None, k=None):
        off = self._size
        if values is not None:
            v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
            self._vals.append(v)
            self._size += v.size
        qs = list(qubits) + [0] * (4 - len(qubits))
        self._ops.append((kind, k if k is not None else len(qubits), qs, off))
        self._compiled = None

    def one_qubit(self, q, R4):
        self._add(native.OP_1Q, [q], ptm.rows3(R4))

    def cx(self, control, target, scale16=None):
        if scale16 is None:
            self._add(native.OP_CX, [control, target])
        else:
            self._add(native.OP_CX_SCALED, [control, target], scale16)

    def two_qubit(self, qa, qb, R16):
        self._add(native.OP_2Q, [qa, qb], R16)

    def block(self, qubits, S):
        self._add(native.OP_BLOCK, list(qubits), S, k=len(qubits))

    def memory_noise(self, f, p, g):
        self._add(native.OP_MEMORY_NOISE, [], [f, p, g])

    def diag(self, tab_nx4):
        self._add(native.OP_DIAG, [], tab_nx4)

    def fused(self, m, qubits, subops):
        """One tile pass: ``subops`` is a list of (kind, qubits, values) with
        kind native.SUB_1Q / SUB_CX / SUB_2Q / SUB_DENSE (see DM_OP_FUSED)."""
        rec = [float(m), float(len(qubits))] + [float(q) for q in qubits] + [float(len(subops))]
        for kind, qs, vals in subops:
            v = [] if vals is None else list(np.asarray(vals, dtype=np.float64).reshape(-1))
            q4 = list(qs) + [0] * (4 - len(qs))
            rec += [float(kind), float(len(qs))] + [float(q) for q in q4] + [float(len(v))] + v
        self._add(native.OP_FUSED, [], rec, k=0)

    def kinds(self):
        return [o[0] for o in self._ops]

    def compile(self):
        if self._compiled is None:
            arr = (native.DmOp * max(1, len(self._ops)))()
            for i, (kind, k, qs, off) in enumerate(self._ops):
                arr[i].kind = kind
                arr[i].k = k
                for j in range(4):
                    arr[i].qubits[j] = qs[j]
                arr[i].param = off
            params = np.concatenate(self._vals) if self._vals else np.zeros(0)
            self._compiled = (arr, params)
        return self._compiled

    def run(self, state):
        if not self._ops:
            return
        arr, params = self.compile()
        ops = (native.DmOp * len(self._ops)).from_buffer(arr)
        state.run_program(ops, params)
