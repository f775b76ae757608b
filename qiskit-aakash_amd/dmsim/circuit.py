"""A small circuit API with the call signatures the reference's users write.

Covers what reaches the dm_simulator path: registers, the standard gates
(qiskit/extensions/standard/*.py), ``measure(qubit, cbit, basis=None,
add_param=None)`` with the fork's bases (qiskit/circuit/measure.py:23-97),
``reset``, ``barrier`` and ``c_if``.  Gates keep their own name and
parameters; ``dmsim.transpiler`` unrolls them with the reference's
decomposition rules.
"""
import itertools
import math

import numpy as np

from .exceptions import BasicAerError


class Register:
    prefix = 'reg'
    bit_type = None
    _counter = itertools.count()

    def __init__(self, size, name=None):
        if int(size) != size or size <= 0:
            raise BasicAerError('Register size must be a positive integer')
        self.size = int(size)
        self.name = name if name is not None else '%s%i' % (self.prefix, next(self._counter))
        self._bits = [self.bit_type(self, i) for i in range(self.size)]

    def __getitem__(self, key):
        if isinstance(key, slice):
            return self._bits[key]
        if isinstance(key, (list, tuple)):
            return [self._bits[k] for k in key]
        if key < 0 or key >= self.size:
            raise BasicAerError('register index out of range')
        return self._bits[key]

    def __len__(self):
        return self.size

    def __iter__(self):
        return iter(self._bits)

    def __repr__(self):
        return "%s(%d, '%s')" % (type(self).__qualname__, self.size, self.name)


class Bit:
    def __init__(self, register, index):
        self.register = register
        self.index = index

    def __repr__(self):
        return '%s(%s, %s)' % (type(self).__name__, self.register, self.index)

    def __eq__(self, other):
        return type(self) is type(other) and self.register is other.register and self.index == other.index

    def __hash__(self):
        return hash((id(self.register), self.index))


class Qubit(Bit):
    pass


class Clbit(Bit):
    pass


class QuantumRegister(Register):
    prefix = 'q'
    bit_type = Qubit
    _counter = itertools.count()


class ClassicalRegister(Register):
    prefix = 'c'
    bit_type = Clbit
    _counter = itertools.count()


class Instruction:
    """One circuit operation: name, parameters, qubits, clbits, condition."""

    __slots__ = ('name', 'params', 'qargs', 'cargs', 'condition')

    def __init__(self, name, params, qargs, cargs=(), condition=None):
        self.name = name
        self.params = list(params)
        self.qargs = list(qargs)
        self.cargs = list(cargs)
        self.condition = condition

    def c_if(self, creg, val):
        self.condition = (creg, int(val))
        return self

    def __repr__(self):
        return 'Instruction(%s, %s, %s)' % (self.name, self.params, [str(q) for q in self.qargs])


class InstructionSet(list):
    def c_if(self, creg, val):
        for inst in self:
            inst.c_if(creg, val)
        return self


_MEASURE_BASES = ('I', 'X', 'Y', 'Z', 'N', 'Bell', 'Ensemble', 'Expect')


def _measure_params(basis, add_param):
    """Parameters of the fork's Measure(basis, add_param) (measure.py:28-60)."""
    if basis == 'N':
        if add_param is None:
            raise BasicAerError('Direction should be provided with this measurement basis.')
        return ['N', add_param]
    if basis == 'Bell':
        if add_param is None:
            raise BasicAerError('Bell measurement should be provided with two qubit locations')
        return ['Bell', str(add_param)]
    if basis == 'Ensemble':
        return ['Ensemble', add_param if add_param is not None else 'Z']
    if basis == 'Expect':
        if add_param is None:
            raise BasicAerError('Expectation measurement should be provided with string of Pauli operators')
        return ['Expect', str(add_param)]
    if basis is not None and add_param is None:
        if basis[0] not in _MEASURE_BASES:
            raise BasicAerError('Invalid basis provided.')
        return [basis[0]]
    if basis is not None and add_param is not None:
        raise BasicAerError('Direction cannot be provided with this measurement basis.')
    return []


class QuantumCircuit:
    """Quantum circuit with the reference's construction API (subset)."""

    _counter = itertools.count()

    def __init__(self, *regs, name=None):
        self.qregs, self.cregs = [], []
        self.data = []
        self.name = name if name is not None else 'circuit%i' % next(self._counter)
        if len(regs) >= 1 and all(isinstance(r, int) for r in regs):
            nq = regs[0]
            nc = regs[1] if len(regs) > 1 else 0
            if nq:
                self.qregs.append(QuantumRegister(nq, 'q'))
            if nc:
                self.cregs.append(ClassicalRegister(nc, 'c'))
        else:
            for r in regs:
                if isinstance(r, QuantumRegister):
                    self.qregs.append(r)
                elif isinstance(r, ClassicalRegister):
                    self.cregs.append(r)
                else:
                    raise BasicAerError('QuantumCircuit expects registers or sizes')

    # bits -----------------------------------------------------------------------
    @property
    def qubits(self):
        return [b for r in self.qregs for b in r]

    @property
    def clbits(self):
        return [b for r in self.cregs for b in r]

    @property
    def num_qubits(self):
        return sum(r.size for r in self.qregs)

    @property
    def num_clbits(self):
        return sum(r.size for r in self.cregs)

    def _bits(self, arg, kind):
        flat = self.qubits if kind == 'q' else self.clbits
        if isinstance(arg, Register):
            return list(arg)
        if isinstance(arg, Bit):
            return [arg]
        if isinstance(arg, (int, np.integer)):
            if arg < 0 or arg >= len(flat):
                raise BasicAerError('%s index %d out of range' % ('qubit' if kind == 'q' else 'clbit', arg))
            return [flat[int(arg)]]
        if isinstance(arg, slice):
            return flat[arg]
        if isinstance(arg, (list, tuple, range)):
            out = []
            for a in arg:
                out.extend(self._bits(a, kind))
            return out
        raise BasicAerError('bad %s argument %r' % ('qubit' if kind == 'q' else 'clbit', arg))

    def append(self, inst):
        self.data.append(inst)
        return inst

    def _gate(self, name, params, *qargs):
        lists = [self._bits(a, 'q') for a in qargs]
        width = max(len(x) for x in lists)
        for x in lists:
            if len(x) not in (1, width):
                raise BasicAerError('register size error')
        out = InstructionSet()
        for i in range(width):
            qs = [x[i] if len(x) > 1 else x[0] for x in lists]
            if len(set(qs)) != len(qs):
                raise BasicAerError('duplicate qubit arguments')
            out.append(self.append(Instruction(name, params, qs)))
        return out if width > 1 else out[0]

    # standard gates (qiskit/extensions/standard) ----------------------------------
    def u1(self, theta, q):
        return self._gate('u1', [theta], q)

    def u2(self, phi, lam, q):
        return self._gate('u2', [phi, lam], q)

    def u3(self, theta, phi, lam, q):
        return self._gate('u3', [theta, phi, lam], q)

    def u0(self, m, q):
        return self._gate('u0', [m], q)

    def iden(self, q):
        return self._gate('id', [], q)

    def i(self, q):
        return self._gate('i', [], q)

    def x(self, q):
        return self._gate('x', [], q)

    def y(self, q):
        return self._gate('y', [], q)

    def z(self, q):
        return self._gate('z', [], q)

    def h(self, q):
        return self._gate('h', [], q)

    def s(self, q):
        return self._gate('s', [], q)

    def sdg(self, q):
        return self._gate('sdg', [], q)

    def t(self, q):
        return self._gate('t', [], q)

    def tdg(self, q):
        return self._gate('tdg', [], q)

    def rx(self, theta, q):
        return self._gate('rx', [theta], q)

    def ry(self, theta, q):
        return self._gate('ry', [theta], q)

    def rz(self, phi, q):
        return self._gate('rz', [phi], q)

    def cx(self, ctl, tgt):
        return self._gate('cx', [], ctl, tgt)

    cnot = cx

    def cy(self, ctl, tgt):
        return self._gate('cy', [], ctl, tgt)

    def cz(self, ctl, tgt):
        return self._gate('cz', [], ctl, tgt)

    def ch(self, ctl, tgt):
        return self._gate('ch', [], ctl, tgt)

    def crz(self, theta, ctl, tgt):
        return self._gate('crz', [theta], ctl, tgt)

    def cu1(self, theta, ctl, tgt):
        return self._gate('cu1', [theta], ctl, tgt)

    def cu3(self, theta, phi, lam, ctl, tgt):
        return self._gate('cu3', [theta, phi, lam], ctl, tgt)

    def swap(self, q1, q2):
        return self._gate('swap', [], q1, q2)

    def rzz(self, theta, q1, q2):
        return self._gate('rzz', [theta], q1, q2)

    def ccx(self, c1, c2, tgt):
        return self._gate('ccx', [], c1, c2, tgt)

    toffoli = ccx

    def cswap(self, ctl, t1, t2):
        return self._gate('cswap', [], ctl, t1, t2)

    fredkin = cswap

    # non-unitary --------------------------------------------------------------------
    def barrier(self, *qargs):
        qs = self.qubits if not qargs else self._bits(list(qargs), 'q')
        return self.append(Instruction('barrier', [], qs))

    def reset(self, q):
        out = InstructionSet()
        for qb in self._bits(q, 'q'):
            out.append(self.append(Instruction('reset', [], [qb])))
        return out if len(out) > 1 else out[0]

    def measure(self, qubit, cbit, basis=None, add_param=None):
        """Measure with the fork's bases; Bell / Ensemble / Expect are wrapped in
        barriers (measure.py:76-97)."""
        params = _measure_params(basis, add_param)
        qs = self._bits(qubit, 'q')
        cs = self._bits(cbit, 'c')
        if len(qs) == len(cs):
            pairs = list(zip(qs, cs))
        elif len(qs) == 1 and cs:
            pairs = [(qs[0], c) for c in cs]
        else:
            raise BasicAerError('register size error')
        wrap = basis in ('Ensemble', 'Expect', 'Bell')
        if wrap:
            self.barrier()
        out = InstructionSet()
        for qb, cb in pairs:
            out.append(self.append(Instruction('measure', list(params), [qb], [cb])))
        if wrap:
            self.barrier()
            return None
        return out if len(out) > 1 else out[0]

    # misc -----------------------------------------------------------------------------
    def size(self):
        return sum(1 for d in self.data if d.name != 'barrier')

    def __repr__(self):
        return 'QuantumCircuit(%s, %d qubits, %d ops)' % (self.name, self.num_qubits, len(self.data))


pi = math.pi
