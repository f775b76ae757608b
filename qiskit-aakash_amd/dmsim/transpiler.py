"""Front end between a QuantumCircuit and the backend's instruction list.

Reproduces what the reference's ``execute()`` does to a circuit before
dm_simulator sees it (execute.py:188-221):

1. unroll every gate to the basis ['u1','u2','u3','cx','id','unitary'] with
   the decomposition rules of qiskit/extensions/standard/*.py
   (transpiler/passes/unroller.py:37-85);
2. drop resets acting on a qubit still in its initial state, repeated until
   the circuit depth stops changing (preset_passmanagers/default.py:95-112,
   transpiler/passes/remove_reset_in_zero_state.py);
3. rebuild the circuit in the order of networkx's lexicographical topological
   sort of the circuit DAG keyed by ``str(qargs)``
   (dagcircuit/dagcircuit.py:774-782, converters/dag_to_circuit.py).  This
   order decides which single-qubit gates merge and how many partition
   levels (hence memory-noise applications) a circuit has;
4. assemble instruction records with integer qubit / clbit indices and
   numeric parameters (assembler/assemble_circuits.py:33-129,
   circuit/instruction.py:168-191).
"""
import heapq
import math
import types

from .circuit import Instruction, QuantumCircuit
from .exceptions import BasicAerError

BASIS = ('u1', 'u2', 'u3', 'cx', 'id', 'unitary')
PASSTHROUGH = ('measure', 'reset', 'barrier', 'snapshot')

PI = math.pi


def _rule(name, params, qs):
    """Decomposition of one gate into (name, params, qubits) steps
    (qiskit/extensions/standard/<gate>.py ``_define``)."""
    p = [float(x) for x in params]
    a = qs[0]
    b = qs[1] if len(qs) > 1 else None
    c = qs[2] if len(qs) > 2 else None
    if name == 'x':
        return [('u3', [PI, 0.0, PI], [a])]
    if name == 'i':
        return [('u3', [PI, 0.0, PI], [a])]
    if name == 'y':
        return [('u3', [PI, PI / 2, PI / 2], [a])]
    if name == 'z':
        return [('u1', [PI], [a])]
    if name == 'h':
        return [('u2', [0.0, PI], [a])]
    if name == 's':
        return [('u1', [PI / 2], [a])]
    if name == 'sdg':
        return [('u1', [-PI / 2], [a])]
    if name == 't':
        return [('u1', [PI / 4], [a])]
    if name == 'tdg':
        return [('u1', [-PI / 4], [a])]
    if name == 'rx':
        return [('u3', [p[0], -PI / 2, PI / 2], [a])]
    if name == 'ry':
        return [('u3', [p[0], 0.0, 0.0], [a])]
    if name == 'rz':
        return [('u1', [p[0]], [a])]
    if name == 'u0':
        return [('u3', [0.0, 0.0, 0.0], [a])]
    if name == 'cz':
        return [('h', [], [b]), ('cx', [], [a, b]), ('h', [], [b])]
    if name == 'cy':
        return [('sdg', [], [b]), ('cx', [], [a, b]), ('s', [], [b])]
    if name == 'ch':
        return [('h', [], [b]), ('sdg', [], [b]), ('cx', [], [a, b]), ('h', [], [b]), ('t', [], [b]),
                ('cx', [], [a, b]), ('t', [], [b]), ('h', [], [b]), ('s', [], [b]), ('x', [], [b]),
                ('s', [], [a])]
    if name == 'crz':
        return [('u1', [p[0] / 2], [b]), ('cx', [], [a, b]), ('u1', [-p[0] / 2], [b]), ('cx', [], [a, b])]
    if name == 'cu1':
        return [('u1', [p[0] / 2], [a]), ('cx', [], [a, b]), ('u1', [-p[0] / 2], [b]), ('cx', [], [a, b]),
                ('u1', [p[0] / 2], [b])]
    if name == 'cu3':
        th, ph, la = p
        return [('u1', [(la - ph) / 2], [b]), ('cx', [], [a, b]), ('u3', [-th / 2, 0.0, -(ph + la) / 2], [b]),
                ('cx', [], [a, b]), ('u3', [th / 2, ph, 0.0], [b])]
    if name == 'swap':
        return [('cx', [], [a, b]), ('cx', [], [b, a]), ('cx', [], [a, b])]
    if name == 'rzz':
        return [('cx', [], [a, b]), ('u1', [p[0]], [b]), ('cx', [], [a, b])]
    if name == 'ccx':
        return [('h', [], [c]), ('cx', [], [b, c]), ('tdg', [], [c]), ('cx', [], [a, c]), ('t', [], [c]),
                ('cx', [], [b, c]), ('tdg', [], [c]), ('cx', [], [a, c]), ('t', [], [b]), ('t', [], [c]),
                ('h', [], [c]), ('cx', [], [a, b]), ('t', [], [a]), ('tdg', [], [b]), ('cx', [], [a, b])]
    if name == 'cswap':
        return [('cx', [], [c, b]), ('ccx', [], [a, b, c]), ('cx', [], [c, b])]
    raise BasicAerError('Cannot unroll the circuit to the given basis, %s. No rule to expand instruction %s.'
                        % (str(list(BASIS)), name))


def unroll(data):
    """Recursively expand non-basis gates (unroller.py:51-84)."""
    out = []
    for inst in data:
        if inst.name in PASSTHROUGH or inst.name in BASIS:
            out.append(inst)
            continue
        steps = _rule(inst.name, inst.params, inst.qargs)
        sub = [Instruction(n, prm, qs, (), inst.condition) for n, prm, qs in steps]
        out.extend(unroll(sub))
    return out


class _Node:
    __slots__ = ('nid', 'inst', 'key', 'succ', 'indeg')

    def __init__(self, nid, inst, key):
        self.nid = nid
        self.inst = inst
        self.key = key
        self.succ = []
        self.indeg = 0


def _wires(inst, creg_bits):
    ws = [('q', q) for q in inst.qargs] + [('c', c) for c in inst.cargs]
    if inst.condition is not None:
        reg = inst.condition[0]
        for b in creg_bits[id(reg)]:
            if ('c', b) not in ws:
                ws.append(('c', b))
    return ws


def _build_dag(circ, ops):
    """Nodes in creation order: input nodes per wire (qubits, then clbits), then
    operations; edges follow each wire."""
    nodes = []
    last = {}
    for q in circ.qubits:
        n = _Node(len(nodes), None, '[]')
        nodes.append(n)
        last[('q', q)] = n
        nodes.append(None)  # the wire's output node (a sink; it never orders operations)
    for c in circ.clbits:
        n = _Node(len(nodes), None, '[]')
        nodes.append(n)
        last[('c', c)] = n
        nodes.append(None)
    creg_bits = {id(r): list(r) for r in circ.cregs}
    op_nodes = []
    for inst in ops:
        key = '[' + ', '.join(repr(q) for q in inst.qargs) + ']'
        n = _Node(len(nodes), inst, key)
        nodes.append(n)
        op_nodes.append(n)
        for w in _wires(inst, creg_bits):
            p = last[w]
            p.succ.append(n)
            n.indeg += 1
            last[w] = n
    return [x for x in nodes if x is not None], op_nodes


def _depth(op_nodes):
    longest = {}
    best = 0
    for n in op_nodes:  # creation order is a topological order
        d = longest.get(n.nid, 0) + 1
        best = max(best, d)
        for s in n.succ:
            if s.inst is not None:
                longest[s.nid] = max(longest.get(s.nid, 0), d)
    return best


def remove_reset_in_zero_state(circ, ops):
    """Drop resets whose predecessor on their qubit is the input node, until the
    depth reaches a fixed point (RemoveResetInZeroState + Depth + FixedPoint)."""
    prev_depth = None
    while True:
        first = {}
        for inst in ops:
            for q in inst.qargs:
                first.setdefault(q, inst)
            for c in inst.cargs:
                first.setdefault(c, inst)
        drop = set()
        for inst in ops:
            if inst.name == 'reset' and first.get(inst.qargs[0]) is inst:
                drop.add(id(inst))
        ops = [o for o in ops if id(o) not in drop]
        _, op_nodes = _build_dag(circ, ops)
        depth = _depth(op_nodes)
        if prev_depth is not None and depth == prev_depth:
            return ops
        prev_depth = depth


def lexicographic_order(circ, ops):
    """Operations in networkx.lexicographical_topological_sort order with key
    str(qargs); ties (only between wire input nodes) broken by creation order."""
    nodes, _ = _build_dag(circ, ops)
    heap = [(n.key, n.nid, n) for n in nodes if n.indeg == 0]
    heapq.heapify(heap)
    out = []
    while heap:
        _, _, n = heapq.heappop(heap)
        if n.inst is not None:
            out.append(n.inst)
        for s in n.succ:
            s.indeg -= 1
            if s.indeg == 0:
                heapq.heappush(heap, (s.key, s.nid, s))
    return out


def transpile(circ):
    """Unroll, drop initial resets and reorder: returns the instruction list."""
    ops = unroll(circ.data)
    ops = remove_reset_in_zero_state(circ, ops)
    return lexicographic_order(circ, ops)


def assemble_experiment(circ, ops=None):
    """Instruction records with the attributes dm_simulator reads."""
    if ops is None:
        ops = transpile(circ)
    qidx = {q: i for i, q in enumerate(circ.qubits)}
    cidx = {c: i for i, c in enumerate(circ.clbits)}
    insts = []
    for inst in ops:
        rec = types.SimpleNamespace(name=inst.name)
        if inst.params:
            rec.params = [float(x) if isinstance(x, (int, float)) and not isinstance(x, bool) else x
                          for x in inst.params]
        if inst.name != 'barrier' or inst.qargs:
            rec.qubits = [qidx[q] for q in inst.qargs]
        if inst.cargs:
            rec.memory = [cidx[c] for c in inst.cargs]
        if inst.condition is not None:
            rec.conditional = inst.condition
        insts.append(rec)
    header = types.SimpleNamespace(name=circ.name, n_qubits=circ.num_qubits, memory_slots=circ.num_clbits)
    header.as_dict = lambda h=header: {'name': h.name, 'n_qubits': h.n_qubits, 'memory_slots': h.memory_slots}
    config = types.SimpleNamespace(n_qubits=circ.num_qubits, memory_slots=circ.num_clbits)
    return types.SimpleNamespace(instructions=insts, header=header, config=config)


def assemble(circuits, qobj_id=None):
    """A minimal in-memory qobj: .experiments, .config, .header, .qobj_id."""
    import uuid
    if isinstance(circuits, QuantumCircuit):
        circuits = [circuits]
    exps = [assemble_experiment(c) for c in circuits]
    cfg = types.SimpleNamespace(n_qubits=max([e.config.n_qubits for e in exps] or [0]),
                                memory_slots=max([e.config.memory_slots for e in exps] or [0]),
                                shots=1)
    header = types.SimpleNamespace()
    header.as_dict = lambda: {}
    return types.SimpleNamespace(experiments=exps, config=cfg, header=header,
                                 qobj_id=qobj_id or str(uuid.uuid4()))
