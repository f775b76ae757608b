"""Host scheduling: single-qubit gate merging and level partitioning.

The level structure decides how often the per-level memory noise is applied,
so both passes keep the reference's semantics exactly.  ``merge_single_gates``
and ``partition`` run the native front end (dm_merge_rotations,
dm_partition in libdmsim); ``merge_single_gates_py`` and ``partition_py``
are the same passes in Python, kept as the cross-check of the native ones:

* ``merge_single_gates`` -- basicaertools.py:129-307 (U3_merge, mergeU,
  merge_gates, single_gate_merge).  Runs of u1/u2/u3 on a qubit between
  separators (cx, measure, bfunc, reset, barrier) fold into one gate; every
  separator flushes all pending runs in qubit order; u2 becomes u3(pi/2, ...);
  id / u0 are dropped.  The u3*u3 fold evaluates the same numpy expressions
  in the same order, so merged angles are bit-identical.
* ``partition`` -- basicaertools.py:536-748 (qubit_stack, partition_helper,
  partition).  Per-qubit stacks are consumed greedily; a CX is placed when it
  heads both of its stacks; measures (resets) are placed together once every
  stack head is a measure (reset) or a placeholder of one; barriers split
  segments; a segment opened by a Bell / Expect / Ensemble measure is one
  level of its own.
"""
import collections
import copy
import itertools

import numpy as np

from .exceptions import BasicAerError

SEPARATORS = frozenset(('CX', 'cx', 'measure', 'bfunc', 'reset', 'barrier'))
ROTATIONS = frozenset(('U', 'u1', 'u2', 'u3'))
OWN_LEVEL = ('Bell', 'Expect', 'Ensemble')


def zyz_from_yzy(xi, th1, th2):
    """Rz(alpha) Ry(beta) Rz(gamma) equal (up to phase) to
    Ry(th1) Rz(xi) Ry(th2); returns (beta, alpha, gamma)."""
    s_x, c_x = np.sin(xi * 0.5), np.cos(xi * 0.5)
    s_p, c_p = np.sin((th1 + th2) * 0.5), np.cos((th1 + th2) * 0.5)
    s_m, c_m = np.sin((th1 - th2) * 0.5), np.cos((th1 - th2) * 0.5)
    t_sum = np.arctan2(s_x * c_m, c_x * c_p)
    t_dif = np.arctan2(-s_x * s_m, c_x * s_p)
    beta = 2 * np.arccos(np.sqrt((c_x * c_p) ** 2 + (s_x * c_m) ** 2))
    return beta, t_sum + t_dif, t_sum - t_dif


def _fold(first, second):
    """Fold two consecutive rotations (each (record, position)); the result
    reuses the earlier record (smaller position)."""
    a, ia = first
    b, ib = second
    out = copy.deepcopy(a if ia < ib else b)
    if a.name == 'u1' and b.name == 'u1':
        out.params[0] = a.params[0] + b.params[0]
    elif a.name == 'u1' and b.name == 'u3':
        if out.name == 'u1':
            out.name = 'u3'
            out.params += [0, 0]
        out.params[0] = b.params[0]
        out.params[1] = b.params[1]
        out.params[2] = b.params[2] + a.params[0]
    elif a.name == 'u3' and b.name == 'u1':
        if out.name == 'u1':
            out.name = 'u3'
            out.params += [0, 0]
        out.params[0] = a.params[0]
        out.params[1] = a.params[1] + b.params[0]
        out.params[2] = a.params[2]
    elif a.name == 'u3' and b.name == 'u3':
        beta, alpha, gamma = zyz_from_yzy(float(b.params[2] + a.params[1]), float(b.params[0]),
                                          float(a.params[0]))
        out.params[0] = beta
        out.params[1] = b.params[1] + alpha
        out.params[2] = a.params[2] + gamma
    else:
        raise BasicAerError('Encountered unrecognized instructions: %s, %s' % (a.name, b.name))
    return out, min(ia, ib)


def _fold_run(run):
    if len(run) == 1:
        return run[0][0]
    acc = _fold(run[0], run[1])
    for item in run[2:]:
        acc = _fold(acc, item)
    return acc[0]


def _normalise(op):
    if op.name == 'CX':
        op.name = 'cx'
    elif op.name == 'U':
        op.name = 'u3'
    elif op.name == 'u2':
        op.name = 'u3'
        op.params.insert(0, np.pi / 2)


def _merge_kind(op):
    if op.name in ROTATIONS:
        return 0
    if op.name in SEPARATORS:
        return 1
    if op.name in ('id', 'u0'):
        return 2
    raise BasicAerError('Encountered unrecognized instruction: %s' % op.name)


def merge_single_gates(insts, n, merge=True):
    """Merged instruction list through the native front end: each output
    rotation is a copy of the first record of its run with the folded name
    and angles (merged angles from the C library's trigonometry, within a
    few ulp of numpy's); separators are the input records, normalised.
    Records are normalised in place, as in the reference (pass copies)."""
    from . import native
    kinds = [_merge_kind(op) for op in insts]
    for op, k in zip(insts, kinds):
        if k == 0:
            _normalise(op)
    kc = [(native.INST_U1 if op.name == 'u1' else native.INST_U3) if k == 0 else
          (native.INST_SEPARATOR if k == 1 else native.INST_ID) for op, k in zip(insts, kinds)]
    qs = [op.qubits[0] if k == 0 else 0 for op, k in zip(insts, kinds)]
    prm = np.zeros((len(insts), 3))
    for i, (op, k) in enumerate(zip(insts, kinds)):
        if k == 0:
            v = [float(x) for x in op.params]
            prm[i, :len(v)] = v
    src, okind, oprm = native.merge_rotations(kc, qs, prm, merge)
    out = []
    for j, i in enumerate(src):
        op = insts[i]
        if okind[j] == native.INST_SEPARATOR:
            _normalise(op)
            out.append(op)
            continue
        rec = copy.deepcopy(op)
        if okind[j] == native.INST_U1:
            rec.name = 'u1'
            rec.params = [float(oprm[j, 0])]
        else:
            rec.name = 'u3'
            rec.params = [float(x) for x in oprm[j]]
        out.append(rec)
    return out


def merge_single_gates_py(insts, n, merge=True):
    """Merged instruction list (records are modified in place, as in the
    reference; pass copies)."""
    out = []
    if not merge:
        for op in insts:
            _normalise(op)
            if op.name not in ('id', 'u0'):
                out.append(op)
        return out
    pending = collections.defaultdict(list)
    for pos, op in enumerate(insts):
        if op.name in SEPARATORS:
            for q in sorted(pending):
                out.append(_fold_run(pending[q]))
            pending.clear()
            _normalise(op)
            out.append(op)
        elif op.name in ROTATIONS:
            _normalise(op)
            pending[op.qubits[0]].append((op, pos))
        elif op.name in ('id', 'u0'):
            continue
        else:
            raise BasicAerError('Encountered unrecognized instruction: %s' % op.name)
    for q in sorted(pending):
        out.append(_fold_run(pending[q]))
    return out


class _Placeholder:
    __slots__ = ('name',)

    def __init__(self, kind):
        self.name = kind


def _stacks(seg, n):
    """Per-qubit stacks; a measure (reset) group leaves one entry (the op or a
    placeholder) on every qubit's stack."""
    st = [collections.deque() for _ in range(n)]
    for op in seg:
        if op.name == 'bfunc' or not hasattr(op, 'qubits'):
            raise BasicAerError('dm_simulator cannot partition "%s": conditional operations are not '
                                'supported on this backend' % op.name)
        if op.name not in ('measure', 'reset'):
            for q in op.qubits:
                st[q].append(op)
            continue
        hold = 'hold_' + op.name
        q = op.qubits[0]
        if st[q] and st[q][-1].name == hold:
            st[q][-1] = op
        else:
            st[q].append(op)
            ph = _Placeholder(hold)
            for other in range(n):
                if other != q:
                    st[other].append(ph)
    return st


def _level_segment(seg, n):
    """Levels of one barrier-free segment (partition_helper semantics).
    Returns (levels, count); ``levels`` is padded with empty lists to the
    deepest stack, as the reference's table is."""
    st = _stacks(seg, n)
    depth = max((len(s) for s in st), default=0)
    levels = [[] for _ in range(depth)]
    left = len(seg)
    level = 0
    stall = 0
    while left:
        placed_before = left
        used = set()
        if level == len(levels):
            levels.append([])
        for q in range(n):
            if not st[q]:
                continue
            head = st[q][0]
            name = head.name
            if name in ('hold_measure', 'hold_reset'):
                continue
            if name in ('u1', 'u3'):
                if q in used:
                    continue
                levels[level].append(head)
                used.add(q)
                st[q].popleft()
                left -= 1
            elif name in ('cx', 'CX'):
                other = head.qubits[1] if head.qubits[0] == q else head.qubits[0]
                if q in used or other in used:
                    continue
                if st[other][0] is head:
                    used.add(q)
                    used.add(other)
                    levels[level].append(head)
                    st[q].popleft()
                    st[other].popleft()
                    left -= 1
            elif name in ('measure', 'reset'):
                hold = 'hold_' + name
                if all((not s) or s[0].name in (name, hold) for s in st):
                    if levels[level]:
                        used = set()
                        level += 1
                        if level == len(levels):
                            levels.append([])
                    for x in range(n):
                        if not st[x]:
                            continue
                        if st[x][0].name == name:
                            used.add(x)
                            levels[level].append(st[x][0])
                            left -= 1
                        st[x].popleft()
                    break
            else:
                raise BasicAerError('dm_simulator encountered unrecognized operation "%s"' % name)
            if not left:
                break
        level += 1
        stall = stall + 1 if left == placed_before else 0
        if stall > 2:
            raise BasicAerError('partition cannot place the remaining %d instructions' % left)
    return levels, level


def _partition_kind(op, n):
    from . import native
    if op.name == 'bfunc' or not hasattr(op, 'qubits') and op.name != 'barrier':
        raise BasicAerError('dm_simulator cannot partition "%s": conditional operations are not '
                            'supported on this backend' % op.name)
    if op.name == 'barrier':
        return native.INST_BARRIER
    if op.name in ('u1', 'u3'):
        return native.INST_ROT
    if op.name in ('cx', 'CX'):
        return native.INST_CX
    if op.name == 'reset':
        return native.INST_RESET
    if op.name == 'measure':
        prm = getattr(op, 'params', None)
        return native.INST_OWN_LEVEL if prm is not None and str(prm[0]) in OWN_LEVEL else native.INST_MEASURE
    raise BasicAerError('dm_simulator encountered unrecognized operation "%s"' % op.name)


def partition(insts, n):
    """Split at barriers and level each segment (native front end).
    Returns (levels, count); the clock loop runs ``count`` levels over
    ``levels`` (which may carry the reference's trailing empty lists)."""
    from . import native
    kinds = [_partition_kind(op, n) for op in insts]
    q0 = [op.qubits[0] if getattr(op, 'qubits', None) else 0 for op in insts]
    q1 = [op.qubits[1] if k == native.INST_CX else -1 for op, k in zip(insts, kinds)]
    try:
        lv, count = native.partition_levels(n, kinds, q0, q1)
    except native.NativeError as e:
        raise BasicAerError(str(e).split(': ', 1)[-1])
    return [[insts[i] for i in level] for level in lv], count


def partition_py(insts, n):
    """Split at barriers and level each segment.  Returns (levels, count);
    the clock loop runs ``count`` levels over ``levels`` (which may carry the
    reference's trailing empty lists)."""
    segments, cur = [], []
    for op in insts:
        if op.name == 'barrier':
            segments.append(cur)
            cur = []
        else:
            cur.append(op)
    if cur:
        segments.append(cur)
    levels, count = [], 0
    for seg in segments:
        if not seg:
            continue
        head = seg[0]
        prm = getattr(head, 'params', None)
        if head.name == 'measure' and prm is not None and str(prm[0]) in OWN_LEVEL:
            levels.append([seg])
            count += 1
        else:
            lv, c = _level_segment(seg, n)
            levels.append(lv)
            count += c
    return list(itertools.chain(*levels)), count
