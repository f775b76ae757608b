"""FP64 numpy restatement of the reference ``dm_simulator`` backend.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Every function names the reference lines it restates.  The state is the
reference's: a float64 array of 4**n Pauli coefficients, shape ``n*[4]``,
qubit q on axis q (stride 4**(n-1-q)).  The numerical steps (full-state copy,
then slice updates in the reference's order) are kept so that rounding
follows the reference; the ensemble measurement is the one place where the
reference's dense 2^n x 2^n Kronecker operator is replaced by the equivalent
per-qubit contraction (same sum, different summation order), because the
dense operator does not fit in memory beyond ~13 qubits.

Instructions are duck-typed records with the attributes the reference reads
from a ``QasmQobjInstruction``: ``name``, ``qubits``, ``memory`` (optional),
``params`` (optional), ``register`` / ``conditional`` (optional).
"""
import copy
import itertools
import logging
import os
import time

import numpy as np

logger = logging.getLogger(__name__)


class RefError(Exception):
    """Raised where the reference raises BasicAerError / QiskitError."""


# ---------------------------------------------------------------------------
# 1. Gate-level state updates (basicaertools.py)
# ---------------------------------------------------------------------------

_ROT_AXES = {'rz': (1, 2), 'ry': (3, 1), 'rx': (2, 3)}


def gate_rotations(name, params):
    """u3 -> [rz lam, ry theta, rz phi]; u1 -> [rz lam].

    Restates ``single_gate_dm_matrix`` (basicaertools.py:69-90).
    """
    p = [float(x) for x in params]
    if name in ('U', 'u3'):
        return [('rz', p[2]), ('ry', p[0]), ('rz', p[1])]
    if name == 'u1':
        return [('rz', p[0])]
    raise RefError('Gate is not among the valid types: %s' % name)


def rotate_components(state3, axis, angle, err):
    """One noisy rotation of two Pauli components of the middle axis.

    ``state3`` has shape (4**q, 4, 4**(n-q-1)).  With c = e0*cos(angle+e1)
    and s = e0*sin(angle+e1) the pair (k0, k1) becomes
    (c*k0 - s*k1, c*k1 + s*k0); a full copy is taken first as in
    ``rot_gate_dm_matrix`` (basicaertools.py:93-126).
    """
    c = err[0] * np.cos(angle + err[1])
    s = err[0] * np.sin(angle + err[1])
    if axis not in _ROT_AXES:
        raise RefError('Gate is not among the valid decomposition types: %s' % axis)
    k0, k1 = _ROT_AXES[axis]
    snapshot = state3.copy()
    a = snapshot[:, k0, :]
    b = snapshot[:, k1, :]
    state3[:, k0, :] = c * a - s * b
    state3[:, k1, :] = c * b + s * a
    return state3


def cx_coefficients(err):
    """(c, s, c2, s2, cs) of the TSP-error CX (basicaertools.py:328-336)."""
    cav = err[0]
    c2av = 4 * cav - 3
    c = cav * np.cos(err[1])
    s = cav * np.sin(err[1])
    c2 = 0.5 * (1 + c2av * np.cos(2 * err[1]))
    s2 = 0.5 * (1 - c2av * np.cos(2 * err[1]))
    cs = c2av * np.sin(err[1]) * np.cos(err[1])
    return c, s, c2, s2, cs


def apply_cx(state, ctrl, tgt, err, n):
    """CX in the Pauli basis, with the transition-selective-pulse error model.

    Restates ``cx_gate_dm_matrix`` (basicaertools.py:310-392).  The two
    branches there (control slower / faster axis, :340-363 and :365-388) are
    the same map written in different axis orders; here the map is written
    once in (control component, target component) coordinates ``T[a, b]``.
    """
    if ctrl == tgt or ctrl >= n or tgt >= n:
        raise RefError('Qubit Labels out of bound in CX Gate')
    c, s, c2, s2, cs = cx_coefficients(err)
    lo, hi = min(ctrl, tgt), max(ctrl, tgt)
    st = np.reshape(state, (4 ** lo, 4, 4 ** (hi - lo - 1), 4, 4 ** (n - hi - 1)))
    old = st.copy()
    if ctrl < tgt:
        def view(arr, a, b):
            return arr[:, a, :, b, :]
    else:
        def view(arr, a, b):
            return arr[:, b, :, a, :]

    def T(a, b):
        return view(old, a, b)

    def put(a, b, val):
        view(st, a, b)[...] = val

    # Z-type block on the target (components 2,3) conditioned on control I/Z.
    put(0, 2, s2 * T(0, 2) + c2 * T(3, 2) - cs * (T(0, 3) - T(3, 3)))
    put(3, 2, c2 * T(0, 2) + s2 * T(3, 2) + cs * (T(0, 3) - T(3, 3)))
    put(0, 3, s2 * T(0, 3) + c2 * T(3, 3) + cs * (T(0, 2) - T(3, 2)))
    put(3, 3, c2 * T(0, 3) + s2 * T(3, 3) - cs * (T(0, 2) - T(3, 2)))
    # Control X / Y rows.
    put(1, 0, c * T(1, 1) - s * T(2, 0))
    put(1, 1, c * T(1, 0) - s * T(2, 1))
    put(1, 2, -s * T(2, 2) + c * T(2, 3))
    put(1, 3, -c * T(2, 2) - s * T(2, 3))
    put(2, 0, s * T(1, 0) + c * T(2, 1))
    put(2, 1, s * T(1, 1) + c * T(2, 0))
    put(2, 2, s * T(1, 2) - c * T(1, 3))
    put(2, 3, c * T(1, 2) + s * T(1, 3))
    return np.reshape(st, n * [4])


def apply_memory_noise(state, n, f, p, g):
    """Decoherence + amplitude decay towards the thermal state, all qubits.

    Per qubit, in qubit order: x,y <- sqrt(g)*f*(x,y); z <- g*z + (1-g)(2p-1)*a_I,
    one full copy per qubit.  Restates ``_add_decoherence_and_amp_decay``
    (dm_simulator.py:397-425).
    """
    od = np.sqrt(g) * f
    dd = (1 - g) * (2 * p - 1)
    for q in range(n):
        st = np.reshape(state, (4 ** q, 4, 4 ** (n - q - 1)))
        old = st.copy()
        st[:, 1, :] = od * old[:, 1, :]
        st[:, 2, :] = od * old[:, 2, :]
        st[:, 3, :] = g * old[:, 3, :] + dd * old[:, 0, :]
        state = st
    return np.reshape(state, n * [4])


def apply_gate_depolarizing(state, n, qubits, prob):
    """Per-gate depolarising channel (no reference counterpart).

    rho -> (1-p) rho + p Tr_q(rho) (x) I/2 on each listed qubit, i.e. the
    X, Y, Z components of that qubit are scaled by (1-p).  Pinned by the dense
    Kraus oracle.
    """
    lam = 1.0 - prob
    for q in qubits:
        st = np.reshape(state, (4 ** q, 4, 4 ** (n - q - 1)))
        st[:, 1:, :] *= lam
        state = st
    return np.reshape(state, n * [4])


# ---------------------------------------------------------------------------
# 2. Host scheduling: merge and partition (basicaertools.py)
# ---------------------------------------------------------------------------

def u3_merge_angles(xi, theta1, theta2):
    """Y-Z-Y to Z-Y-Z conversion used when two U3 gates merge.

    Returns (beta, alpha, gamma).  Restates ``U3_merge``
    (basicaertools.py:129-178), same numpy calls in the same order.
    """
    sxi = np.sin(xi * 0.5)
    cxi = np.cos(xi * 0.5)
    sp = np.sin((theta1 + theta2) * 0.5)
    cp = np.cos((theta1 + theta2) * 0.5)
    sm = np.sin((theta1 - theta2) * 0.5)
    cm = np.cos((theta1 - theta2) * 0.5)
    half_sum = np.arctan2(sxi * cm, cxi * cp)
    half_diff = np.arctan2(-sxi * sm, cxi * sp)
    alpha = half_sum + half_diff
    gamma = half_sum - half_diff
    cb2 = np.sqrt((cxi * cp) ** 2 + (sxi * cm) ** 2)
    beta = 2 * np.arccos(cb2)
    return beta, alpha, gamma


def merge_pair(g1, g2):
    """Merge two consecutive single-qubit gates ``[inst, index]``.

    Restates ``mergeU`` (basicaertools.py:181-228): the result keeps the
    record of the earlier gate (smaller index), u1+u1 adds angles, u1 with u3
    folds into lambda / phi, u3+u3 goes through ``u3_merge_angles``.
    """
    base = copy.deepcopy(g1 if g1[1] < g2[1] else g2)
    a, b = g1[0], g2[0]
    rec = base[0]
    if a.name == 'u1' and b.name == 'u1':
        rec.params[0] = a.params[0] + b.params[0]
    elif a.name == 'u1' or b.name == 'u1':
        if rec.name == 'u1':
            rec.name = 'u3'
            rec.params.extend([0, 0])
        if a.name == 'u1' and b.name == 'u3':
            rec.params[0] = b.params[0]
            rec.params[1] = b.params[1]
            rec.params[2] = b.params[2] + a.params[0]
        elif a.name == 'u3' and b.name == 'u1':
            rec.params[0] = a.params[0]
            rec.params[1] = a.params[1] + b.params[0]
            rec.params[2] = a.params[2]
    elif a.name == 'u3' and b.name == 'u3':
        xi = float(b.params[2] + a.params[1])
        th1 = float(b.params[0])
        th2 = float(a.params[0])
        beta, alpha, gamma = u3_merge_angles(xi, th1, th2)
        rec.params[0] = beta
        rec.params[1] = b.params[1] + alpha
        rec.params[2] = a.params[2] + gamma
    else:
        raise RefError('Encountered unrecognized instructions: %s, %s' % (a.name, b.name))
    return base


def merge_run(run):
    """Fold a run of ``[inst, index]`` pairs left to right (merge_gates :231-248)."""
    if len(run) < 2:
        return run[0][0]
    acc = merge_pair(run[0], run[1])
    for item in run[2:]:
        acc = merge_pair(acc, item)
    return acc[0]


_SEPARATORS = ('CX', 'cx', 'measure', 'bfunc', 'reset', 'barrier')


def single_gate_merge(insts, n, merge_flag=True):
    """Merge runs of u1/u2/u3 per qubit between separators.

    Restates ``single_gate_merge`` (basicaertools.py:251-307): any separator
    flushes every qubit's pending run (in qubit order) before it is emitted;
    u2(phi, lam) becomes u3(pi/2, phi, lam); id / u0 are dropped.
    Mutates the records (as the reference does) -- pass copies.
    """
    pending = [[] for _ in range(n)]
    out = []
    if merge_flag:
        for idx, op in enumerate(insts):
            if op.name in _SEPARATORS:
                for q in range(n):
                    if pending[q]:
                        out.append(merge_run(pending[q]))
                        pending[q] = []
                if op.name == 'CX':
                    op.name = 'cx'
                out.append(op)
            elif op.name in ('U', 'u1', 'u2', 'u3'):
                if op.name == 'U':
                    op.name = 'u3'
                elif op.name == 'u2':
                    op.name = 'u3'
                    op.params.insert(0, np.pi / 2)
                pending[op.qubits[0]].append([op, idx])
            elif op.name in ('id', 'u0'):
                continue
            else:
                raise RefError('Encountered unrecognized instruction: %s' % op.name)
        for q in range(n):
            if pending[q]:
                out.append(merge_run(pending[q]))
    else:
        for op in insts:
            if op.name == 'CX':
                op.name = 'cx'
            elif op.name == 'U':
                op.name = 'u3'
            elif op.name == 'u2':
                op.name = 'u3'
                op.params.insert(0, np.pi / 2)
            if op.name not in ('id', 'u0'):
                out.append(op)
    return out


def _dummy_of(inst, kind):
    d = copy.deepcopy(inst)
    d.name = kind
    d.qubits[0] = -1
    return d


def qubit_stack(insts, n):
    """Per-qubit instruction stacks with measure/reset placeholders.

    Restates ``qubit_stack`` (basicaertools.py:536-591): a measure (reset)
    on qubit q joins the previous measure group if q's stack ends with a
    placeholder of that group, otherwise it opens a new group and every other
    qubit receives a placeholder.
    """
    stacks = [[] for _ in range(n)]
    for inst in insts:
        if inst.name not in ('measure', 'reset'):
            for q in inst.qubits:
                stacks[q].append(inst)
            continue
        kind = 'dummy_measure' if inst.name == 'measure' else 'dummy_reset'
        q = inst.qubits[0]
        if stacks[q] and stacks[q][-1].name == kind:
            stacks[q][-1] = inst
        else:
            stacks[q].append(inst)
            d = _dummy_of(inst, kind)
            for other in set(range(n)).difference(set(inst.qubits)):
                stacks[other].append(d)
    depth = max(len(s) for s in stacks)
    return stacks, depth


def _remove_identity(lst, item):
    for i, x in enumerate(lst):
        if x is item:
            del lst[i]
            return
    raise ValueError('instruction not in list')


def partition_levels(insts, n):
    """Greedy level assignment of one barrier-free segment.

    Restates ``partition_helper`` (basicaertools.py:594-712).  Returns
    (levels_list, level_count).
    """
    remaining = list(insts)
    stacks, depth = qubit_stack(insts, n)
    level = 0
    seq = [[] for _ in range(depth)]
    guard = 0
    while remaining:
        guard += 1
        if guard > 4 * (len(insts) + 2) * (n + 2):
            raise RefError('partition made no progress')
        used = []
        if level == len(seq):
            seq.append([])
        for q in range(n):
            if not stacks[q]:
                continue
            gate = stacks[q][0]
            if gate.name in ('dummy_measure', 'dummy_reset'):
                continue
            if gate.name in ('u3', 'u1'):
                if q in used:
                    continue
                seq[level].append(gate)
                used.append(q)
                _remove_identity(remaining, gate)
                stacks[q].pop(0)
            elif gate.name in ('CX', 'cx'):
                other = list(set(gate.qubits).difference({q}))[0]
                partner_top = stacks[other][0]
                if q in used or other in used:
                    continue
                if gate is partner_top:
                    used.append(q)
                    used.append(other)
                    seq[level].append(gate)
                    _remove_identity(remaining, gate)
                    stacks[q].pop(0)
                    stacks[other].pop(0)
                else:
                    continue
            elif gate.name in ('measure', 'reset'):
                kind = gate.name
                dummy = 'dummy_' + kind
                ready = True
                for x in range(n):
                    if not stacks[x]:
                        continue
                    if stacks[x][0].name not in (kind, dummy):
                        ready = False
                        break
                if ready:
                    if seq[level]:
                        used = []
                        level += 1
                        if level == len(seq):
                            seq.append([])
                    for x in range(n):
                        if not stacks[x]:
                            continue
                        if stacks[x][0].name == kind:
                            used.append(x)
                            seq[level].append(stacks[x][0])
                            _remove_identity(remaining, stacks[x][0])
                        stacks[x].pop(0)
                    break
            if not remaining:
                break
        level += 1
    return seq, level


_OWN_LEVEL_BASES = ('Bell', 'Expect', 'Ensemble')


def partition(insts, n):
    """Split at barriers, then level each segment.

    Restates ``partition`` (basicaertools.py:714-748).  A segment that starts
    with a Bell / Expect / Ensemble measure is one level on its own.  The
    reference flattens such a segment into bare instructions with
    ``itertools.chain`` (:747) while counting it as one level; here it is kept
    as one level holding those measures, the behaviour the user guide's
    recorded outputs show.  Returns (levels, count); ``levels`` may carry
    trailing empty lists from the pre-sized per-segment tables, exactly as in
    the reference.
    """
    segments, cur = [], []
    for inst in insts:
        if inst.name != 'barrier':
            cur.append(inst)
        else:
            segments.append(cur)
            cur = []
    if cur:
        segments.append(cur)
    out, count = [], 0
    for seg in segments:
        if not seg:
            continue
        head = seg[0]
        hp = getattr(head, 'params', None)
        if head.name == 'measure' and hp is not None and str(hp[0]) in _OWN_LEVEL_BASES:
            out.append([seg])
            count += 1
        else:
            seq, lv = partition_levels(seg, n)
            out.append(seq)
            count += lv
    return list(itertools.chain(*out)), count


# ---------------------------------------------------------------------------
# 3. Pauli -> matrix conversion (dm_simulator.py:1198-1269)
# ---------------------------------------------------------------------------

def pauli_to_matrix(coeffs, n):
    """2^n x 2^n complex density matrix from the Pauli coefficients.

    Per qubit (in qubit order) the component index i in {I,X,Y,Z} becomes the
    matrix-unit index mu in {00,01,10,11}: mu0 = I+Z, mu1 = X - iY,
    mu2 = X + iY, mu3 = I - Z, accumulated in that order; then mu_q = (r_q, c_q)
    is scattered with qubit 0 as the most significant bit of row and column.
    Restates ``_compute_densitymatrix`` (:1198-1255) with the final O(4^n)
    Python scatter replaced by an equivalent reshape/transpose.
    """
    cur = np.asarray(coeffs, dtype=complex).reshape(4 ** n)
    mi = complex(0, -1)
    pi_ = complex(0, 1)
    for q in range(n):
        src = cur.reshape(4 ** q, 4, 4 ** (n - q - 1))
        dst = np.zeros_like(src)
        dst[:, 0, :] += 1 * src[:, 0, :]
        dst[:, 0, :] += 1 * src[:, 3, :]
        dst[:, 1, :] += 1 * src[:, 1, :]
        dst[:, 1, :] += mi * src[:, 2, :]
        dst[:, 2, :] += 1 * src[:, 1, :]
        dst[:, 2, :] += pi_ * src[:, 2, :]
        dst[:, 3, :] += 1 * src[:, 0, :]
        dst[:, 3, :] += -1 * src[:, 3, :]
        cur = dst
    # index of mu_q = 2*r_q + c_q; axes (r0,c0,r1,c1,...) -> (r0..r_{n-1}, c0..c_{n-1})
    t = cur.reshape([2, 2] * n)
    order = [2 * q for q in range(n)] + [2 * q + 1 for q in range(n)]
    return np.ascontiguousarray(t.transpose(order)).reshape(2 ** n, 2 ** n)


# ---------------------------------------------------------------------------
# 4. Measurement reductions (dm_simulator.py:427-777)
# ---------------------------------------------------------------------------

_SUPPORT = {'X': [0, 1], 'Y': [0, 2], 'Z': [0, 3], 'N': [0, 1, 2, 3]}


def ensemble_probabilities(state, n, basis, add_param, err):
    """Probabilities of all 2^n outcomes of an all-qubit measurement.

    The reference (:427-481) gathers the coefficients with components in
    {0, k}^n ({0..3}^n for N), builds the Kronecker operator
    [[1, e],[1, -e]]^(x)n (rows [1, n*e], [1, -n*e] for N) and takes 2^n dot
    products.  The same sum is evaluated here one qubit at a time.  Keys are
    '0'/'1' strings with qubit 0 leftmost.
    """
    st = np.reshape(state, n * [4])
    sup = _SUPPORT[basis]
    v = st[np.ix_(*([sup] * n))] if n > 0 else st
    if basis != 'N':
        row = np.array([[1.0, err], [1.0, -err]])
    else:
        nv = np.asarray(add_param, dtype=float) * err
        row = np.array([[1.0, nv[0], nv[1], nv[2]], [1.0, -nv[0], -nv[1], -nv[2]]])
    cur = np.asarray(v, dtype=float)
    for q in range(n):
        cur = np.moveaxis(np.tensordot(row, cur, axes=([1], [q])), 0, q)
    probs = np.reshape(cur, 2 ** n)
    keys = [''.join(s) for s in itertools.product('01', repeat=n)]
    return dict(zip(keys, probs))


def ensemble_probabilities_kron(state, n, basis, add_param, err):
    """Literal form of the reference's dense Kronecker operator (small n only)."""
    sup = _SUPPORT[basis]
    st = np.reshape(state, n * [4])
    ind = list(itertools.product(sup, repeat=n))
    vals = [st[x] for x in ind]
    if basis != 'N':
        one = np.array([[1, err], [1, -err]], dtype=float)
    else:
        nv = np.asarray(add_param, dtype=float) * err
        one = np.array([[1, nv[0], nv[1], nv[2]], [1, -nv[0], -nv[1], -nv[2]]])
    op = one
    for _ in range(n - 1):
        op = np.kron(one, op)
    probs = np.reshape(np.array([np.sum(np.multiply(vals, r)) for r in op]), 2 ** n)
    keys = [''.join(s) for s in itertools.product('01', repeat=n)]
    return dict(zip(keys, probs))


def collapse_axis(state, n, q, comp, err):
    """Single-qubit X / Y / Z measurement: keep component ``comp`` times err,
    zero the other two; P(0) = (1 + 2^n a_{e_q})/2.  Restates
    ``_add_qasm_measure_{X,Y,Z}`` (:574-664)."""
    st = np.reshape(state, (4 ** q, 4, 4 ** (n - q - 1)))
    for k in (1, 2, 3):
        if k == comp:
            st[:, k, :] *= err
        else:
            st[:, k, :] = 0
    st = np.reshape(st, n * [4])
    idx = [0] * n
    idx[q] = comp
    p0 = 0.5 * (1 + st[tuple(idx)] * 2 ** n)
    return st, p0


def collapse_direction(state, n, q, nvec, err):
    """Measurement along unit vector n: (x,y,z) <- err*(n.v) n.
    Restates ``_add_qasm_measure_N`` (:666-704)."""
    st = np.reshape(state, (4 ** q, 4, 4 ** (n - q - 1)))
    proj = nvec[0] * st[:, 1, :] + nvec[1] * st[:, 2, :] + nvec[2] * st[:, 3, :]
    proj *= err
    st[:, 1, :] = proj * nvec[0]
    st[:, 2, :] = proj * nvec[1]
    st[:, 3, :] = proj * nvec[2]
    st = np.reshape(st, n * [4])
    idx = [0] * n
    pn = 0.0
    for k in range(3):
        idx[q] = k + 1
        pn += st[tuple(idx)]
    pn *= 2 ** n
    return st, 0.5 * (1 + pn)


def reset_qubit(state, n, q):
    """x, y <- 0 and z <- a_I on qubit q (``_add_qasm_reset`` :810-823)."""
    st = np.reshape(state, (4 ** q, 4, 4 ** (n - q - 1)))
    st[:, 1, :] = 0
    st[:, 2, :] = 0
    st[:, 3, :] = st[:, 0, :].copy()
    return np.reshape(st, n * [4])


def pauli_expectation(state, n, pstring, err):
    """Collapse every qubit onto its Pauli letter and return 2^n a_index.
    Restates ``_pauli_string_expectation`` (:540-572)."""
    letter = {'I': 0, 'X': 1, 'Y': 2, 'Z': 3}
    for i in range(n):
        st = np.reshape(state, (4 ** i, 4, 4 ** (n - i - 1)))
        k = letter[pstring[i]]
        if k != 0:
            for j in (1, 2, 3):
                if j == k:
                    st[:, j, :] *= err
                else:
                    st[:, j, :] = 0
        state = st
    state = np.reshape(state, n * [4])
    return state, state[tuple(letter[x] for x in pstring)] * 2 ** n


def bell_measure(state, n, qa, qb, err):
    """Bell-basis measurement of the pair given as (qa, qb).

    The reference (:721-777) reshapes to (4^(n-q2-1), 4, 4^(q2-q1-1), 4, 4^q1)
    with q1 < q2, i.e. it addresses the axes of qubits n-1-q2 and n-1-q1 (the
    mirrored stride convention).  Restated literally.
    """
    q1, q2 = min(qa, qb), max(qa, qb)
    st = np.reshape(state, (4 ** (n - q2 - 1), 4, 4 ** (q2 - q1 - 1), 4, 4 ** q1))
    reduced = np.zeros((4, 4))
    for i in range(4):
        for j in range(4):
            reduced[i, j] = st[0, i, 0, j, 0] * 2 ** (n - 2)
    for i in range(4):
        for j in range(4):
            if i != j:
                st[:, i, :, j, :] = 0
    for i in (1, 2, 3):
        st[:, i, :, i, :] *= err
    k = [st[0, i, 0, i, 0] * 2 ** n for i in range(4)]
    probs = [0.25 * (k[0] + k[1] - k[2] + k[3]),
             0.25 * (k[0] - k[1] + k[2] + k[3]),
             0.25 * (k[0] + k[1] + k[2] - k[3]),
             0.25 * (k[0] - k[1] - k[2] - k[3])]
    return np.reshape(st, n * [4]), dict(zip(['Bell_1', 'Bell_2', 'Bell_3', 'Bell_4'], probs)), reduced


# ---------------------------------------------------------------------------
# 5. Initial states (dm_simulator.py:284-365)
# ---------------------------------------------------------------------------

def initial_state(n, custom=None, initial=None, thermal_factor=1.0):
    """Product initial states built with np.kron, normalised by 2^-n."""
    if initial is None:
        if custom is None:
            one = [1, 0, 0, 1]
        elif custom == 'max_mixed':
            one = [1, 0, 0, 0]
        elif custom == 'uniform_superpos':
            one = [1, 1, 0, 0]
        elif custom == 'thermal_state':
            one = [1, 0, 0, 2 * thermal_factor - 1]
        else:
            raise RefError('_custom_densitymatrix value is invalid')
        st = np.array(one, dtype=float)
        for _ in range(n - 1):
            st = np.kron(one, st)
        st = st * 0.5 ** n
    elif custom == 'binary_string':
        if len(initial) != n:
            raise RefError('Wrong input binary string length')
        st = np.array([1, 0, 0, 1] if initial[0] == '0' else [1, 0, 0, -1], dtype=float)
        for ch in initial[1:]:
            st = np.kron([1, 0, 0, 1] if ch == '0' else [1, 0, 0, -1], st)
        st = st * 0.5 ** n
    elif custom == 'stored_density_matrix':
        st = np.load('stored_density_matrix.npy')
        if len(st) != 4 ** n:
            raise RefError('Wrong input stored density matrix')
    else:
        # An ``initial_densitymatrix`` without a matching ``custom_densitymatrix``
        # is rejected by the reference (:344-345).
        raise RefError('_custom_densitymatrix value is invalid')
    st = np.array(st, dtype=float).reshape(-1)
    if initial is not None:
        # _validate_initial_densitymatrix (:351-365)
        if st.size != 4 ** n:
            raise RefError('initial densitymatrix is incorrect length')
        if st[0] != 2.0 ** (-n):
            raise RefError('Trace of initial densitymatrix is not one')
    return np.reshape(st, n * [4])


# ---------------------------------------------------------------------------
# 6. The driver (dm_simulator.py:950-1196)
# ---------------------------------------------------------------------------

DEFAULTS = {
    'chop_threshold': 1e-15,
    'thermal_factor': 1.0,
    'decoherence_factor': 1.0,
    'depolarization_factor': 1.0,
    'bell_depolarization_factor': 1.0,
    'decay_factor': 1.0,
    'rotation_error': {'rx': [1.0, 0.0], 'ry': [1.0, 0.0], 'rz': [1.0, 0.0]},
    'tsp_model_error': [1.0, 0.0],
    'gate_depolarization': 0.0,
}


def _unit(v):
    """``_unit_vector_normalisation`` (:706-719)."""
    v = np.asarray(v, dtype=float)
    nrm = np.linalg.norm(v)
    return v if nrm == 1 else v / nrm


def run_experiment(instructions, n, options=None, strict_noise=True):
    """Simulate one experiment's (assembled) instruction list.

    ``options`` uses the reference's ``backend_options`` keys (:177-271), plus
    ``gate_depolarization`` (probability p of the per-gate depolarising
    channel, applied after every executed u1/u3/cx on its qubits; 0 = off).
    Memory noise runs after every level (:1172-1177) as in the reference;
    with ``strict_noise=False`` it is skipped when it is the identity
    (f = g = 1), which differs from the reference only in the sign of zeros.

    Returns ``(data, levels, timings)`` where ``data`` is the reference's
    ``data`` dict (coeffmatrix, densitymatrix, probabilities ...).
    """
    opt = dict(DEFAULTS)
    opt.update(options or {})
    insts = copy.deepcopy(list(instructions))
    custom = opt.get('custom_densitymatrix')
    initial = opt.get('initial_densitymatrix')
    rot_err = dict(DEFAULTS['rotation_error'])
    rot_err.update(opt.get('rotation_error') or {})
    tsp = list(opt['tsp_model_error'])
    f, p, g = opt['decoherence_factor'], opt['thermal_factor'], opt['decay_factor']
    meas_err = opt['depolarization_factor']
    bell_err = 1.0  # bell_depolarization_factor is never applied (:240)
    gdep = float(opt.get('gate_depolarization', 0.0))
    want_dense = opt.get('compute_densitymatrix', True)

    t0 = time.time()
    state = initial_state(n, custom, initial, opt['thermal_factor'])
    insts = single_gate_merge(insts, n, opt.get('merge', True))
    levels, count = partition(insts, n)
    t1 = time.time()
    data = {}
    creg = 0
    cmem = 0
    part_measure = False
    for clock in range(count):
        level = levels[clock]
        for op in level:
            if op.name == 'measure':
                basis = str(getattr(level[0], 'params', ['Z'])[0])
                prm = getattr(op, 'params', ['Z'])
                if str(prm[0]) in ('X', 'Y', 'Z', 'N'):
                    if str(prm[0]) == basis:
                        part_measure = True
                    else:
                        part_measure = False
                        break
                else:
                    part_measure = False
                    continue
        for op in level:
            cond = getattr(op, 'conditional', None)
            if isinstance(cond, int):
                if not (creg >> cond) & 1:
                    continue
            elif cond is not None:
                mask = int(cond.mask, 16)
                if mask > 0:
                    value = cmem & mask
                    while (mask & 0x1) == 0:
                        mask >>= 1
                        value >>= 1
                    if value != int(cond.val, 16):
                        continue
            if op.name in ('u1', 'u3'):
                q = op.qubits[0]
                st3 = np.reshape(state, (4 ** q, 4, 4 ** (n - q - 1)))
                for axis, ang in gate_rotations(op.name, getattr(op, 'params', None)):
                    st3 = rotate_components(st3, axis, ang, rot_err[axis])
                state = np.reshape(st3, n * [4])
                if gdep:
                    state = apply_gate_depolarizing(state, n, [q], gdep)
            elif op.name == 'cx':
                state = apply_cx(state, op.qubits[0], op.qubits[1], tsp, n)
                if gdep:
                    state = apply_gate_depolarizing(state, n, op.qubits[:2], gdep)
            elif op.name == 'reset':
                state = reset_qubit(state, n, op.qubits[0])
            elif op.name == 'barrier':
                pass
            elif op.name == 'measure':
                prm = getattr(op, 'params', None)
                if prm is None:
                    prm = ['Z']
                    op.params = prm
                q = op.qubits[0]
                kind = str(prm[0])
                single = expect = ens = False
                if kind == 'Ensemble':
                    ens = True
                elif kind == 'Expect':
                    expect = True
                elif len(level) == 1:
                    single = True
                elif part_measure is not True:
                    single = True
                if len(prm) == 1:
                    prm.append(None)
                if single:
                    if kind == 'X':
                        state, _ = collapse_axis(state, n, q, 1, meas_err)
                    elif kind == 'Y':
                        state, _ = collapse_axis(state, n, q, 2, meas_err)
                    elif kind == 'N':
                        prm[1] = _unit(prm[1])
                        state, _ = collapse_direction(state, n, q, prm[1], meas_err)
                    elif kind == 'Bell':
                        pair = str(prm[1])
                        state, bp, red = bell_measure(state, n, int(pair[0]), int(pair[1]), bell_err)
                        data['bell_probabilities%s%s' % (pair[0], pair[1])] = bp
                        data['reduced_bell_densitymatrix%s%s' % (pair[0], pair[1])] = red
                    else:
                        state, _ = collapse_axis(state, n, q, 3, meas_err)
                    _remove_identity(level, op)
                elif part_measure:
                    mqs = [x.qubits[0] for x in level]
                    if kind == 'N':
                        addp = prm[1]
                    else:
                        addp = None
                    probs = ensemble_probabilities(state, n, kind, addp, meas_err)
                    arr = np.array(list(probs.values()))
                    rest = tuple(sorted(set(range(n)) - set(mqs)))
                    marg = np.reshape(np.sum(np.reshape(arr, n * [2]), axis=rest), 2 ** len(mqs))
                    keys = [''.join(s) for s in itertools.product('01', repeat=len(mqs))]
                    data['partial_probability'] = dict(zip(keys, marg))
                    for mq in mqs:
                        if kind == 'N' and addp is not None:
                            state, _ = collapse_direction(state, n, mq, addp, meas_err)
                        else:
                            state, _ = collapse_axis(state, n, mq, {'X': 1, 'Y': 2, 'Z': 3}[kind], meas_err)
                    break
                elif expect:
                    state, val = pauli_expectation(state, n, str(prm[1]), meas_err)
                    data['Pauli_string_expectation'] = val
                    break
                elif ens:
                    spec = prm[1]
                    if isinstance(spec, (tuple, list)) and len(spec) == 2 and str(spec[0]) == 'N':
                        addp, basis = _unit(spec[1]), 'N'
                    else:
                        addp, basis = None, str(spec)
                    data['ensemble_probability'] = ensemble_probabilities(state, n, basis, addp, meas_err)
                    break
            elif op.name == 'bfunc':
                mask = int(op.mask, 16)
                val = int(op.val, 16)
                cmp_ = (creg & mask) - val
                rel = op.relation
                outcome = {'==': cmp_ == 0, '!=': cmp_ != 0, '<': cmp_ < 0,
                           '<=': cmp_ <= 0, '>': cmp_ > 0, '>=': cmp_ >= 0}.get(rel)
                if outcome is None:
                    raise RefError('Invalid boolean function relation.')
                creg = (creg & ~(1 << op.register)) | (int(outcome) << op.register)
                mb = getattr(op, 'memory', None)
                if mb is not None:
                    cmem = (cmem & ~(1 << mb)) | (int(outcome) << mb)
            else:
                raise RefError('dm_simulator encountered unrecognized operation "%s"' % op.name)
        if strict_noise or not (f == 1 and g == 1):
            state = apply_memory_noise(state, n, f, p, g)
    t2 = time.time()
    vec = np.reshape(state, 4 ** n)
    if want_dense:
        dense = pauli_to_matrix(vec, n)
    vec[abs(vec) < opt['chop_threshold']] = 0.0
    data['coeffmatrix'] = vec
    if want_dense:
        data['densitymatrix'] = dense
    t3 = time.time()
    return data, count, {'processing': t1 - t0, 'levels': t2 - t1, 'output': t3 - t2}
