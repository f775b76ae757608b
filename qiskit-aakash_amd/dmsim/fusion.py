"""Gate fusion: groups of noisy gates become one k-qubit Pauli-transfer pass.

A noisy gate is a PTM on one or two qubits (the gate followed by its
channels).  A *block* is a set of gates on at most k qubits that can be
applied together: its gates keep their order on every qubit, and every gate
left out of it that comes earlier in the program acts on other qubits only
(see ``_grow``), so running the block first commutes past them.  A block of one gate runs on the plain kernels; a
larger block is one tile pass (DM_OP_FUSED): the gates applied in sequence to
LDS-staged tiles, or the block's dense PTM on the matrix cores.  ``plan`` chooses the blocks:

* ``greedy``  -- grow each block up to k_max qubits (the classic rule);
* ``learned`` -- at every step score the candidate blocks seeded at the first
  pending gate (the greedy growth for every k, and every qubit set of at most
  k_max qubits drawn from the nearby two-qubit gates, ``candidate_blocks``)
  by their predicted net saving -- their gates valued at the plan's running
  cost per gate, minus TIME_WEIGHT times their predicted time -- the
  predictions coming from the MLP cost model (dmsim.costmodel), and take, of
  the best few, the one whose saving plus the best saving of the blocks
  that can follow it (LOOKAHEAD) is largest; a single gate on its plain
  kernel is always a candidate;
* ``search``  -- the same candidates, scored by gate count alone (every pass
  costed alike): the learned planner without its network.

A sequential tile pass holds the block's qubits and the lowest other
positions (its contiguous HBM runs).  With tile passes of the default size
the learned planner also weighs, for each candidate, the block that takes
every gate its tile holds (k_max then counts only the qubits outside the
three lowest positions): more gates per pass, at the price of more LDS work
per tile, which the network prices.
"""
import collections
import itertools

import numpy as np

from . import native, ptm
from .program import Program

NoisyGate = collections.namedtuple('NoisyGate', 'qubits R kind')

# largest block each planner forms (qubits): 4, a Pauli transfer matrix of at
# most 256 x 256 (the learned planner's tile-wide blocks count only the
# qubits outside the LOW_DIGITS lowest positions)
KMAX = {'greedy': 4, 'learned': 4, 'search': 4}

# every sequential tile pass holds the LOW_DIGITS lowest positions (its runs
# of 4^3 contiguous coefficients), so gates on their qubits join a block
# without widening its tile
LOW_DIGITS = 3

# weight of a block's predicted time against the gates it takes, both in
# seconds (gates at the plan's running cost per gate).  Measured on MI355X
# (16 qubits, FP32, three held-out circuits, tile-wide candidates): weights
# 0.25 and 0.5 plan alike, 1.0 is 4 % slower.
TIME_WEIGHT = 0.25
# lookahead widths of the learned planner: its 3 best candidates are each
# credited with the best saving of the two blocks that can follow (the 2
# best next, then the best after each).  Five held-out 400-gate circuits at
# 16 qubits: 173 passes and 5.1 % less predicted time than choosing without
# lookahead (187), against 178 and 3.4 % for one block of lookahead (3,).
LOOKAHEAD = (3, 2)
# kind: 'u3' (4x4 R), 'cx' (R = per-pair scale16 or None, signed permutation),
#       '2q' (16x16 R)


def noisy_gates(gates, gate_depolarization=0.0, rotation_error=None, tsp=(1.0, 0.0)):
    """NoisyGate list from (name, qubits, params) gates: u1/u3 (and u2) as
    4x4 PTMs followed by the depolarising channel; cx as the signed
    permutation with the channel as a per-pair scale (or the dense TSP PTM)."""
    out = []
    p = float(gate_depolarization)
    for name, qs, prm in gates:
        if name in ('u3', 'u1', 'u2'):
            if name == 'u2':
                name, prm = 'u3', (np.pi / 2,) + tuple(prm)
            R = ptm.gate_ptm(name, prm, rotation_error)
            if p:
                R = ptm.depolarizing1(p) @ R
            out.append(NoisyGate((qs[0],), R, 'u3'))
        elif name == 'cx':
            if list(tsp) == [1.0, 0.0]:
                out.append(NoisyGate((qs[0], qs[1]), ptm.depolarizing_pair_scale(p) if p else None, 'cx'))
            else:
                R = ptm.cx_ptm(tsp)
                if p:
                    R = np.diag(ptm.depolarizing_pair_scale(p)) @ R
                out.append(NoisyGate((qs[0], qs[1]), R, '2q'))
        else:
            raise ValueError('fusion handles u1/u2/u3/cx gates, got %s' % name)
    return out


def gate_ptm16(g):
    """16x16 PTM (pair index 4*i_q0 + i_q1) of a two-qubit NoisyGate."""
    if g.kind == '2q':
        return g.R
    R = ptm.cx_ptm((1.0, 0.0))
    if g.R is not None:
        R = np.diag(g.R) @ R
    return R


def block_ptm(gates, qubits):
    """4^k x 4^k PTM of the gates (in order) on the block qubits; component
    index sum_j i_{qubits[j]} 4^j.  Each gate is applied to the rows of the
    accumulated matrix along its own tensor axes."""
    k = len(qubits)
    dim = 4 ** k
    digit = {q: j for j, q in enumerate(qubits)}
    M = np.eye(dim).reshape([4] * k + [dim])   # row axes: axis k-1-j is digit j
    for g in gates:
        if len(g.qubits) == 1:
            ax = k - 1 - digit[g.qubits[0]]
            M = np.moveaxis(np.tensordot(g.R, M, axes=([1], [ax])), 0, ax)
        else:
            a0 = k - 1 - digit[g.qubits[0]]
            a1 = k - 1 - digit[g.qubits[1]]
            R = gate_ptm16(g).reshape(4, 4, 4, 4)
            M = np.tensordot(R, M, axes=([2, 3], [a0, a1]))
            M = np.moveaxis(M, [0, 1], [a0, a1])
    return np.ascontiguousarray(M.reshape(dim, dim))


def qubit_masks(gates):
    """Bitmask of each gate's qubits (the planners' scans test masks)."""
    return [sum(1 << q for q in g.qubits) for g in gates]


def _grow(gates, done, start, kmax, n, free=frozenset(), masks=None):
    """Largest block (<= kmax qubits outside ``free``) seeded at the first
    pending gate ``start``.  Scan forward: a gate joins when none of its
    qubits is blocked and the union stays within kmax; a gate that cannot
    join blocks all of its qubits (every later gate on them must run after
    it).  Returns the qubits counted against kmax and the members; with
    ``free`` the scan stops as soon as the counted qubits are final, so only
    they are complete (the caller recollects the members)."""
    if masks is None:
        masks = qubit_masks(gates)
    fm = sum(1 << q for q in free)
    full = (1 << n) - 1
    qm = 0
    qs = []
    members = []
    blocked = 0
    for i in range(start, len(gates)):
        if done[i]:
            continue
        gm = masks[i]
        if gm & blocked:
            blocked |= gm
        else:
            new = gm & ~qm & ~fm
            if len(qs) + new.bit_count() <= kmax:
                if new:
                    qs.extend(q for q in gates[i].qubits if (new >> q) & 1)
                    qm |= new
                members.append(i)
            else:
                blocked |= gm
        if fm:
            if len(qs) >= kmax or (blocked | qm | fm) == full:
                break  # no further counted qubit can join
        elif qm and not (qm & ~blocked) and (len(qs) >= kmax or (blocked | qm) == full):
            break  # nothing further can join
    return qs, members


def _absorb(gates, done, start, qset, masks=None):
    """Gates of the block on qubit set ``qset`` seeded at the first pending
    gate ``start``: scanning forward, a gate joins when its qubits lie in the
    set and none is blocked; any other gate blocks all of its qubits."""
    if masks is None:
        masks = qubit_masks(gates)
    S = sum(1 << q for q in qset)
    members = []
    blocked = 0
    for i in range(start, len(gates)):
        if done[i]:
            continue
        gm = masks[i]
        if (gm & blocked) or (gm & ~S):
            blocked |= gm
            if not (S & ~blocked):
                break
        else:
            members.append(i)
    return members


def _near_qubits(gates, done, start, seed, limit=6, window=48):
    """Up to ``limit`` other qubits, in order of first appearance, of the
    pending two-qubit gates within ``window`` gates that connect to the seed
    qubits (directly or through qubits already collected)."""
    reach = set(seed)
    out = []
    seen = 0
    for i in range(start, len(gates)):
        if done[i]:
            continue
        seen += 1
        if seen > window or len(out) >= limit:
            break
        gq = gates[i].qubits
        if len(gq) == 2 and (gq[0] in reach or gq[1] in reach):
            for q in gq:
                if q not in reach:
                    reach.add(q)
                    out.append(q)
    return out[:limit]


def low_qubits(n):
    """Qubits at the LOW_DIGITS lowest positions."""
    return frozenset(range(max(0, n - LOW_DIGITS), n))


def tile_qubits(n, T, precision='fp32'):
    """Qubits of the tile a sequential pass over the qubits ``T`` (outside the
    low positions) runs in: T, the low positions, and the next lowest
    positions up to 6 digits (|T| <= 2, and |T| = 3 in FP64) or 7 digits
    (|T| = 4, and |T| = 3 in FP32, whose runs are then 4^4 coefficients).
    Its runs are at least 4^3 coefficients long."""
    m = min(n, 6 if len(T) <= 2 or (len(T) == 3 and precision == 'fp64') else 7)
    S = set(T) | low_qubits(n)
    p = 0
    while len(S) < m:
        S.add(n - 1 - p)
        p += 1
    return S


def used_qubits(gates, members):
    """Distinct qubits of the member gates, in order of first use."""
    qs = []
    for i in members:
        for q in gates[i].qubits:
            if q not in qs:
                qs.append(q)
    return qs


def candidate_blocks(gates, done, start, kmax, n, tiles=False, masks=None, precision='fp32'):
    """Distinct candidate blocks (qubits, members) seeded at ``start``.

    The greedy growth for every k <= kmax, and every seed-containing set of
    at most kmax qubits drawn from the nearby qubits (its members by
    ``_absorb``).  With ``tiles`` (sequential tile passes, n >=
    MIN_TILE_QUBITS) also the tile-wide blocks: for each set T of qubits
    outside the low positions (the greedy growth counting only such qubits,
    for every k, and the seed's qubits with nearby ones), every gate
    ``_absorb`` admits into T's whole tile (``tile_qubits``).  A block's
    qubits are those its members use; the single seed gate is always a
    candidate.  ``precision`` selects the tile sizes (``tile_qubits``)."""
    if masks is None:
        masks = qubit_masks(gates)
    seed = list(gates[start].qubits)
    cands = {}

    def add(members):
        key = tuple(members)
        if key in cands or not members:
            return
        cands[key] = (used_qubits(gates, members), members)

    add([start])
    for k in range(2, kmax + 1):
        add(_grow(gates, done, start, k, n, masks=masks)[1])
    near = _near_qubits(gates, done, start, seed)
    for extra in range(1, kmax - len(seed) + 1):
        for T in itertools.combinations(near, extra):
            add(_absorb(gates, done, start, seed + list(T), masks))
    if not tiles:
        return list(cands.values())
    low = low_qubits(n)
    sets = set()
    for k in range(1, kmax + 1):
        sets.add(frozenset(_grow(gates, done, start, k, n, low, masks)[0]))
    hi = [q for q in seed if q not in low]
    near = [q for q in _near_qubits(gates, done, start, seed, limit=6 + LOW_DIGITS) if q not in low][:6]
    for extra in range(0, kmax - len(hi) + 1):
        for X in itertools.combinations(near, extra):
            sets.add(frozenset(hi) | frozenset(X))
    for T in sorted(sets, key=sorted):
        add(_absorb(gates, done, start, tile_qubits(n, T, precision), masks))
    return list(cands.values())


# blocks of 3-4 qubits run as tile passes of at least 6 digits (the MFMA form
# needs k + column digits = 3+3 or 4+2), so smaller registers use k <= 2
MIN_TILE_QUBITS = 6


def _state_after(gates, done, first, members):
    """done list and first pending gate once ``members`` have run."""
    d2 = list(done)
    for j in members:
        d2[j] = True
    while first < len(gates) and d2[first]:
        first += 1
    return d2, first


def _lookahead_pick(gates, done, first, cands, score, per_gate, widths, ctx):
    """Index of the learned planner's choice among ``cands`` (scored by
    ``score``), and the candidates of the state after it if the search built
    them, as (first pending gate, candidates) or None.

    Each of the widths[0] best candidates is credited with the best total
    saving of the blocks that can follow it, searched over the widths[1:]
    best at each further step; the step after the last width takes its best
    alone.  With no widths (or a single width of 1) the best score wins."""
    km, n, tiles, masks, cost = ctx
    top = sorted(range(len(cands)), key=lambda i: -score[i])[:max(1, widths[0] if widths else 1)]
    if not (len(widths) >= 2 or (widths and widths[0] > 1)):
        return top[0], None
    seen = {}  # path of member tuples -> (first pending gate, candidates, scores)

    def best_after(d2, f2, depth, key):
        if f2 >= len(gates):
            return 0.0
        if key not in seen:
            c2 = candidate_blocks(gates, d2, f2, km, n, tiles, masks, cost.precision)
            t2 = cost.block_times(n, [(q2, [gates[j] for j in m2]) for q2, m2 in c2])
            seen[key] = (f2, c2, [len(m2) * per_gate - TIME_WEIGHT * tt for (_, m2), tt in zip(c2, t2)])
        _, c2, s2 = seen[key]
        if depth >= len(widths):
            return max(s2)
        order = sorted(range(len(c2)), key=lambda i: -s2[i])[:widths[depth]]
        return max(s2[i] + best_after(*_state_after(gates, d2, f2, c2[i][1]), depth + 1, key + tuple(c2[i][1]))
                   for i in order)

    total = {i: score[i] + best_after(*_state_after(gates, done, first, cands[i][1]), 1, tuple(cands[i][1]))
             for i in top}
    pick = max(top, key=lambda i: total[i])
    nxt = seen.get(tuple(cands[pick][1]))
    return pick, ((nxt[0], nxt[1]) if nxt is not None else None)


def plan(gates, n, kmax=4, method='greedy', cost=None, tiles=True, lookahead=LOOKAHEAD, precision=None):
    """Blocks as (qubits, [gate indices]) in execution order, from the native
    planner (dm_plan_fusion, the same rules as ``plan_py``).  With ``tiles``
    (blocks run as sequential tile passes of the default size) the learned
    planner's candidates include the tile-wide blocks (``candidate_blocks``);
    ``lookahead`` gives its search widths (LOOKAHEAD), () none.  The state's
    ``precision`` defaults to the cost model's, else FP32."""
    if method not in native.PLAN_METHODS:
        raise ValueError('unknown fusion method %r' % method)
    prec = precision or (cost.precision if cost is not None else 'fp32')
    gq = [(g.qubits[0], g.qubits[1] if len(g.qubits) == 2 else -1) for g in gates]
    if method == 'learned':
        return native.plan_fusion(n, gq, method, kmax, tiles, tuple(lookahead), cost.params, prec)
    return native.plan_fusion(n, gq, method, kmax, tiles, precision=prec)


def plan_py(gates, n, kmax=4, method='greedy', cost=None, tiles=True, lookahead=LOOKAHEAD):
    """The planner in Python (the cross-check of the native one): blocks as
    (qubits, [gate indices]) in execution order."""
    if any(len(g.qubits) > kmax for g in gates):
        raise ValueError('kmax=%d cannot hold a gate on 2 qubits' % kmax)
    done = [False] * len(gates)
    blocks = []
    first = 0
    allow_mfma = n >= MIN_TILE_QUBITS
    tiles = tiles and allow_mfma and method == 'learned'
    masks = qubit_masks(gates)
    spent = [0.0, 0]  # predicted seconds and gates of the blocks chosen so far
    ahead = None  # candidates of the current state, found by the last lookahead
    while True:
        while first < len(gates) and done[first]:
            first += 1
        if first >= len(gates):
            break
        if method == 'greedy':
            km = kmax if allow_mfma else min(kmax, 2)
            qs, mem = _grow(gates, done, first, km, n, masks=masks)
        elif method in ('learned', 'search'):
            km = kmax if allow_mfma else min(kmax, 2)
            if ahead is not None and ahead[0] == first:
                cands = ahead[1]
            else:
                cands = candidate_blocks(gates, done, first, km, n, tiles, masks,
                                         cost.precision if cost is not None else 'fp32')
            ahead = None
            if method == 'learned':
                # predicted net saving: the gates a block takes, valued at the
                # plan's running cost per gate, minus its weighted own time
                t = cost.block_times(n, [(qs_c, [gates[i] for i in mem_c]) for qs_c, mem_c in cands])
                g = [len(c[1]) for c in cands]
                if spent[1]:
                    per_gate = spent[0] / spent[1]
                else:
                    j = int(np.argmax(g))
                    per_gate = t[j] / g[j]
                score = [gc * per_gate - TIME_WEIGHT * tc for gc, tc in zip(g, t)]
                pick, ahead = _lookahead_pick(gates, done, first, cands, score, per_gate, tuple(lookahead),
                                              (km, n, tiles, masks, cost))
                spent[0] += t[pick]
                spent[1] += g[pick]
            else:
                # the same search with every pass costed alike: most gates
                pick = int(np.argmax([len(c[1]) for c in cands]))
            qs, mem = cands[pick]
        else:
            raise ValueError('unknown fusion method %r' % method)
        for i in mem:
            done[i] = True
        blocks.append((qs, mem))
    return blocks


def compose_single_qubit_runs(gates):
    """Same gates with each run of 1-qubit PTMs on one qubit (no two-qubit gate
    on that qubit in between) folded into one PTM; gates on different qubits
    commute, so each folded PTM is emitted just before the next two-qubit
    gate on its qubit (or at the end)."""
    pending = {}
    out = []
    for g in gates:
        if len(g.qubits) == 1:
            q = g.qubits[0]
            pending[q] = g.R @ pending[q] if q in pending else g.R
            continue
        for q in g.qubits:
            if q in pending:
                out.append(NoisyGate((q,), pending.pop(q), 'u3'))
        out.append(g)
    for q in sorted(pending):
        out.append(NoisyGate((q,), pending[q], 'u3'))
    return out


def _subop(g):
    if g.kind == 'u3':
        return (native.SUB_1Q, list(g.qubits), ptm.rows3(g.R))
    if g.kind == 'cx':
        return (native.SUB_CX, list(g.qubits), g.R)
    return (native.SUB_2Q, list(g.qubits), g.R)


def run_digits(n, positions, m):
    """w: the number of lowest positions a tile of m digits holds when it is
    the block's positions completed with the lowest other positions; a tile
    reads contiguous runs of 4^w coefficients from HBM."""
    inside = set(positions)
    p = 0
    while len(inside) < m and p < n:
        inside.add(p)
        p += 1
    w = 0
    while w in inside:
        w += 1
    return w


def tile_for(n, qs, precision='fp32'):
    """Tile size of a sequential pass over block ``qs``: 6 digits (32 KB of
    LDS in FP64) unless that leaves runs of only 4^2 coefficients, which
    stream at about two thirds of the rate of longer runs, or (FP32) 7 digits
    turn runs of 4^3 into runs of 4^4 (measured faster on MI355X); then 7
    digits (512 threads, 64 / 128 KB) when the register allows it.  A block
    of 7 qubits takes 7 digits."""
    pos = [n - 1 - q for q in qs]
    w6, w7 = run_digits(n, pos, 6), run_digits(n, pos, 7)
    if n >= 7 and (len(qs) > 6 or (w6 < 3 and w7 > w6) or (precision != 'fp64' and w6 == 3 and w7 == 4)):
        return 7
    return 6


def add_block(prog, gates, qs, mem, mode='seq', tile=None, precision='fp32'):
    """Append one block to a Program.

    A single gate runs on its plain kernel.  Otherwise, with n >= 6:
      mode 'seq'   -- one tile pass applying the gates one after another in LDS
                      (tile digits: ``tile``, or ``tile_for`` when None);
      mode 'dense' -- the block PTM on the matrix cores (dm_apply_block; 3-4
                      qubits; 1-2 qubit blocks use the dense 16x16 kernel).
    """
    if len(mem) == 1:
        g = gates[mem[0]]
        if g.kind == 'u3':
            prog.one_qubit(g.qubits[0], g.R)
        elif g.kind == 'cx':
            prog.cx(g.qubits[0], g.qubits[1], g.R)
        else:
            prog.two_qubit(g.qubits[0], g.qubits[1], g.R)
        return
    sub = [gates[i] for i in mem]
    n = prog.n
    if tile is None:
        tile = tile_for(n, qs, precision) if n >= 6 else 6
    if len(qs) == 1:
        prog.one_qubit(qs[0], block_ptm(sub, qs))
    elif (mode == 'seq' or len(qs) > 4) and n >= tile:
        prog.fused(tile, qs, [_subop(g) for g in compose_single_qubit_runs(sub)])
    elif len(qs) == 2:
        # block_ptm's pair index is i_qs0 + 4 i_qs1: qs[1] is the leading qubit
        prog.two_qubit(qs[1], qs[0], block_ptm(sub, qs).reshape(16, 16))
    elif n >= 6:
        # the standalone matrix-core block kernel: 2.4x faster than the dense
        # operation inside a tile pass (16 qubits, k = 4, MI355X)
        prog.block(qs, block_ptm(sub, qs))
    else:
        for i in mem:
            add_block(prog, gates, list(gates[i].qubits), [i])


def plain_program(gates, n):
    """One pass per noisy gate."""
    prog = Program(n)
    for i in range(len(gates)):
        add_block(prog, gates, list(gates[i].qubits), [i])
    return prog


def fused_program(gates, n, kmax=4, method='greedy', cost=None, mode='seq', tile=None, lookahead=LOOKAHEAD,
                  precision=None):
    """Program of fused passes; with a cost model each block also gets the
    cheaper of the sequential and dense forms.  ``tile`` fixes the tile size
    of sequential passes (default: ``tile_for`` for the state's
    ``precision``, by default the cost model's, else FP32); blocks take the
    gates of their whole tile only with the default size in mode 'seq'."""
    prec = precision or (cost.precision if cost is not None else 'fp32')
    blocks = plan(gates, n, kmax, method, cost, tiles=(mode == 'seq' and tile is None), lookahead=lookahead,
                  precision=prec)
    prog = Program(n)
    for qs, mem in blocks:
        m = mode
        if cost is not None and len(mem) > 1:
            m = cost.best_mode(n, qs, [gates[i] for i in mem])
        add_block(prog, gates, qs, mem, m, tile, prec)
    return prog, blocks


def _fusable(level):
    return bool(level) and all(op.name in ('u1', 'u3', 'cx') and getattr(op, 'conditional', None) is None
                               for op in level)


def plan_levels(levels, count, n, backend):
    """Fused execution of the backend's level loop: maximal runs of levels that
    hold only u1/u3/cx become one fused program; memory noise between levels
    enters the gate stream as one 1-qubit channel per qubit and level.
    Returns {first_clock: (next_clock, runner)}."""
    f, p, g = backend._decoherence_factor, backend._thermal_factor, backend._decay_factor
    noisy_levels = not (f == 1 and g == 1)
    mem_R = ptm.memory_noise(f, p, g)
    method = backend._fusion
    cost = None
    if method == 'learned':
        from . import costmodel
        cost = costmodel.default_model()
        cost.precision = backend._precision
    out = {}
    clock = 0
    while clock < count:
        if not _fusable(levels[clock]):
            clock += 1
            continue
        start = clock
        gates = []
        while clock < count and _fusable(levels[clock]):
            seq = [(op.name, tuple(op.qubits), tuple(float(x) for x in getattr(op, 'params', ())))
                   for op in levels[clock]]
            gates += noisy_gates(seq, backend._gate_depolarization, backend._rotation_error,
                                 backend._tsp_model_error)
            if noisy_levels:
                gates += [NoisyGate((q,), mem_R, 'u3') for q in range(n)]
            clock += 1
        # below 15 qubits a pass takes about a millisecond and the lookahead's
        # host time (about 2x the plain choice) costs more than it saves
        prog, _ = fused_program(gates, n, KMAX[method], method, cost, lookahead=LOOKAHEAD if n >= 15 else (),
                                precision=backend._precision)
        out[start] = (clock, prog.run)
    return out
