"""Op-record programs against the CPU oracle, without a GPU.

A numpy interpreter of the dm_op records (DM_OP_1Q, DM_OP_CX[_SCALED],
DM_OP_2Q, DM_OP_BLOCK, DM_OP_MEMORY_NOISE and DM_OP_FUSED with its 1-qubit,
CX, 16x16 and dense sub-operations, applied in record order; the tile
geometry does not change the result) runs the programs of small circuits.

* Batch circuit programs built by the library (dm_batch_circuit_program):
  their final Pauli coefficients must match the oracle's simulation of the
  same circuit with a barrier after every gate (refsim.run_experiment: the
  gate, then memory noise on every qubit, at every level;
  dm_simulator.py:1005-1177) to 1e-13, in both program forms (one record
  per gate and level, and fused tile passes) and with every noise option
  the builder takes.  This pins the builder's gate PTMs, the order of the
  per-gate depolarising channel and the memory noise, the CX pair scales,
  the folding of single-qubit runs and the fused-record layout against the
  reference's semantics rather than against the Python builders.
* Fused programs of noisy gate lists (fusion.fused_program: the greedy and
  learned planners, sequential and dense blocks, TSP-error CX as 16x16
  PTMs): they must match the oracle's gate functions applied gate by gate
  (rotations, CX, per-gate depolarising; basicaertools.py:69-126, 310-392)
  to 1e-13.
* The backend's level fusion (fusion.plan_levels: runs of u1/u3/cx levels
  with the memory noise of every level as 1-qubit channels): the programs
  of a whole circuit must match the oracle's run_experiment (merge,
  partition, level loop with memory noise after every level) to 1e-12
  with every noise option.
"""
import numpy as np
import pytest

from dmsim import assemble, batch, native, workloads
from oracle import refsim


def _apply_ptm_rows(st, n, q, rows):
    """X, Y, Z rows of a trace-preserving 4x4 PTM on qubit q (axis q)."""
    R = np.vstack([[1.0, 0.0, 0.0, 0.0], np.asarray(rows).reshape(3, 4)])
    return np.moveaxis(np.tensordot(R, st, axes=([1], [q])), 0, q)


def _apply_dense(st, n, qs, S):
    """A 4^k x 4^k PTM on qubits qs, component index sum_j i_{qs[j]} 4^j."""
    k = len(qs)
    S = np.asarray(S).reshape((4,) * (2 * k))   # (out_{k-1} .. out_0, in_{k-1} .. in_0)
    axes = [qs[j] for j in range(k - 1, -1, -1)]
    moved = np.moveaxis(st, axes, list(range(k)))
    out = np.tensordot(S, moved, axes=(list(range(k, 2 * k)), list(range(k))))
    return np.moveaxis(out, list(range(k)), axes)


def _apply_cx(st, n, c, t, scale):
    st = refsim.apply_cx(st.reshape(-1), c, t, (1.0, 0.0), n)
    if scale is not None:
        s = np.asarray(scale).reshape(4, 4)  # [control component, target component]
        shape = [1] * n
        shape[c], shape[t] = 4, 4
        s_b = s if c < t else s.T
        st = st * s_b.reshape(shape)
    return st


def run_records(n, prog, st=None):
    """Final coefficients (shape n*[4]) of ``st`` (default |0...0>) under
    the program."""
    ops, vals = prog.compile()
    st = (refsim.initial_state(n) if st is None else st).reshape(n * [4])
    for i in range(len(prog)):
        o = ops[i]
        p = o.param
        if o.kind == native.OP_1Q:
            st = _apply_ptm_rows(st, n, o.qubits[0], vals[p:p + 12])
        elif o.kind == native.OP_CX:
            st = _apply_cx(st, n, o.qubits[0], o.qubits[1], None)
        elif o.kind == native.OP_CX_SCALED:
            st = _apply_cx(st, n, o.qubits[0], o.qubits[1], vals[p:p + 16])
        elif o.kind == native.OP_2Q:
            # pair index 4 i_qa + i_qb
            st = _apply_dense(st, n, [o.qubits[1], o.qubits[0]], vals[p:p + 256])
        elif o.kind == native.OP_BLOCK:
            k = o.k
            st = _apply_dense(st, n, list(o.qubits[:k]), vals[p:p + 4 ** (2 * k)])
        elif o.kind == native.OP_MEMORY_NOISE:
            st = refsim.apply_memory_noise(st.reshape(-1), n, vals[p], vals[p + 1], vals[p + 2])
        elif o.kind == native.OP_FUSED:
            m, nq = int(vals[p]), int(vals[p + 1])
            assert m in (5, 6, 7) and 1 <= nq <= 7
            at = p + 2 + nq
            nsub = int(vals[at])
            at += 1
            for _ in range(nsub):
                kind, kq = int(vals[at]), int(vals[at + 1])
                q = [int(x) for x in vals[at + 2:at + 2 + kq]]
                nv = int(vals[at + 6])
                v = vals[at + 7:at + 7 + nv]
                if kind == native.SUB_1Q:
                    assert kq == 1 and nv == 12
                    st = _apply_ptm_rows(st, n, q[0], v)
                elif kind == native.SUB_CX:
                    assert kq == 2 and nv in (0, 16)
                    st = _apply_cx(st, n, q[0], q[1], v if nv else None)
                elif kind == native.SUB_2Q:
                    assert kq == 2 and nv == 256
                    st = _apply_dense(st, n, [q[1], q[0]], v)
                else:
                    assert kind == native.SUB_DENSE and nv == 4 ** (2 * kq)
                    st = _apply_dense(st, n, q, v)
                at += 7 + nv
            assert at == (ops[i + 1].param if i + 1 < len(prog) else vals.size)
        else:
            raise AssertionError('unexpected op kind %d' % o.kind)
        st = np.reshape(st, n * [4])
    return st


def oracle_state(n, gates, **opts):
    qc = workloads.to_circuit(n, gates, barrier_after_each=True)
    exp = assemble(qc).experiments[0]
    data, _, _ = refsim.run_experiment(exp.instructions, n, dict(opts, compute_densitymatrix=False))
    return np.asarray(data['coeffmatrix']).reshape(n * [4])


def _mixed_gates(n, count, seed):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(count):
        r = rng.random()
        if r < 0.15:
            gates.append(('u1', (int(rng.integers(n)),), (float(rng.uniform(0, 6.3)),)))
        elif r < 0.6:
            gates.append(('u3', (int(rng.integers(n)),), tuple(float(x) for x in rng.uniform(0, 6.3, 3))))
        else:
            c, t = rng.choice(n, 2, replace=False)
            gates.append(('cx', (int(c), int(t)), ()))
    return gates


@pytest.mark.parametrize('fused', [False, True])
def test_native_records_match_the_oracle_on_the_batch_workload(fused):
    specs = workloads.batch(6, 21, n_range=(6, 8), depth_range=(30, 70))
    for n, gates, dr, _ in specs:
        prog = batch.native_program(n, gates, decay_factor=1.0 - dr, fused=fused)
        got = run_records(n, prog)
        want = oracle_state(n, gates, decay_factor=1.0 - dr)
        assert np.max(np.abs(got - want)) <= 1e-13, (n, len(gates))


@pytest.mark.parametrize('fused', [False, True])
def test_native_records_match_the_oracle_with_every_noise_option(fused):
    """Per-gate depolarising (1q rows and CX pair scales), dephasing and a
    thermal feed on the memory noise, u1 gates; and noiseless memory (no
    channels in the fused form)."""
    n = 7
    gates = _mixed_gates(n, 60, 4)
    noise = dict(decay_factor=0.97, decoherence_factor=0.95, thermal_factor=0.8)
    prog = batch.native_program(n, gates, gate_depolarization=2e-3, fused=fused, **noise)
    want = oracle_state(n, gates, gate_depolarization=2e-3, **noise)
    assert np.max(np.abs(run_records(n, prog) - want)) <= 1e-13
    prog = batch.native_program(n, gates, fused=fused)
    assert np.max(np.abs(run_records(n, prog) - oracle_state(n, gates))) <= 1e-13


def test_interpreter_sees_a_planted_mistake():
    """A record with one sub-operation's control and target swapped no longer
    matches: the comparison is sharp enough to see a wrong CX direction."""
    n = 6
    gates = _mixed_gates(n, 40, 9)
    prog = batch.native_program(n, gates, decay_factor=0.99, fused=True)
    ops, vals = prog.compile()
    cx_at = []  # value offsets of every fused CX sub-operation
    for i in range(len(prog)):
        if ops[i].kind != native.OP_FUSED:
            continue
        p = ops[i].param
        at = p + 3 + int(vals[p + 1])
        for _ in range(int(vals[at - 1])):
            if int(vals[at]) == native.SUB_CX:
                cx_at.append(at)
            at += 7 + int(vals[at + 6])
    assert cx_at
    # the last one acts mid-circuit, on qubits no longer in |00> (on |00>
    # both directions of a CX agree)
    at = cx_at[-1]
    bad = vals.copy()
    bad[at + 2], bad[at + 3] = vals[at + 3], vals[at + 2]
    prog._compiled = (ops, bad)
    want = oracle_state(n, gates, decay_factor=0.99)
    assert np.max(np.abs(run_records(n, prog) - want)) > 1e-6


def oracle_gates(n, gates, p, tsp=(1.0, 0.0)):
    """The oracle's gate functions, gate by gate: rotations of u1 / u3, CX
    with the TSP error, then the per-gate depolarising channel."""
    st = refsim.initial_state(n)
    for name, qs, prm in gates:
        if name == 'cx':
            st = refsim.apply_cx(np.reshape(st, -1), qs[0], qs[1], tsp, n)
        else:
            for axis, ang in refsim.gate_rotations(name, prm):
                q = qs[0]
                st3 = np.reshape(st, (4 ** q, 4, 4 ** (n - q - 1)))
                st = refsim.rotate_components(st3, axis, ang, (1.0, 0.0))
        if p:
            st = refsim.apply_gate_depolarizing(np.reshape(st, -1), n, list(qs), p)
    return np.reshape(st, n * [4])


def _sub_kinds(prog):
    ops, vals = prog.compile()
    kinds = set()
    for i in range(len(prog)):
        if ops[i].kind != native.OP_FUSED:
            continue
        p = ops[i].param
        at = p + 3 + int(vals[p + 1])
        for _ in range(int(vals[at - 1])):
            kinds.add(int(vals[at]))
            at += 7 + int(vals[at + 6])
    return kinds


@pytest.mark.parametrize('method,kmax,mode,tsp,want_ops,want_subs', [
    ('greedy', 4, 'seq', (1.0, 0.0), {native.OP_FUSED}, {native.SUB_1Q, native.SUB_CX}),
    ('greedy', 4, 'dense', (1.0, 0.0), {native.OP_BLOCK}, set()),
    ('greedy', 2, 'dense', (1.0, 0.0), {native.OP_2Q}, set()),
    ('learned', 4, 'seq', (1.0, 0.0), {native.OP_FUSED}, {native.SUB_1Q, native.SUB_CX}),
    ('greedy', 4, 'seq', (0.97, 0.03), {native.OP_FUSED}, {native.SUB_1Q, native.SUB_2Q})])
@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_fused_programs_match_the_oracle(method, kmax, mode, tsp, want_ops, want_subs, precision):
    """Sequential tile passes (6- and 7-digit, the learned planner's
    tile-wide blocks), dense 3-4 qubit blocks and 2-qubit PTMs, and the
    TSP-error CX inside tile passes."""
    from dmsim import costmodel, fusion
    n = 8
    gates = workloads.local_u3_cx(n, 120, 5) + [('u1', (3,), (0.7,)), ('cx', (7, 2), ())]
    noisy = fusion.noisy_gates(gates, 1e-3, tsp=tsp)
    cost = None
    if method == 'learned':
        cost = costmodel.default_model()
        cost.precision = precision
    prog, blocks = fusion.fused_program(noisy, n, kmax, method, cost, mode, precision=precision)
    assert want_ops <= set(prog.kinds()) and want_subs <= _sub_kinds(prog)
    got = run_records(n, prog)
    assert np.max(np.abs(got - oracle_gates(n, gates, 1e-3, tsp))) <= 1e-13


@pytest.mark.parametrize('method', ['greedy', 'learned'])
@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_backend_level_fusion_matches_the_oracle(method, precision):
    """Levels of several gates (no barriers), single-qubit runs merged, and
    every noise option: thermal memory noise, dephasing, per-gate
    depolarising, rotation error, TSP CX error."""
    import types
    from dmsim import fusion, schedule
    n = 7
    gates = _mixed_gates(n, 90, 12)
    opts = {'decay_factor': 0.98, 'decoherence_factor': 0.96, 'thermal_factor': 0.85,
            'gate_depolarization': 1e-3, 'rotation_error': {'rx': [1.0, 0.0], 'ry': [0.999, 0.01], 'rz': [0.998, -0.02]},
            'tsp_model_error': [0.99, 0.02]}
    backend = types.SimpleNamespace(
        _decoherence_factor=opts['decoherence_factor'], _thermal_factor=opts['thermal_factor'],
        _decay_factor=opts['decay_factor'], _fusion=method, _precision=precision,
        _gate_depolarization=opts['gate_depolarization'], _rotation_error=opts['rotation_error'],
        _tsp_model_error=opts['tsp_model_error'])
    qc = workloads.to_circuit(n, gates)
    insts = assemble(qc).experiments[0].instructions
    levels, count = schedule.partition(schedule.merge_single_gates(insts, n, True), n)
    plan = fusion.plan_levels(levels, count, n, backend)
    st, clock = None, 0
    while clock < count:
        nxt, runner = plan[clock]   # every level holds only u1/u3/cx
        st = run_records(n, runner.__self__, st)
        clock = nxt
    data, ocount, _ = refsim.run_experiment(insts, n, dict(opts, compute_densitymatrix=False))
    assert ocount == count
    want = np.asarray(data['coeffmatrix']).reshape(n * [4])
    assert np.max(np.abs(st - want)) <= 1e-12


def test_native_records_over_random_configurations():
    """24 seeded configurations: widths 6-9, depths 10-60, random memory
    noise (decay, dephasing, thermal feed or none), per-gate depolarising on
    or off, both forms and both planning precisions."""
    rng = np.random.default_rng(2024)
    for trial in range(24):
        n = int(rng.integers(6, 10))
        gates = _mixed_gates(n, int(rng.integers(10, 61)), 1000 + trial)
        if rng.random() < 0.25:
            noise = {}
        else:
            noise = dict(decay_factor=float(rng.uniform(0.95, 1.0)), decoherence_factor=float(rng.uniform(0.95, 1.0)),
                         thermal_factor=float(rng.choice([1.0, rng.uniform(0.5, 1.0)])))
        gd = float(rng.choice([0.0, rng.uniform(1e-4, 1e-2)]))
        fused = bool(rng.integers(2))
        precision = 'fp32' if rng.integers(2) else 'fp64'
        prog = batch.native_program(n, gates, gate_depolarization=gd, fused=fused, precision=precision, **noise)
        want = oracle_state(n, gates, gate_depolarization=gd, **noise)
        assert np.max(np.abs(run_records(n, prog) - want)) <= 1e-13, (trial, n, fused, precision, noise, gd)
