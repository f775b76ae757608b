"""The native host front end (merge, partition, fusion planner, cost model)
against the same passes in Python, on the CPU.

* merge: same records, names and qubits; merged angles to 1e-14 (the C
  library's atan2 / acos against numpy's, a few ulp apart);
* partition: identical level lists and level count;
* planner: identical blocks (members and qubit order) for greedy, search
  and the learned planner with the shipped cost model (both evaluate the
  network with the host forward pass), with and without lookahead;
* cost model: the native block seconds equal CostModel.block_times;
* batch circuit programs: the same op records (kinds, qubits, value
  offsets, fused-record layout) as the Python builders, values to 1e-14
  (the C library's sin / cos against numpy's, and 4x4 products summed in
  another order).
"""
import copy
import ctypes
import time

import numpy as np
import pytest

from dmsim import assemble, batch, costmodel, fusion, native, schedule, workloads

from test_frontend import FIX, assert_same_merge, build_fixture_circuit, random_circuit


@pytest.mark.parametrize('merge', [True, False])
def test_merge_matches_python(merge):
    rng = np.random.default_rng(11)
    for trial in range(60):
        n = int(rng.integers(2, 8))
        qc = random_circuit(rng, n, 80)
        insts = assemble(qc).experiments[0].instructions
        a = schedule.merge_single_gates(copy.deepcopy(insts), n, merge)
        b = schedule.merge_single_gates_py(copy.deepcopy(insts), n, merge)
        assert_same_merge(a, b)


def _partition_inputs(rng, n, count):
    qc = random_circuit(rng, n, count)
    insts = assemble(qc).experiments[0].instructions
    return schedule.merge_single_gates_py(copy.deepcopy(insts), n)


def test_partition_matches_python_on_random_circuits():
    rng = np.random.default_rng(12)
    for trial in range(80):
        n = int(rng.integers(2, 8))
        merged = _partition_inputs(rng, n, int(rng.integers(1, 90)))
        la, ca = schedule.partition(merged, n)
        lb, cb = schedule.partition_py(merged, n)
        assert ca == cb, trial
        assert len(la) == len(lb), trial
        assert [[id(o) for o in x] for x in la] == [[id(o) for o in y] for y in lb], trial


@pytest.mark.parametrize('f', FIX['fixtures'], ids=lambda f: f['name'])
def test_partition_matches_python_on_fixtures(f):
    qc = build_fixture_circuit(f)
    n = qc.num_qubits
    merged = schedule.merge_single_gates_py(copy.deepcopy(assemble(qc).experiments[0].instructions), n)
    la, ca = schedule.partition(merged, n)
    lb, cb = schedule.partition_py(merged, n)
    assert ca == cb
    assert [[id(o) for o in x] for x in la] == [[id(o) for o in y] for y in lb]


def _gates(n, G, seed, kind):
    if kind == 'local':
        return fusion.noisy_gates(workloads.local_u3_cx(n, G, seed), 1e-3)
    if kind == 'matching':
        return fusion.noisy_gates(workloads.matching_layers(n, max(1, G // (n + n // 2)), seed), 5e-4)
    return fusion.noisy_gates(workloads.random_u3_cx(n, G, seed), 1e-3)


@pytest.mark.parametrize('method', ['greedy', 'search'])
def test_plan_matches_python(method):
    for n, G, seed, kind in ((4, 80, 1, 'random'), (6, 120, 2, 'local'), (9, 150, 3, 'random'),
                             (12, 200, 4, 'matching'), (16, 200, 5, 'local'), (5, 60, 6, 'local')):
        gates = _gates(n, G, seed, kind)
        for tiles in (True, False):
            assert fusion.plan(gates, n, 4, method, tiles=tiles) == fusion.plan_py(gates, n, 4, method, tiles=tiles)


@pytest.mark.parametrize('lookahead', [(), (3,), (3, 2)])
@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_learned_plan_matches_python(lookahead, precision):
    cost = costmodel.CostModel.load(precision=precision)
    for n, G, seed, kind in ((7, 100, 7, 'local'), (10, 150, 8, 'random'), (16, 160, 9, 'local'),
                             (13, 150, 10, 'matching')):
        gates = _gates(n, G, seed, kind)
        for tiles in (True, False):
            a = fusion.plan(gates, n, 4, 'learned', cost, tiles=tiles, lookahead=lookahead)
            b = fusion.plan_py(gates, n, 4, 'learned', cost, tiles=tiles, lookahead=lookahead)
            assert a == b, (n, kind, tiles)


def test_plan_refuses_blocks_too_small_for_a_gate():
    """kmax = 1 cannot hold a CX: the planner refuses instead of emitting
    empty blocks without end (which overran its output arrays)."""
    gates = _gates(6, 20, 3, 'local')
    with pytest.raises(native.NativeError, match='kmax=1 cannot hold'):
        fusion.plan(gates, 6, 1, 'greedy', tiles=False)
    with pytest.raises(ValueError, match='kmax=1 cannot hold'):
        fusion.plan_py(gates, 6, 1, 'greedy', tiles=False)
    ones = [g for g in gates if len(g.qubits) == 1]
    assert fusion.plan(ones, 6, 1, 'greedy', tiles=False) == fusion.plan_py(ones, 6, 1, 'greedy', tiles=False)


def test_cost_block_seconds_match_python():
    cost = costmodel.CostModel.load(precision='fp32')
    n = 14
    gates = _gates(n, 150, 12, 'local')
    blocks = fusion.plan(gates, n, 4, 'search')
    gq = [(g.qubits[0], g.qubits[1] if len(g.qubits) == 2 else -1) for g in gates]
    got = native.cost_block_seconds(n, gq, blocks, cost.params, 'fp32')
    want = cost.block_times(n, [(qs, [gates[i] for i in mem]) for qs, mem in blocks])
    assert np.array_equal(got, np.array(want))


def test_host_forward_matches_the_oracle():
    from oracle import mlp as omlp
    cost = costmodel.CostModel.load()
    X = np.random.default_rng(3).random((50, costmodel.N_FEATURES)).astype(np.float32)
    got = native.mlp_predict_host(cost.params, costmodel.N_FEATURES, costmodel.HIDDEN, X)
    want = omlp.forward(cost.params, costmodel.N_FEATURES, costmodel.HIDDEN, X.astype(np.float64))[0]
    assert np.max(np.abs(got - want)) <= 1e-4 * max(1.0, float(np.max(np.abs(want))))


def _same_program(a, b):
    (ao, av), (bo, bv) = a.compile(), b.compile()
    assert len(a) == len(b)
    ra = np.frombuffer(ao, dtype=native.DMOP_DTYPE, count=len(a))
    rb = np.frombuffer(bo, dtype=native.DMOP_DTYPE, count=len(b))
    assert np.array_equal(ra, rb)
    assert av.shape == bv.shape
    if av.size:
        assert np.max(np.abs(av - bv)) <= 1e-14
    # the integer fields of fused records are exact
    for r in ra[ra['kind'] == native.OP_FUSED]:
        p = int(r['param'])
        m, nq = av[p], int(av[p + 1])
        assert m in (6.0, 7.0) and np.array_equal(av[p:p + 3 + nq], bv[p:p + 3 + nq])


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_batch_circuit_program_matches_python(precision):
    """The batching workload's circuits (widths 6-12, both forms) and one
    circuit with u1 / u2, per-gate depolarising, dephasing and a thermal
    feed."""
    specs = workloads.batch(40, 3, n_range=(6, 12))
    for n, gates, dr, _ in specs:
        for fused in (False, True):
            got = batch.native_program(n, gates, decay_factor=1.0 - dr, fused=fused, precision=precision)
            want = (batch.fused_circuit_program(n, gates, decay_factor=1.0 - dr, precision=precision) if fused
                    else batch.circuit_program(n, gates, decay_factor=1.0 - dr))
            _same_program(got, want)
    rng = np.random.default_rng(7)
    n = 8
    gates = []
    for i in range(120):
        r = rng.random()
        if r < 0.2:
            gates.append(('u1', (int(rng.integers(n)),), (float(rng.uniform(0, 6.3)),)))
        elif r < 0.6:
            gates.append(('u3', (int(rng.integers(n)),), tuple(float(x) for x in rng.uniform(0, 6.3, 3))))
        else:
            c, t = rng.choice(n, 2, replace=False)
            gates.append(('cx', (int(c), int(t)), ()))
    kw = dict(decay_factor=0.995, decoherence_factor=0.98, thermal_factor=0.9, gate_depolarization=1e-3)
    _same_program(batch.native_program(n, gates, fused=False, precision=precision, **kw),
                  batch.circuit_program(n, gates, **kw))
    _same_program(batch.native_program(n, gates, fused=True, precision=precision, **kw),
                  batch.fused_circuit_program(n, gates, precision=precision, **kw))
    # u2(phi, lam) is u3(pi/2, phi, lam); identity memory noise adds no channels
    g2 = [('u2', (1,), (0.3, 0.7)), ('cx', (0, 1), ()), ('u2', (0,), (1.1, -0.4))]
    g3 = [('u3', (1,), (np.pi / 2, 0.3, 0.7)), ('cx', (0, 1), ()), ('u3', (0,), (np.pi / 2, 1.1, -0.4))]
    _same_program(batch.native_program(6, g2, fused=True, precision=precision),
                  batch.fused_circuit_program(6, g3, precision=precision))


def test_batch_circuit_program_errors():
    lib = native.load()
    with pytest.raises(native.NativeError, match='n >= 6'):
        batch.native_program(5, [('cx', (0, 1), ())], fused=True)
    with pytest.raises(native.NativeError, match='bad qubits'):
        batch.native_program(6, [('cx', (2, 2), ())])
    with pytest.raises(native.NativeError, match='bad qubits'):
        batch.native_program(6, [('u3', (6,), (0.1, 0.2, 0.3))])
    with pytest.raises(ValueError):
        batch.native_program(6, [('rx', (0,), (0.1,))])
    # a short parameter list would shift every later gate's angles
    with pytest.raises(ValueError, match='3 parameter'):
        batch.native_program(6, [('u3', (0,), (0.1, 0.2)), ('u3', (1,), (0.1, 0.2, 0.3))])
    with pytest.raises(ValueError, match='2 qubit'):
        batch.native_program(6, [('cx', (0,), ())])
    # a kind outside u1 / u3 / cx, the size query, and too small an output
    kind, q = native._i32([native.INST_U3, 9]), native._i32([0, 0])
    ang = np.zeros(6)
    nops, nv = ctypes.c_int(0), ctypes.c_int64(0)
    rc = lib.dm_batch_circuit_program(6, 2, native._ip(kind), native._ip(q), native._ip(q), native._dp(ang), 1.0, 1.0,
                                      1.0, 0.0, 0, native.FP64, None, 0, None, 0, ctypes.byref(nops), ctypes.byref(nv))
    assert rc == native.ERR_ARG and 'not u1, u3 or cx' in native.last_error()
    kind = native._i32([native.INST_U3, native.INST_CX])
    q0, q1 = native._i32([0, 0]), native._i32([0, 1])
    rc = lib.dm_batch_circuit_program(6, 2, native._ip(kind), native._ip(q0), native._ip(q1), native._dp(ang), 0.99,
                                      1.0, 1.0, 0.0, 0, native.FP64, None, 0, None, 0, ctypes.byref(nops),
                                      ctypes.byref(nv))
    assert rc == 0 and (nops.value, nv.value) == (4, 12 + 3 + 0 + 3)
    ops = (native.DmOp * 3)()
    vals = np.zeros(18)
    rc = lib.dm_batch_circuit_program(6, 2, native._ip(kind), native._ip(q0), native._ip(q1), native._dp(ang), 0.99,
                                      1.0, 1.0, 0.0, 0, native.FP64, ctypes.cast(ops, ctypes.c_void_p), 3,
                                      native._dp(vals), 18, ctypes.byref(nops), ctypes.byref(nv))
    assert rc == native.ERR_ARG and 'need 4' in native.last_error()
    # a gate count whose records would overflow the op count is refused
    # before the gate arrays are read
    rc = lib.dm_batch_circuit_program(12, 200_000_000, native._ip(kind), native._ip(q0), native._ip(q1),
                                      native._dp(ang), 1.0, 1.0, 1.0, 0.0, 1, native.FP64, None, 0, None, 0,
                                      ctypes.byref(nops), ctypes.byref(nv))
    assert rc == native.ERR_ARG and 'too many' in native.last_error()


def test_concatenated_programs_keep_each_circuits_values():
    """The batch's op table: circuit c's records in [starts[c], starts[c+1]),
    each record's values found at its shifted offset."""
    specs = workloads.batch(12, 5, n_range=(4, 9))
    progs = [batch.native_program(n, gates, decay_factor=1.0 - dr, fused=n > 6) for n, gates, dr, _ in specs]
    progs.append(batch.Program(5))  # an empty program
    ops, vals, starts, nvals = batch.concat_programs(progs)
    assert starts[-1] == sum(len(p) for p in progs) and nvals == sum(p.compile()[1].size for p in progs)
    for c, p in enumerate(progs):
        arr, pv = p.compile()
        own = np.frombuffer(arr, dtype=native.DMOP_DTYPE, count=len(p))
        mine = ops[starts[c]:starts[c + 1]]
        assert np.array_equal(mine[['kind', 'k', 'qubits']], own[['kind', 'k', 'qubits']])
        ends = list(own['param'][1:]) + [pv.size]
        for r, o, e in zip(mine, own, ends):
            assert np.array_equal(vals[r['param']:r['param'] + (e - o['param'])], pv[o['param']:e])
    assert batch.concat_programs([])[3] == 0


def test_native_batch_programs_are_fast():
    """The 512-circuit batching workload's programs (HBM tier fused) build
    in well under a second (the Python builders take several seconds)."""
    specs = workloads.batch(512, 0, n_range=(6, 12))
    native.load()
    t0 = time.perf_counter()
    for n, gates, dr, _ in specs:
        batch.native_program(n, gates, decay_factor=1.0 - dr, fused=n > 7, precision='fp32')
    assert time.perf_counter() - t0 < 1.0


def test_native_front_end_is_fast():
    """The 16-qubit 1000-gate learned plan (with lookahead) and the greedy
    plans of the 512-circuit batching workload's HBM-tier circuits, each in
    well under a second."""
    cost = costmodel.CostModel.load(precision='fp32')
    gates = _gates(16, 1000, 0, 'local')
    t0 = time.perf_counter()
    fusion.plan(gates, 16, 4, 'learned', cost)
    t_learned = time.perf_counter() - t0
    specs = workloads.batch(512, 0, n_range=(6, 12))
    noisy = []
    for n, g, dr, _ in specs:
        mem = batch.ptm.memory_noise(1.0, 1.0, 1.0 - dr)
        seq = []
        for gate in g:
            seq += fusion.noisy_gates([gate])
            seq += [fusion.NoisyGate((q,), mem, 'u3') for q in range(n)]
        noisy.append((n, seq))
    t0 = time.perf_counter()
    for n, seq in noisy:
        fusion.plan(seq, n, 4, 'greedy', tiles=False)
    t_batch = time.perf_counter() - t0
    assert t_learned < 0.5, t_learned
    assert t_batch < 1.0, t_batch
