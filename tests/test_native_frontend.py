"""The native host front end (merge, partition, fusion planner, cost model)
against the same passes in Python, on the CPU.

* merge: same records, names and qubits; merged angles to 1e-14 (the C
  library's atan2 / acos against numpy's, a few ulp apart);
* partition: identical level lists and level count;
* planner: identical blocks (members and qubit order) for greedy, search
  and the learned planner with the shipped cost model (both evaluate the
  network with the host forward pass), with and without lookahead;
* cost model: the native block seconds equal CostModel.block_times.
"""
import copy
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
