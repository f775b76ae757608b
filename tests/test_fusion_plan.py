"""Fusion planner (CPU only): every plan applies each gate once, keeps the
order of gates on every qubit, and fits its tiles; the tile-wide candidates
are valid blocks.  The learned planner's network runs through the numpy
forward pass of the oracle here (the GPU forward pass is tested against it
in test_gpu_costmodel.py)."""
import numpy as np
import pytest

from dmsim import costmodel, fusion, native, workloads
from oracle import mlp


@pytest.fixture
def cpu_model(monkeypatch):
    monkeypatch.setattr(native, 'mlp_predict', lambda p, F, H, X: mlp.forward(
        p.astype(np.float64), F, H, X.astype(np.float64))[0])
    m = costmodel.CostModel.load()
    m.precision = 'fp32'
    return m


def check_block(gates, done, members):
    """Members run together at the first pending gate: every pending gate
    left out before a member shares no qubit with it."""
    mem = set(members)
    for i in members:
        for j in range(i):
            if not done[j] and j not in mem:
                assert not set(gates[j].qubits) & set(gates[i].qubits), (j, i)


def check_plan(gates, n, blocks):
    order = [i for _, mem in blocks for i in mem]
    assert sorted(order) == list(range(len(gates)))
    done = [False] * len(gates)
    for qs, mem in blocks:
        assert set(qs) == {q for i in mem for q in gates[i].qubits}
        check_block(gates, done, mem)
        if len(mem) > 1 and len(qs) > 1 and n >= fusion.MIN_TILE_QUBITS:
            assert len(qs) <= fusion.tile_for(n, qs) <= max(n, 6)
        for i in mem:
            done[i] = True


@pytest.mark.parametrize('n', [6, 7, 9, 16])
@pytest.mark.parametrize('method', ['greedy', 'search', 'learned'])
def test_plans_are_valid(n, method, cpu_model):
    gates = fusion.noisy_gates(workloads.local_u3_cx(n, 300, n), 1e-3)
    blocks = fusion.plan(gates, n, 4, method, cpu_model if method == 'learned' else None)
    check_plan(gates, n, blocks)


@pytest.mark.parametrize('lookahead', [(), (3,), (3, 2)])
def test_learned_plans_with_and_without_lookahead_are_valid(lookahead, cpu_model):
    n = 12
    gates = fusion.noisy_gates(workloads.local_u3_cx(n, 300, 11), 1e-3)
    check_plan(gates, n, fusion.plan(gates, n, 4, 'learned', cpu_model, lookahead=lookahead))


def test_tile_wide_candidates():
    """With tiles the candidates include the classic ones; every candidate
    is a valid block of at most 7 qubits, the three lowest positions free."""
    n = 16
    gates = fusion.noisy_gates(workloads.local_u3_cx(n, 300, 3), 1e-3)
    for start in (1, 2, 50, 100):
        # start is the first pending gate; some later ones are done already
        done = [i < start or i % 7 == 0 for i in range(300)]
        classic = fusion.candidate_blocks(gates, done, start, 4, n)
        wide = fusion.candidate_blocks(gates, done, start, 4, n, tiles=True)
        assert {tuple(m) for _, m in classic} <= {tuple(m) for _, m in wide}
        low = fusion.low_qubits(n)
        for qs, mem in wide:
            assert start in mem
            check_block(gates, done, mem)
            assert len(qs) <= 7 and len(set(qs) - low) <= 4
    assert fusion.tile_qubits(n, {0, 1, 2, 3}) == {0, 1, 2, 3, 13, 14, 15}
    assert fusion.tile_qubits(n, {0}) == {0, 12, 13, 14, 15, 11}
    assert fusion.tile_for(n, [0, 1, 2, 3, 13, 14, 15]) == 7


def test_learned_plan_uses_tile_wide_blocks(cpu_model):
    """On the 16-qubit workload the learned planner takes blocks of more than
    4 qubits and needs fewer passes than the classic candidates give it."""
    n = 16
    gates = fusion.noisy_gates(workloads.local_u3_cx(n, 400, 5000), 1e-3)
    wide = fusion.plan(gates, n, 4, 'learned', cpu_model, tiles=True)
    classic = fusion.plan(gates, n, 4, 'learned', cpu_model, tiles=False)
    assert max(len(qs) for qs, _ in wide) > 4
    assert len(wide) < len(classic)
