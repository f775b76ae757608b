"""Host front end (CPU only): transpile order, merge/partition, options, ABI exports."""
import copy
import ctypes
import json
import os
import re
import types

import numpy as np
import pytest

from dmsim import (BasicAer, BasicAerError, ClassicalRegister, QuantumCircuit, QuantumRegister, assemble,
                   schedule, transpiler)
from dmsim import native
from oracle import refsim

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIX = json.load(open(os.path.join(HERE, 'golden', 'notebook_fixtures.json')))


def build_fixture_circuit(f):
    qr = QuantumRegister(f['n_qubits'], 'q')
    cr = ClassicalRegister(f['n_clbits'], 'c') if f['n_clbits'] else None
    qc = QuantumCircuit(qr, cr) if cr else QuantumCircuit(qr)
    for step in f['circuit']:
        name = step[0]
        if name == 'measure':
            kw = step[3] if len(step) > 3 else {}
            qq = qr if step[1] == 'all' else step[1][0]
            cc = cr if step[2] == 'all' else step[2][0]
            qc.measure(qq, cc, **kw)
        else:
            prm = step[2] if len(step) > 2 else []
            getattr(qc, name)(*prm, *step[1])
    return qc


def _rec_view(r):
    d = {'name': r.name}
    if hasattr(r, 'qubits'):
        d['qubits'] = list(r.qubits)
    if hasattr(r, 'params'):
        d['params'] = [x if isinstance(x, str) else (list(x) if isinstance(x, (list, np.ndarray)) else float(x))
                       for x in r.params]
    if hasattr(r, 'memory'):
        d['memory'] = list(r.memory)
    return d


@pytest.mark.parametrize('f', FIX['fixtures'] + FIX['partition_fixtures'], ids=lambda f: f['name'])
def test_transpile_matches_fixture(f):
    exp = assemble(build_fixture_circuit(f)).experiments[0]
    got = [_rec_view(r) for r in exp.instructions]
    assert len(got) == len(f['instructions'])
    for g, w in zip(got, f['instructions']):
        assert g['name'] == w['name']
        assert g.get('qubits') == w.get('qubits')
        if 'params' in w:
            assert g['params'] == w['params']


@pytest.mark.parametrize('f', FIX['partition_fixtures'], ids=lambda f: f['name'])
def test_product_partition_matches_notebook(f):
    exp = assemble(build_fixture_circuit(f)).experiments[0]
    insts = copy.deepcopy(exp.instructions)
    merged = schedule.merge_single_gates(insts, f['n_qubits'])
    levels, count = schedule.partition(merged, f['n_qubits'])
    table = []
    for lv in levels[:count]:
        row = []
        for op in lv:
            prm = getattr(op, 'params', None)
            if op.name in ('u1', 'u3'):
                prm = [round(float(x), 6) for x in prm]
            elif op.name == 'measure':
                prm = [str(x) for x in (prm or ['Z'])]
            else:
                prm = []
            row.append([op.name, list(op.qubits), prm])
        table.append(row)
    assert table == f['levels']


def random_circuit(rng, n, ngates, with_meas=True):
    qc = QuantumCircuit(n, n)
    for _ in range(ngates):
        r = rng.random()
        if r < 0.35:
            qc.u3(*rng.uniform(0, 2 * np.pi, 3), int(rng.integers(n)))
        elif r < 0.45:
            qc.u1(rng.uniform(0, 2 * np.pi), int(rng.integers(n)))
        elif r < 0.5:
            qc.u2(*rng.uniform(0, 2 * np.pi, 2), int(rng.integers(n)))
        elif r < 0.55:
            getattr(qc, ['h', 's', 't', 'x', 'y', 'z', 'sdg', 'tdg'][int(rng.integers(8))])(int(rng.integers(n)))
        elif r < 0.8:
            a, b = rng.choice(n, 2, replace=False)
            qc.cx(int(a), int(b))
        elif r < 0.85:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            kind = ['cz', 'cu1', 'swap', 'crz', 'cy'][int(rng.integers(5))]
            if kind in ('cu1', 'crz'):
                getattr(qc, kind)(rng.uniform(0, 2 * np.pi), a, b)
            else:
                getattr(qc, kind)(a, b)
        elif r < 0.9:
            qc.barrier()
        elif r < 0.95 and with_meas:
            q = int(rng.integers(n))
            qc.measure(q, q, basis=['X', 'Y', 'Z'][int(rng.integers(3))])
        elif with_meas:
            qc.reset(int(rng.integers(n)))
    return qc


def test_lexicographic_order_matches_networkx():
    nx = pytest.importorskip('networkx')
    rng = np.random.default_rng(1)
    for trial in range(30):
        n = int(rng.integers(2, 6))
        qc = random_circuit(rng, n, 40)
        ops = transpiler.unroll(qc.data)
        ops = transpiler.remove_reset_in_zero_state(qc, ops)
        mine = transpiler.lexicographic_order(qc, ops)
        # independent: networkx's own lexicographical topological sort
        G = nx.MultiDiGraph()
        last = {}
        order_id = {}
        nid = 0
        for q in qc.qubits:
            G.add_node(nid, key='[]', inst=None)
            last[('q', q)] = nid
            nid += 2
        for c in qc.clbits:
            G.add_node(nid, key='[]', inst=None)
            last[('c', c)] = nid
            nid += 2
        for inst in ops:
            key = '[' + ', '.join(repr(q) for q in inst.qargs) + ']'
            G.add_node(nid, key=key, inst=inst)
            for w in [('q', q) for q in inst.qargs] + [('c', c) for c in inst.cargs]:
                G.add_edge(last[w], nid)
                last[w] = nid
            nid += 1
        ref = [G.nodes[x]['inst'] for x in nx.lexicographical_topological_sort(G, key=lambda x: (G.nodes[x]['key'], x))
               if G.nodes[x]['inst'] is not None]
        assert [id(x) for x in mine] == [id(x) for x in ref], trial


def assert_same_merge(a, b):
    """Same records and names; merged angles to 1e-14 (the native merge uses
    the C library's atan2 / acos, numpy its own: a few ulp apart)."""
    assert [(x.name, x.qubits) for x in a] == [(x.name, x.qubits) for x in b]
    for x, y in zip(a, b):
        if x.name in ('u1', 'u3'):
            px = np.array([float(p) for p in x.params])
            py = np.array([float(p) for p in y.params])
            assert px.shape == py.shape and np.max(np.abs(px - py)) <= 1e-14


def test_product_schedule_matches_oracle_on_random_circuits():
    rng = np.random.default_rng(2)
    for trial in range(40):
        n = int(rng.integers(2, 7))
        qc = random_circuit(rng, n, 60)
        exp = assemble(qc).experiments[0]
        a = schedule.merge_single_gates(copy.deepcopy(exp.instructions), n)
        b = refsim.single_gate_merge(copy.deepcopy(exp.instructions), n)
        assert_same_merge(a, b)
        la, ca = schedule.partition(a, n)
        lb, cb = refsim.partition(b, n)
        assert ca == cb, trial
        for x, y in zip(la[:ca], lb[:cb]):
            assert [(o.name, o.qubits) for o in x] == [(o.name, o.qubits) for o in y]


def test_remove_reset_in_zero_state():
    qc = QuantumCircuit(2)
    qc.reset(0)
    qc.reset(0)
    qc.h(1)
    qc.reset(1)
    names = [(o.name, o.qubits) for o in assemble(qc).experiments[0].instructions]
    assert ('reset', [1]) in names
    assert ('reset', [0]) not in names


def test_options_errors_mirror_reference():
    be = BasicAer.get_backend('dm_simulator')
    with pytest.raises(BasicAerError, match='Rotation Error'):
        be._set_options(backend_options={'rotation_error': {'rq': [1, 0]}})
    with pytest.raises(BasicAerError, match='transition model'):
        be._set_options(backend_options={'tsp_model_error': [1.0]})
    with pytest.raises(BasicAerError):
        be._set_options(backend_options={'precision': 'fp16'})
    be._set_options(backend_options={})
    assert BasicAer.get_backend('dm_simulator') is be
    assert BasicAer.get_backend('dm_simulator_py') is be
    with pytest.raises(BasicAerError):
        BasicAer.get_backend('qasm_simulator')


def test_measure_api_errors():
    qc = QuantumCircuit(2, 2)
    with pytest.raises(BasicAerError):
        qc.measure(0, 0, basis='N')
    with pytest.raises(BasicAerError):
        qc.measure(0, 0, basis='Bell')
    with pytest.raises(BasicAerError):
        qc.measure(0, 0, basis='Q')


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, 'include', 'dmsim.h')).read()
    declared = set(re.findall(r'^int\s+(dm_\w+)\s*\(', header, flags=re.M))
    assert declared == set(native.SIGNATURES)
    lib = ctypes.CDLL(native.library_path())
    for name in declared:
        assert hasattr(lib, name), name


def test_fused_tile_choice():
    """Sequential passes take 7-digit tiles where 6 digits would leave
    contiguous HBM runs of 4^2 coefficients, and in FP32 where 7 digits turn
    runs of 4^3 into 4^4; FP64 keeps 6 digits there."""
    from dmsim import fusion
    n = 16
    assert fusion.run_digits(n, [15, 14, 13, 12], 6) == 2
    assert fusion.run_digits(n, [15, 14, 13, 12], 7) == 3
    assert fusion.run_digits(n, [0, 7], 6) == 5
    assert fusion.tile_for(n, [0, 1, 2, 3]) == 7      # positions 12..15
    assert fusion.tile_for(n, [0, 1, 2]) == 7         # FP32: w 3 -> 4
    assert fusion.tile_for(n, [0, 1, 2], 'fp64') == 6  # FP64: m=6 already gives w=3
    assert fusion.tile_for(n, [0, 1, 12]) == 6        # position 3 in the block: m=6 gives w=4
    assert fusion.tile_for(n, [15, 14]) == 6          # block at the lowest positions
    assert fusion.tile_for(6, [0, 1, 2, 3]) == 6      # register too small for 7
    assert fusion.tile_qubits(n, {0, 1, 2}) == {0, 1, 2, 12, 13, 14, 15}
    assert fusion.tile_qubits(n, {0, 1, 2}, 'fp64') == {0, 1, 2, 13, 14, 15}


def test_fp64_short_run_blocks_take_7_digits_at_any_stage_count():
    """An FP64 block whose 6-digit tile would leave runs of 4^2 takes the
    7-digit tile also when its operations need many register stages (a
    6-digit rule for 3+ stages was measured slower with the current FP64
    7-digit kernels and removed)."""
    from dmsim import fusion, native
    from dmsim.program import Program
    n = 16
    qs = [0, 1, 2, 3]                                  # positions 15..12: w6 = 2
    gates = [fusion.NoisyGate(pair, None, 'cx') for pair in ((0, 1), (1, 2), (2, 3), (0, 3), (1, 3), (0, 2))]
    for prec in ('fp64', 'fp32'):
        prog = Program(n)
        fusion.add_block(prog, gates, qs, list(range(len(gates))), 'seq', None, prec)
        arr, vals = prog.compile()
        assert len(prog) == 1 and arr[0].kind == native.OP_FUSED
        assert vals[arr[0].param] == 7.0, prec


def test_shipped_cost_model_file():
    """The trained parameters load without pickle and carry their training
    report (held-out fit of the log-time target)."""
    from dmsim import costmodel
    m = costmodel.CostModel.load()
    P = costmodel.N_FEATURES * costmodel.HIDDEN + costmodel.HIDDEN + costmodel.HIDDEN ** 2 + 2 * costmodel.HIDDEN + 1
    assert m.params.shape == (P,) and np.isfinite(m.params).all()
    assert m.meta['features'] == costmodel.N_FEATURES and m.meta['hidden'] == costmodel.HIDDEN
    assert m.meta['holdout_r2'] > 0.9
