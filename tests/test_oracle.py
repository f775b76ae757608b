"""Pins the CPU oracle (oracle/refsim.py) before anything is checked against it.

* Each gate / channel update against the brute-force dense Kraus oracle.
* The ensemble contraction against the reference's literal Kronecker form.
* The whole driver against the user-guide notebooks' recorded outputs.
"""
import itertools
import json
import os
import types

import numpy as np
import pytest

from oracle import dense, refsim

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, 'golden', 'notebook_fixtures.json')))


def random_pauli_state(n, rng):
    """Random mixed state (a mixture of random pure states) in the Pauli basis."""
    rho = np.zeros((2 ** n, 2 ** n), dtype=complex)
    ws = rng.random(3)
    ws /= ws.sum()
    for w in ws:
        v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
        v /= np.linalg.norm(v)
        rho += w * np.outer(v, v.conj())
    return rho, dense.to_pauli(rho, n)


def as_state(a, n):
    return np.array(a, dtype=float).reshape(n * [4])


@pytest.mark.parametrize('n,q', [(1, 0), (3, 0), (3, 1), (3, 2)])
def test_u3_matches_dense(n, q):
    rng = np.random.default_rng(10 + q)
    rho, a = random_pauli_state(n, rng)
    th, ph, la = rng.uniform(0, 2 * np.pi, 3)
    st = as_state(a, n)
    st3 = st.reshape(4 ** q, 4, 4 ** (n - q - 1))
    for axis, ang in refsim.gate_rotations('u3', [th, ph, la]):
        st3 = refsim.rotate_components(st3, axis, ang, [1.0, 0.0])
    got = st3.reshape(-1)
    want = dense.to_pauli(dense.unitary(rho, dense.u3_matrix(th, ph, la), [q], n), n)
    assert np.max(np.abs(got - want)) < 1e-15


def test_rotation_error_is_averaged_rotation():
    """e0 = <cos d>, e1 = mean: the rotation averaged over d = +-acos(e0)."""
    n, q = 2, 1
    rng = np.random.default_rng(3)
    rho, a = random_pauli_state(n, rng)
    e0, e1, lam = 0.93, 0.07, 1.1
    st3 = as_state(a, n).reshape(4 ** q, 4, 4 ** (n - q - 1))
    got = refsim.rotate_components(st3, 'rz', lam, [e0, e1]).reshape(-1)
    d = np.arccos(e0)
    mix = 0.5 * dense.unitary(rho, dense.u3_matrix(0, 0, lam + e1 + d), [q], n) + \
        0.5 * dense.unitary(rho, dense.u3_matrix(0, 0, lam + e1 - d), [q], n)
    assert np.max(np.abs(got - dense.to_pauli(mix, n))) < 1e-15


@pytest.mark.parametrize('c,t', [(0, 1), (1, 0), (0, 2), (2, 0), (1, 2)])
def test_cx_matches_dense(c, t):
    n = 3
    rng = np.random.default_rng(c * 7 + t)
    rho, a = random_pauli_state(n, rng)
    got = refsim.apply_cx(as_state(a, n), c, t, [1.0, 0.0], n).reshape(-1)
    want = dense.to_pauli(dense.unitary(rho, dense.cx_matrix(), [c, t], n), n)
    assert np.max(np.abs(got - want)) < 1e-15


@pytest.mark.parametrize('c,t', [(0, 1), (2, 0)])
def test_tsp_cx_unit_cos_matches_documented_unitary(c, t):
    """With <cos a> = 1 the TSP CX is the PTM of U(a) = diag(1,1,[[i sin a, cos a],[cos a, i sin a]])."""
    n = 3
    rng = np.random.default_rng(5)
    rho, a = random_pauli_state(n, rng)
    mean = 0.2
    got = refsim.apply_cx(as_state(a, n), c, t, [1.0, mean], n).reshape(-1)
    want = dense.to_pauli(dense.unitary(rho, dense.tsp_cx_matrix(mean), [c, t], n), n)
    assert np.max(np.abs(got - want)) < 1e-15


def test_tsp_cx_is_linear_and_trace_preserving():
    """<cos> < 1 (moment substitution, formula-text pinned): linear, keeps a_I."""
    n = 3
    rng = np.random.default_rng(6)
    _, a = random_pauli_state(n, rng)
    _, b = random_pauli_state(n, rng)
    err = [0.97, 0.05]
    fa = refsim.apply_cx(as_state(a, n), 2, 1, err, n).reshape(-1)
    fb = refsim.apply_cx(as_state(b, n), 2, 1, err, n).reshape(-1)
    fab = refsim.apply_cx(as_state(0.3 * a + 0.7 * b, n), 2, 1, err, n).reshape(-1)
    assert np.max(np.abs(fab - (0.3 * fa + 0.7 * fb))) < 1e-16
    assert fa[0] == a[0]


@pytest.mark.parametrize('f,p,g', [(0.9, 0.0, 0.99), (0.95, 1.0, 0.97), (1.0, 0.3, 0.8), (0.7, 0.5, 1.0)])
def test_memory_noise_matches_kraus(f, p, g):
    n = 3
    rng = np.random.default_rng(int(f * 100))
    rho, a = random_pauli_state(n, rng)
    got = refsim.apply_memory_noise(as_state(a, n), n, f, p, g).reshape(-1)
    r = rho
    for q in range(n):
        r = dense.channel(r, dense.gad_kraus(g, p), [q], n)
        r = dense.channel(r, dense.dephase_kraus(f), [q], n)
    assert np.max(np.abs(got - dense.to_pauli(r, n))) < 1e-15


def test_gate_depolarizing_matches_kraus():
    n = 3
    rng = np.random.default_rng(8)
    rho, a = random_pauli_state(n, rng)
    got = refsim.apply_gate_depolarizing(as_state(a, n), n, [0, 2], 0.03).reshape(-1)
    r = dense.channel(rho, dense.depol_kraus(0.03), [0], n)
    r = dense.channel(r, dense.depol_kraus(0.03), [2], n)
    assert np.max(np.abs(got - dense.to_pauli(r, n))) < 1e-16


@pytest.mark.parametrize('q', [0, 1, 2])
def test_reset_matches_kraus(q):
    n = 3
    rng = np.random.default_rng(9 + q)
    rho, a = random_pauli_state(n, rng)
    got = refsim.reset_qubit(as_state(a, n), n, q).reshape(-1)
    assert np.max(np.abs(got - dense.to_pauli(dense.reset(rho, q, n), n))) < 1e-16


@pytest.mark.parametrize('axis,err', [(1, 1.0), (2, 0.9), (3, 0.95), (3, 1.0)])
def test_axis_collapse_matches_kraus(axis, err):
    n, q = 3, 1
    rng = np.random.default_rng(axis)
    rho, a = random_pauli_state(n, rng)
    got, p0 = refsim.collapse_axis(as_state(a, n), n, q, axis, err)
    r = dense.measure_collapse(rho, q, axis, err, n)
    assert np.max(np.abs(got.reshape(-1) - dense.to_pauli(r, n))) < 1e-16
    # P(0) of the (lossy) measurement, from the dense post-measurement state
    P = 0.5 * (dense.PAULI[0] + dense.PAULI[axis])
    p0_dense = np.real(np.trace(dense.embed(P, [q], n) @ r))
    assert abs(p0 - p0_dense) < 1e-14


def test_direction_collapse_matches_kraus():
    n, q = 2, 0
    rng = np.random.default_rng(21)
    rho, a = random_pauli_state(n, rng)
    nv = np.array([1.0, 2.0, 3.0]) / np.sqrt(14)
    got, _ = refsim.collapse_direction(as_state(a, n), n, q, nv, 0.9)
    r = dense.measure_direction(rho, q, nv, 0.9, n)
    assert np.max(np.abs(got.reshape(-1) - dense.to_pauli(r, n))) < 1e-15


@pytest.mark.parametrize('basis', ['X', 'Y', 'Z'])
@pytest.mark.parametrize('err', [1.0, 0.9])
def test_ensemble_matches_dense_and_kron(basis, err):
    n = 4
    rng = np.random.default_rng(ord(basis))
    rho, a = random_pauli_state(n, rng)
    got = refsim.ensemble_probabilities(as_state(a, n), n, basis, None, err)
    lit = refsim.ensemble_probabilities_kron(as_state(a, n), n, basis, None, err)
    den = dense.basis_probabilities(rho, n, basis, err)
    assert list(got) == list(den) == [''.join(s) for s in itertools.product('01', repeat=n)]
    for k in got:
        assert abs(got[k] - lit[k]) < 1e-15
        assert abs(got[k] - den[k]) < 1e-14


def test_ensemble_n_basis_matches_kron():
    n = 3
    rng = np.random.default_rng(2)
    _, a = random_pauli_state(n, rng)
    nv = np.array([0.3, -0.4, 0.5])
    got = refsim.ensemble_probabilities(as_state(a, n), n, 'N', nv, 0.8)
    lit = refsim.ensemble_probabilities_kron(as_state(a, n), n, 'N', nv, 0.8)
    for k in got:
        assert abs(got[k] - lit[k]) < 1e-15


def test_pauli_to_matrix_matches_dense():
    n = 3
    rng = np.random.default_rng(4)
    rho, a = random_pauli_state(n, rng)
    assert np.max(np.abs(refsim.pauli_to_matrix(a, n) - rho)) < 1e-15


def test_u3_merge_is_product_up_to_phase():
    """Merged U3 equals the product of the merged gates (semantic pin of mergeU)."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        g1 = [types.SimpleNamespace(name='u3', qubits=[0], params=list(rng.uniform(0, 2 * np.pi, 3))), 0]
        g2 = [types.SimpleNamespace(name='u3', qubits=[0], params=list(rng.uniform(0, 2 * np.pi, 3))), 1]
        m = refsim.merge_pair(g1, g2)[0]
        U = dense.u3_matrix(*g2[0].params) @ dense.u3_matrix(*g1[0].params)
        V = dense.u3_matrix(*m.params)
        ph = np.vdot(V.reshape(-1), U.reshape(-1))
        ph /= abs(ph)
        assert np.max(np.abs(U - ph * V)) < 1e-7  # arccos near 0 amplifies rounding


def _rec(d):
    r = types.SimpleNamespace(**{k: (list(v) if isinstance(v, list) else v) for k, v in d.items()})
    return r


def _levels_table(levels, count):
    out = []
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
        out.append(row)
    return out


@pytest.mark.parametrize('fx', FIX['partition_fixtures'], ids=lambda f: f['name'])
def test_partition_matches_notebook(fx):
    insts = [_rec(d) for d in fx['instructions']]
    merged = refsim.single_gate_merge(insts, fx['n_qubits'], True)
    levels, count = refsim.partition(merged, fx['n_qubits'])
    assert count == len(fx['levels'])
    assert _levels_table(levels, count) == fx['levels']


def _check_expect(data, exp):
    atol = exp['atol']
    if 'densitymatrix_real' in exp:
        dm = data['densitymatrix']
        assert np.max(np.abs(dm.real - np.array(exp['densitymatrix_real']))) <= atol
        assert np.max(np.abs(dm.imag - np.array(exp['densitymatrix_imag']))) <= atol
    for key in ('ensemble_probability', 'bell_probabilities01'):
        if key in exp:
            assert list(data[key]) == list(exp[key])
            for k, v in exp[key].items():
                assert abs(data[key][k] - v) <= atol
    if 'Pauli_string_expectation' in exp:
        assert abs(data['Pauli_string_expectation'] - exp['Pauli_string_expectation']) <= atol


@pytest.mark.parametrize('fx', FIX['fixtures'], ids=lambda f: f['name'])
def test_driver_matches_notebook(fx):
    insts = [_rec(d) for d in fx['instructions']]
    data, _, _ = refsim.run_experiment(insts, fx['n_qubits'], fx['options'])
    _check_expect(data, fx['expect'])


def test_partial_measure_marginals_and_collapse():
    """Partial Z measure of qubits {0, 2} of a 3-qubit state: marginals of the
    dense diagonal, then each measured qubit collapsed."""
    n = 3
    rng = np.random.default_rng(17)
    rho, a = random_pauli_state(n, rng)
    insts = [types.SimpleNamespace(name='measure', qubits=[0], memory=[0], params=['Z']),
             types.SimpleNamespace(name='measure', qubits=[2], memory=[1], params=['Z'])]
    opts = {'custom_densitymatrix': 'stored_density_matrix'}
    # feed the state via the stored-state path
    cwd = os.getcwd()
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        os.chdir(td)
        try:
            np.save('stored_density_matrix.npy', a)
            opts['initial_densitymatrix'] = 'x'
            data, count, _ = refsim.run_experiment(insts, n, opts)
        finally:
            os.chdir(cwd)
    diag = np.real(np.diag(rho)).reshape(2, 2, 2)
    marg = diag.sum(axis=1).reshape(-1)
    got = data['partial_probability']
    assert list(got) == ['00', '01', '10', '11']
    assert np.max(np.abs(np.array(list(got.values())) - marg)) < 1e-14
    r = dense.measure_collapse(rho, 0, 3, 1.0, n)
    r = dense.measure_collapse(r, 2, 3, 1.0, n)
    assert np.max(np.abs(data['coeffmatrix'] - dense.to_pauli(r, n))) < 1e-15
    assert count == 1
