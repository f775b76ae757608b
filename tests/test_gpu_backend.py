"""End-to-end parity of the GPU backend with the CPU oracle.

Bar (FP64): max |coeff - oracle| <= 1e-12 and |trace - 1| <= 1e-14 on the
same circuits and options; probabilities and expectations to 1e-12.
FP32 storage: errors reported against the FP64 oracle with the bound stated
per test; the identity coefficient stays exactly 2^-n.
"""
import copy
import os

import numpy as np
import pytest

from dmsim import BasicAer, BasicAerError, QuantumCircuit, assemble, execute, workloads
from oracle import refsim

from test_frontend import FIX, build_fixture_circuit

pytestmark = pytest.mark.gpu

TOL = 1e-12


def backend():
    return BasicAer.get_backend('dm_simulator')


@pytest.fixture(autouse=True)
def fresh_backend_state():
    """The backend keeps some options across runs, as the reference's does
    (custom_densitymatrix, compute_densitymatrix, merge, store/compare flags,
    rotation_error); every test starts from the defaults."""
    be = backend()

    def reset():
        be._custom_densitymatrix = None
        be._get_den_mat = True
        for flag in ('MERGE', 'PLOTTING', 'SHOW_PARTITION', 'STORE_LOCAL', 'COMPARE', 'FILE_EXIST'):
            be.__dict__.pop(flag, None)
        be.DEFAULT_OPTIONS['rotation_error'].update({'rx': [1.0, 0.0], 'ry': [1.0, 0.0], 'rz': [1.0, 0.0]})
    reset()
    yield
    reset()


def run_both(qc, opts, precision='fp64', **backend_only):
    exp = assemble(qc).experiments[0]
    ref, count, _ = refsim.run_experiment(copy.deepcopy(exp.instructions), qc.num_qubits, dict(opts))
    res = execute(qc, backend(), precision=precision, **opts, **backend_only).result()
    out = res['results'][0]
    assert out['number_of_clock_cycles'] == count
    return ref, out['data'], res


def compare(ref, got, tol=TOL):
    assert np.max(np.abs(ref['coeffmatrix'] - got['coeffmatrix'])) <= tol
    for key in ('ensemble_probability', 'partial_probability'):
        if key in ref:
            assert list(ref[key]) == list(got[key])
            assert max(abs(ref[key][k] - got[key][k]) for k in ref[key]) <= tol
    if 'Pauli_string_expectation' in ref:
        assert abs(ref['Pauli_string_expectation'] - got['Pauli_string_expectation']) <= tol
    if 'densitymatrix' in ref:
        assert np.max(np.abs(ref['densitymatrix'] - got['densitymatrix'])) <= tol * 4


@pytest.mark.parametrize('f', FIX['fixtures'], ids=lambda f: f['name'])
def test_notebook_fixtures_on_gpu(f):
    qc = build_fixture_circuit(f)
    res = execute(qc, backend(), **f['options']).result()
    data = res['results'][0]['data']
    exp = f['expect']
    if 'densitymatrix_real' in exp:
        assert np.max(np.abs(data['densitymatrix'].real - np.array(exp['densitymatrix_real']))) <= exp['atol']
        assert np.max(np.abs(data['densitymatrix'].imag - np.array(exp['densitymatrix_imag']))) <= exp['atol']
    for key in ('ensemble_probability', 'bell_probabilities01'):
        if key in exp:
            assert list(data[key]) == list(exp[key])
            for k, v in exp[key].items():
                assert abs(data[key][k] - v) <= exp['atol']
    if 'Pauli_string_expectation' in exp:
        assert abs(data['Pauli_string_expectation'] - exp['Pauli_string_expectation']) <= exp['atol']


def test_cpu_reference_workload_10q_fp64():
    """10 qubits, 200 seeded U3/CX gates, barrier after each, decay 0.999,
    measurement depolarisation 0.999, ensemble Z measure."""
    for seed in (0, 1):
        gates = workloads.random_u3_cx(10, 200, seed)
        qc = workloads.to_circuit(10, gates, barrier_after_each=True, measure=('ensemble', 'Z'))
        opts = {'decay_factor': 0.999, 'depolarization_factor': 0.999, 'compute_densitymatrix': False}
        ref, got, _ = run_both(qc, opts)
        compare(ref, got)
        assert abs(got['coeffmatrix'][0] * 2 ** 10 - 1.0) <= 1e-14


def test_thermal_brickwork_qft_partial_fp64():
    """Brickwork + QFT with thermal relaxation after every gate, partial Z
    measure of qubits 0-3 (8 qubits so the oracle finishes in seconds)."""
    n = 8
    g, f = workloads.thermal_factors(50e-6, 70e-6, 50e-9)
    gates = workloads.brickwork_qft(n, 4, 0)
    qc = workloads.to_circuit(n, gates, barrier_after_each=True, measure=('partial', [0, 1, 2, 3], 'Z'))
    opts = {'decay_factor': g, 'decoherence_factor': f, 'thermal_factor': 1.0}
    ref, got, _ = run_both(qc, opts)
    compare(ref, got)


@pytest.mark.parametrize('seed', [3, 4, 5])
def test_random_circuits_all_options_fp64(seed):
    from test_frontend import random_circuit
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 7))
    qc = random_circuit(rng, n, 80)
    qc.measure(0, 0, basis='Ensemble', add_param='Y')
    opts = {'decay_factor': 0.98, 'decoherence_factor': 0.97, 'thermal_factor': 0.3,
            'depolarization_factor': 0.95, 'rotation_error': {'rx': [1.0, 0.0], 'ry': [0.99, 0.01],
                                                              'rz': [0.995, -0.02]},
            'tsp_model_error': [0.98, 0.03], 'gate_depolarization': 0.002}
    ref, got, _ = run_both(qc, opts)
    compare(ref, got)


@pytest.mark.parametrize('fusion', ['greedy', 'learned'])
@pytest.mark.parametrize('seed,barriers', [(11, False), (12, True)])
def test_fused_execution_matches_oracle_fp64(fusion, seed, barriers):
    """Fused execution (tile passes, memory noise as per-qubit channels)
    against the oracle's level-by-level run: every noise option, TSP CX
    (dense two-qubit operations), with and without barriers."""
    n = 8
    gates = workloads.local_u3_cx(n, 120, seed)
    qc = workloads.to_circuit(n, gates, barrier_after_each=barriers, measure=('ensemble', 'Z'))
    opts = {'decay_factor': 0.98, 'decoherence_factor': 0.97, 'thermal_factor': 0.3,
            'rotation_error': {'rx': [1.0, 0.0], 'ry': [0.99, 0.01], 'rz': [0.995, -0.02]},
            'tsp_model_error': [0.98, 0.03], 'gate_depolarization': 0.002}
    ref, got, _ = run_both(qc, opts, fusion=fusion)
    compare(ref, got)


def test_measure_modes_fp64():
    n = 4
    for mode in ('partial_X', 'partial_N', 'expect', 'bell', 'single_mixed'):
        qc = QuantumCircuit(n, n)
        qc.h(0)
        qc.cx(0, 1)
        qc.u3(0.3, 0.2, 0.1, 2)
        qc.cx(2, 3)
        if mode == 'partial_X':
            qc.measure(1, 1, basis='X')
            qc.measure(3, 3, basis='X')
        elif mode == 'partial_N':
            qc.measure(0, 0, basis='N', add_param=np.array([0.6, 0.0, 0.8]))
            qc.measure(2, 2, basis='N', add_param=np.array([0.6, 0.0, 0.8]))
        elif mode == 'expect':
            qc.measure(0, 0, basis='Expect', add_param='ZZXI')
        elif mode == 'bell':
            qc.measure(0, 0, basis='Bell', add_param='02')
        else:
            qc.measure(0, 0, basis='Y')
            qc.measure(1, 1, basis='X')
            qc.measure(2, 2, basis='Z')
        ref, got, _ = run_both(qc, {'depolarization_factor': 0.9})
        compare(ref, got)
        if mode == 'bell':
            b = got['bell_probabilities02']
            r = ref['bell_probabilities02']
            assert max(abs(b[k] - r[k]) for k in r) <= TOL
            assert np.max(np.abs(got['reduced_bell_densitymatrix02'] - ref['reduced_bell_densitymatrix02'])) <= TOL


@pytest.mark.parametrize('n,pair', [(4, '13'), (4, '31'), (5, '04'), (5, '32')])
def test_bell_measure_asymmetric_states_fp64(n, pair):
    """Bell measurement after a random noisy circuit (no symmetry between
    the qubits), both pair orders: probabilities, reduced matrix and the
    collapsed state against the oracle, whose Bell function is pinned by
    dense Bell projectors on the mirrored pair."""
    rng = np.random.default_rng(int(pair) + n)
    qc = QuantumCircuit(n, n)
    for _ in range(12):
        if rng.random() < 0.6:
            qc.u3(*rng.uniform(0, 2 * np.pi, 3), int(rng.integers(n)))
        else:
            c, t = (int(x) for x in rng.choice(n, 2, replace=False))
            qc.cx(c, t)
    qc.measure(0, 0, basis='Bell', add_param=pair)
    ref, got, _ = run_both(qc, {'decay_factor': 0.99, 'decoherence_factor': 0.98, 'thermal_factor': 0.9})
    compare(ref, got)
    key = 'bell_probabilities' + pair
    assert max(abs(got[key][k] - ref[key][k]) for k in ref[key]) <= TOL
    red = 'reduced_bell_densitymatrix' + pair
    assert np.max(np.abs(got[red] - ref[red])) <= TOL


def test_initial_states_fp64(tmp_path):
    n = 3
    for opts in ({'custom_densitymatrix': 'max_mixed'}, {'custom_densitymatrix': 'uniform_superpos'},
                 {'custom_densitymatrix': 'thermal_state', 'thermal_factor': 0.8},
                 {'custom_densitymatrix': 'binary_string', 'initial_densitymatrix': '011'}):
        qc = QuantumCircuit(n)
        qc.h(1)
        qc.cx(1, 2)
        ref, got, _ = run_both(qc, opts)
        compare(ref, got)


def test_store_and_compare_roundtrip(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    qc = QuantumCircuit(3, 3)
    qc.h(0)
    qc.cx(0, 1)
    qc.measure(0, 0, basis='Ensemble', add_param='Z')
    be = backend()
    res = execute(qc, be, store_densitymatrix=True).result()
    stored = np.load('stored_coefficients.npy')
    assert np.max(np.abs(stored - res['results'][0]['data']['coeffmatrix'])) <= 1e-15
    res2 = execute(qc, be, compare=True, store_densitymatrix=False).result()
    assert abs(res2['results'][0]['data']['fidelity'] - 2 ** 3 * np.dot(stored, stored)) <= 1e-12


@pytest.mark.parametrize('n', [8, 10])
def test_fp32_storage_against_fp64_oracle(n):
    gates = workloads.random_u3_cx(n, 150, 7)
    qc = workloads.to_circuit(n, gates, barrier_after_each=True, measure=('ensemble', 'Z'))
    opts = {'decay_factor': 0.999, 'depolarization_factor': 0.999, 'compute_densitymatrix': False}
    ref, got, _ = run_both(qc, opts, precision='fp32')
    err = np.max(np.abs(ref['coeffmatrix'] - got['coeffmatrix']))
    assert err * 2 ** n <= 2e-4          # expectation-scale error of FP32 storage
    assert got['coeffmatrix'][0] == 2.0 ** -n
    pe = max(abs(ref['ensemble_probability'][k] - got['ensemble_probability'][k]) for k in ref['ensemble_probability'])
    assert pe <= 1e-5


@pytest.mark.parametrize('n', [4, 9])
def test_fused_fp32_execution_against_fp64_oracle(n):
    """Learned fusion with FP32 storage through execute(): small registers
    (n < 6: blocks of at most 2 qubits) and tile passes (n = 9)."""
    gates = workloads.local_u3_cx(n, 150, 21)
    qc = workloads.to_circuit(n, gates, barrier_after_each=True, measure=('ensemble', 'Z'))
    opts = {'decay_factor': 0.995, 'decoherence_factor': 0.99, 'thermal_factor': 0.8,
            'depolarization_factor': 0.999, 'gate_depolarization': 0.001, 'compute_densitymatrix': False}
    ref, got, _ = run_both(qc, opts, precision='fp32', fusion='learned')
    err = np.max(np.abs(ref['coeffmatrix'] - got['coeffmatrix']))
    assert err * 2 ** n <= 2e-4
    assert got['coeffmatrix'][0] == 2.0 ** -n
    pe = max(abs(ref['ensemble_probability'][k] - got['ensemble_probability'][k]) for k in ref['ensemble_probability'])
    assert pe <= 1e-5


def test_sticky_options_mirror_reference():
    """compute_densitymatrix and custom_densitymatrix persist on the backend
    between runs, as on the reference's cached backend instance."""
    qc = QuantumCircuit(2)
    qc.h(0)
    be = backend()
    r1 = execute(qc, be, compute_densitymatrix=False, custom_densitymatrix='max_mixed').result()
    r2 = execute(qc, be).result()
    assert 'densitymatrix' not in r2['results'][0]['data']
    assert np.allclose(r1['results'][0]['data']['coeffmatrix'], r2['results'][0]['data']['coeffmatrix'])


def test_qubit_count_limit_and_bad_ops():
    be = backend()
    qc = QuantumCircuit(be.configuration().n_qubits + 1)
    with pytest.raises(BasicAerError):
        execute(qc, be).result()
