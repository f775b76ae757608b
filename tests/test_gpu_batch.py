"""Batched execution (LDS-resident small circuits + HBM circuits in one
hipGraph) against the CPU oracle, circuit by circuit."""
import numpy as np
import pytest

from dmsim import batch, workloads
from oracle import refsim

pytestmark = pytest.mark.gpu


def oracle_probs(n, gates, decay_rate, depol_rate):
    qc = workloads.to_circuit(n, gates, barrier_after_each=True, measure=('ensemble', 'Z'))
    from dmsim import assemble
    exp = assemble(qc).experiments[0]
    data, _, _ = refsim.run_experiment(exp.instructions, n, {
        'decay_factor': 1.0 - decay_rate, 'depolarization_factor': 1.0 - depol_rate,
        'compute_densitymatrix': False})
    return np.array(list(data['ensemble_probability'].values()))


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-12), ('fp32', 2e-5)])
@pytest.mark.parametrize('graph', [True, False])
def test_batch_matches_oracle(precision, tol, graph):
    specs = workloads.batch(10, 77, n_range=(2, 9), depth_range=(10, 40))
    b = batch.from_workload(specs, precision, lds_max_qubits=7, lds_threads=256)
    for rep in range(2):  # graph capture then replay
        probs, us, ms = b.run(graph)
        for (n, gates, dr, pr), got in zip(specs, probs):
            want = oracle_probs(n, gates, dr, pr)
            assert np.max(np.abs(got - want)) <= tol, (n, rep)
        assert np.all(us > 0) and ms > 0
    b.close()


@pytest.mark.parametrize('lmax,threads', [(0, 256), (5, 512), (7, 1024)])
def test_batch_split_choices_agree(lmax, threads):
    specs = workloads.batch(6, 5, n_range=(3, 8), depth_range=(10, 30))
    ref = batch.from_workload(specs, 'fp64', lds_max_qubits=7, lds_threads=256)
    want, _, _ = ref.run(False)
    ref.close()
    b = batch.from_workload(specs, 'fp64', lds_max_qubits=lmax, lds_threads=threads)
    got, _, _ = b.run(True)
    b.close()
    for g, w in zip(got, want):
        assert np.max(np.abs(g - w)) <= 1e-13


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-12), ('fp32', 2e-5)])
def test_fused_programs_in_lds_and_hbm_tiers(precision, tol):
    """Fused tile programs (memory noise as per-qubit channels) as batch
    circuits: inlined into the LDS kernel for 6-7 qubits, tile passes in HBM
    for 8 qubits."""
    specs = workloads.batch(6, 300, n_range=(6, 8), depth_range=(20, 60))
    circuits = [(n, batch.fused_circuit_program(n, gates, decay_factor=1.0 - dr), 1.0 - pr)
                for n, gates, dr, pr in specs]
    b = batch.Batch(circuits, precision, lds_max_qubits=7, lds_threads=256)
    probs, us, ms = b.run(True)
    b.close()
    assert any(n <= 7 for n, _, _, _ in specs) and any(n == 8 for n, _, _, _ in specs)
    for (n, gates, dr, pr), got in zip(specs, probs):
        want = oracle_probs(n, gates, dr, pr)
        assert np.max(np.abs(got - want)) <= tol, n
