"""Batched execution (LDS-resident small circuits + HBM circuits as up to
three launches per round of passes, one hipGraph) against the CPU oracle,
circuit by circuit: FP64 to 1e-12, FP32 to 2e-5 (probabilities)."""
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


@pytest.mark.parametrize('lmax,threads,ipp', [(5, 256, 1), (6, 512, 64), (7, 1024, 1024)])
def test_batch_split_choices_agree(lmax, threads, ipp):
    specs = workloads.batch(8, 5, n_range=(3, 10), depth_range=(10, 30))
    ref = batch.from_workload(specs, 'fp64', lds_max_qubits=7, lds_threads=256)
    want, _, _ = ref.run(False)
    ref.close()
    b = batch.from_workload(specs, 'fp64', lds_max_qubits=lmax, lds_threads=threads, items_per_pass=ipp)
    got, _, _ = b.run(True)
    b.close()
    for g, w in zip(got, want):
        assert np.max(np.abs(g - w)) <= 1e-13


def pass_tiles(prog):
    """Tile digits of each pass of a fused circuit program (single-gate
    operations run as 6-digit passes)."""
    from dmsim import native
    arr, params = prog.compile()
    return [int(params[arr[i].param]) if arr[i].kind == native.OP_FUSED else 6 for i in range(len(prog))]


def round_launches(progs):
    """Launches of one batch run: round r holds INIT (r = 0), pass r - 1 or
    PROBS of every circuit; one launch for its INIT and PROBS items, one for
    its 6-digit tile items and one for its 7-digit ones."""
    tiles = [pass_tiles(p) for p in progs]
    rounds = max(len(t) for t in tiles) + 2
    n = 0
    for r in range(rounds):
        ends = six = seven = False
        for t in tiles:
            if r == 0 or r == len(t) + 1:
                ends = True
            elif r <= len(t):
                six, seven = six or t[r - 1] == 6, seven or t[r - 1] == 7
        n += ends + six + seven
    return n


def test_batching_workload_reduced_against_oracle():
    """The batching workload's generator (widths 6-10 here, depths 20-60,
    per-circuit decay and measurement rates) with every HBM-tier circuit in
    the round kernels, graph replayed three times, FP64 and FP32 (whose
    passes over high qubits take 7-digit tiles)."""
    specs = workloads.batch(24, 3000, n_range=(6, 10), depth_range=(20, 60))
    wants = [oracle_probs(n, gates, dr, pr) for n, gates, dr, pr in specs]
    for precision, tol in (('fp64', 1e-12), ('fp32', 2e-5)):
        b = batch.from_workload(specs, precision, lds_max_qubits=5, lds_threads=256, items_per_pass=64)
        info = b.info()
        assert info['hbm_circuits'] == 24
        progs = [batch.fused_circuit_program(n, gates, decay_factor=1.0 - dr, precision=precision)
                 for n, gates, dr, _ in specs]
        sevens = sum(t == 7 for p in progs for t in pass_tiles(p))
        assert (sevens > 0) == (precision == 'fp32')
        assert info['launches'] == round_launches(progs)
        for rep in range(3):
            probs, us, ms = b.run(True)
            for (n, _, _, _), got, want in zip(specs, probs, wants):
                assert np.max(np.abs(got - want)) <= tol, (precision, n, rep)
                assert abs(got.sum() - 1.0) <= (1e-12 if precision == 'fp64' else 1e-5)
            assert np.all(us > 0) and ms > 0
        b.close()


def test_plain_programs_with_memory_noise_in_the_hbm_tier():
    """Circuits built one op per gate plus a memory-noise op per level (the
    HBM tier turns each into 6-digit tile passes) against the oracle."""
    specs = workloads.batch(5, 41, n_range=(8, 11), depth_range=(10, 25))
    circuits = [(n, batch.circuit_program(n, gates, decay_factor=1.0 - dr, decoherence_factor=0.99,
                                          thermal_factor=0.8), 1.0 - pr) for n, gates, dr, pr in specs]
    b = batch.Batch(circuits, 'fp64', lds_max_qubits=7, lds_threads=256, items_per_pass=16)
    probs, _, _ = b.run(True)
    b.close()
    for (n, gates, dr, pr), got in zip(specs, probs):
        qc = workloads.to_circuit(n, gates, barrier_after_each=True, measure=('ensemble', 'Z'))
        from dmsim import assemble
        exp = assemble(qc).experiments[0]
        data, _, _ = refsim.run_experiment(exp.instructions, n, {
            'decay_factor': 1.0 - dr, 'decoherence_factor': 0.99, 'thermal_factor': 0.8,
            'depolarization_factor': 1.0 - pr, 'compute_densitymatrix': False})
        want = np.array(list(data['ensemble_probability'].values()))
        assert np.max(np.abs(got - want)) <= 1e-12, n


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
