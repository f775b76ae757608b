"""The benchmark configurations on the GPU: against the oracle at reduced
size, and at full size through properties the oracle cannot afford.

* Precision form (17 qubits, layers of U3 on every qubit + CX on a random
  perfect matching, depolarising 5e-4 per gate): the same workload at 9
  qubits against the oracle in FP64 (max |da| <= 1e-12, |trace - 1| <=
  1e-14) and FP32 (2^n max |da| <= 2e-4, identity coefficient exact); at 17
  qubits, FP32 and FP64 states side by side (206 GB) for
  - a circuit on 4 qubits spread over the register (positions 0, 5, 11, 16):
    the state must be the oracle's 4-qubit state times |0><0| on every idle
    qubit, checked coefficient by coefficient on the 4^4 x {I,Z} lattice
    and over the whole vector through the purity (any stray coefficient
    anywhere adds to the sum of squares) and a zero X/Y sample;
  - two full-width layers: trace 1, identity coefficient exact, marginals in
    [0, 1], FP32 against FP64 within the FP32 storage bound.
* Plain FP64 brickwork + QFT with thermal relaxation (14 qubits, barrier
  after every gate, partial Z on qubits 0-3) through execute(): the plain
  per-gate path and both fused paths agree to 1e-12 at full size (two
  independent kernel paths), and the same circuit family on 4 qubits
  embedded in 14 must equal the oracle's 4-qubit run (idle qubits stay
  |0>: thermal relaxation with thermal_factor 1 keeps the ground state).
* Learned fusion (16 qubits, 1000 gates, depolarising 1e-3): fused FP64
  against plain FP64 to 1e-12 at full size, fused FP32 within 2^n max |da|
  <= 1e-5 of plain FP64, identity coefficient exact.
"""
import copy

import numpy as np
import pytest

from dmsim import BasicAer, assemble, costmodel, execute, fusion, native, workloads
from oracle import refsim

pytestmark = pytest.mark.gpu


def oracle_noisy(n, gates, p):
    """The oracle applied gate by gate, each gate followed by the
    depolarising channel p on its qubits."""
    st = refsim.initial_state(n)
    for name, qs, prm in gates:
        if name == 'u3':
            s3 = st.reshape(4 ** qs[0], 4, 4 ** (n - qs[0] - 1))
            for axis, ang in refsim.gate_rotations('u3', prm):
                s3 = refsim.rotate_components(s3, axis, ang, [1.0, 0.0])
            st = s3.reshape(n * [4])
        else:
            st = refsim.apply_cx(st, qs[0], qs[1], [1.0, 0.0], n)
        st = refsim.apply_gate_depolarizing(st, n, list(qs), p)
    return st.reshape(-1)


def zero_state(st, n):
    st.init_product(np.tile([1.0, 0.0, 0.0, 1.0], n), 0.5 ** n)


def learned(precision):
    cost = costmodel.default_model()
    cost.precision = precision
    return cost


def backend():
    return BasicAer.get_backend('dm_simulator')


def embed_index(sub_index, S, n):
    """Flat index of the n-qubit register whose digits on qubits S are the
    digits of sub_index (qubit order kept) and 0 elsewhere."""
    k = len(S)
    out = 0
    for j, q in enumerate(S):
        d = (sub_index >> (2 * (k - 1 - j))) & 3
        out |= d << (2 * (n - 1 - q))
    return out


# ---------------------------------------------------------------------------
# precision form
# ---------------------------------------------------------------------------

@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_precision_workload_reduced_against_oracle(precision):
    n, p = 9, 5e-4
    gates = workloads.matching_layers(n, 6, 0)
    ref = oracle_noisy(n, gates, p)
    noisy = fusion.noisy_gates(gates, p)
    prog, _ = fusion.fused_program(noisy, n, 4, 'learned', learned(precision), 'seq')
    st = native.DeviceState(n, precision)
    zero_state(st, n)
    prog.run(st)
    got = st.download()
    tr, _ = st.trace_purity()
    st.close()
    err = float(np.max(np.abs(got - ref)))
    if precision == 'fp64':
        assert err <= 1e-12
        assert abs(tr - 1.0) <= 1e-14
    else:
        assert err * 2 ** n <= 2e-4
        assert got[0] == 2.0 ** -n


def test_precision_17q_local_circuit_is_oracle_state_times_idle_qubits():
    n, p = 17, 5e-4
    S = [0, 5, 11, 16]
    k = len(S)
    local = workloads.matching_layers(k, 6, 3)
    gates = [(name, tuple(S[q] for q in qs), prm) for name, qs, prm in local]
    ref = oracle_noisy(k, local, p)
    ref_purity = 2 ** k * float(np.dot(ref, ref))
    noisy = fusion.noisy_gates(gates, p)
    rng = np.random.default_rng(0)
    states = {}
    try:
        for precision in ('fp64', 'fp32'):
            prog, _ = fusion.fused_program(noisy, n, 4, 'learned', learned(precision), 'seq')
            st = native.DeviceState(n, precision)
            states[precision] = st
            zero_state(st, n)
            prog.run(st)
            # the 4^4 coefficients with I on every idle qubit, and the same with Z
            # on a random set of idle qubits (the |0><0| factor is (I + Z) / 2)
            idx = np.array([embed_index(i, S, n) for i in range(4 ** k)], dtype=np.uint64)
            idle = [q for q in range(n) if q not in S]
            zmask = 0
            for q in idle:
                if rng.random() < 0.5:
                    zmask |= 3 << (2 * (n - 1 - q))
            want = ref * 0.5 ** (n - k)
            for mask in (0, zmask):
                got = st.read(idx | np.uint64(mask))
                err = float(np.max(np.abs(got - want)))
                assert err <= (1e-12 * 0.5 ** (n - k) if precision == 'fp64' else 2e-4 * 2.0 ** -n)
            # an X or Y on any idle qubit: exactly zero
            xy = [int(idx[rng.integers(len(idx))]) | (int(rng.integers(1, 3)) << (2 * (n - 1 - int(q))))
                  for q in rng.choice(idle, 64)]
            assert np.all(st.read(np.array(xy, dtype=np.uint64)) == 0.0)
            tr, pur = st.trace_purity()
            assert abs(tr - 1.0) <= (1e-14 if precision == 'fp64' else 1e-7)
            assert abs(pur - ref_purity) <= (1e-12 if precision == 'fp64' else 1e-5)
        cmp = states['fp32'].compare(states['fp64'])
        assert cmp['max_abs_diff'] * 2 ** n <= 2e-4
    finally:
        for st in states.values():
            st.close()


def test_precision_17q_full_width_fp32_against_fp64():
    n = 17
    gates = workloads.matching_layers(n, 2, 0)
    noisy = fusion.noisy_gates(gates, 5e-4)
    states = {}
    try:
        for precision in ('fp64', 'fp32'):
            prog, _ = fusion.fused_program(noisy, n, 4, 'learned', learned(precision), 'seq')
            st = native.DeviceState(n, precision)
            states[precision] = st
            zero_state(st, n)
            prog.run(st)
            tr, pur = st.trace_purity()
            assert abs(tr - 1.0) <= (1e-14 if precision == 'fp64' else 1e-7)
            assert 0.0 < pur <= 1.0 + 1e-9
            assert float(st.read([0])[0]) == 2.0 ** -n
        m = {prec: np.array([states[prec].probabilities([q], 'Z')[0] for q in range(n)]) for prec in states}
        for prec in m:
            assert np.all(m[prec] >= -1e-12) and np.all(m[prec] <= 1 + 1e-12)
        assert float(np.max(np.abs(m['fp32'] - m['fp64']))) <= 1e-5
        cmp = states['fp32'].compare(states['fp64'])
        assert cmp['max_abs_diff'] * 2 ** n <= 2e-4
        assert cmp['dot'] / np.sqrt(cmp['norm2_self'] * cmp['norm2_other']) >= 1 - 1e-6
    finally:
        for st in states.values():
            st.close()


# ---------------------------------------------------------------------------
# plain FP64, brickwork + QFT with thermal relaxation
# ---------------------------------------------------------------------------

THERMAL = dict(zip(('decay_factor', 'decoherence_factor'), workloads.thermal_factors(50e-6, 70e-6, 50e-9)),
               thermal_factor=1.0)


def test_thermal_14q_plain_and_fused_agree_at_full_size():
    n = 14
    gates = workloads.brickwork_qft(n, 30, 0)
    qc = workloads.to_circuit(n, gates, barrier_after_each=True, measure=('partial', [0, 1, 2, 3], 'Z'))
    out = {}
    for mode in ('none', 'greedy', 'learned'):
        res = execute(qc, backend(), fusion=mode, compute_densitymatrix=False, **THERMAL).result()
        out[mode] = res['results'][0]
    base = out['none']
    assert abs(base['data']['coeffmatrix'][0] * 2 ** n - 1.0) <= 1e-14
    pp = base['data']['partial_probability']
    assert abs(sum(pp.values()) - 1.0) <= 1e-12
    for mode in ('greedy', 'learned'):
        assert out[mode]['number_of_clock_cycles'] == base['number_of_clock_cycles']
        assert np.max(np.abs(out[mode]['data']['coeffmatrix'] - base['data']['coeffmatrix'])) <= 1e-12
        got = out[mode]['data']['partial_probability']
        assert list(got) == list(pp)
        assert max(abs(got[k] - pp[k]) for k in pp) <= 1e-12


@pytest.mark.parametrize('mode', ['none', 'learned'])
def test_thermal_14q_embedded_circuit_against_oracle(mode):
    n, S = 14, [0, 3, 9, 13]
    k = len(S)
    local = workloads.brickwork_qft(k, 3, 1)
    qc_small = workloads.to_circuit(k, local, barrier_after_each=True, measure=('partial', list(range(k)), 'Z'))
    big = [(name, tuple(S[q] for q in qs), prm) for name, qs, prm in local]
    qc_big = workloads.to_circuit(n, big, barrier_after_each=True, measure=('partial', S, 'Z'))
    opts = dict(THERMAL, compute_densitymatrix=False)
    exp = assemble(qc_small).experiments[0]
    ref, count, _ = refsim.run_experiment(copy.deepcopy(exp.instructions), k, dict(opts))
    res = execute(qc_big, backend(), fusion=mode, **opts).result()['results'][0]
    assert res['number_of_clock_cycles'] == count
    a = res['data']['coeffmatrix']
    idx = np.array([embed_index(i, S, n) for i in range(4 ** k)])
    assert np.max(np.abs(a[idx] - ref['coeffmatrix'] * 0.5 ** (n - k))) <= 1e-12 * 0.5 ** (n - k)
    # every other coefficient: I or Z on the idle qubits with the same factor, else 0
    r = ref['coeffmatrix']
    assert abs(float(np.dot(a, a)) * 2 ** n - 2 ** k * float(np.dot(r, r))) <= 1e-12
    got = res['data']['partial_probability']
    want = ref['partial_probability']
    assert list(got) == list(want)
    assert max(abs(got[key] - want[key]) for key in want) <= 1e-12


# ---------------------------------------------------------------------------
# learned fusion at 16 qubits
# ---------------------------------------------------------------------------

def test_learned_16q_full_size_fused_against_plain():
    n, G = 16, 1000
    noisy = fusion.noisy_gates(workloads.local_u3_cx(n, G, 0), 1e-3)
    states = {}
    try:
        st = native.DeviceState(n, 'fp64')
        states['plain'] = st
        zero_state(st, n)
        fusion.plain_program(noisy, n).run(st)
        for precision in ('fp64', 'fp32'):
            prog, _ = fusion.fused_program(noisy, n, 4, 'learned', learned(precision), 'seq')
            st = native.DeviceState(n, precision)
            states[precision] = st
            zero_state(st, n)
            prog.run(st)
        c64 = states['fp64'].compare(states['plain'])
        assert c64['max_abs_diff'] <= 1e-12
        c32 = states['fp32'].compare(states['plain'])
        assert c32['max_abs_diff'] * 2 ** n <= 1e-5
        for prec in ('plain', 'fp64', 'fp32'):
            tr, _ = states[prec].trace_purity()
            assert abs(tr - 1.0) <= 1e-14
            assert float(states[prec].read([0])[0]) == 2.0 ** -n
    finally:
        for st in states.values():
            st.close()
