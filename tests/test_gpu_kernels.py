"""Kernel-level parity: each C-ABI entry point against the CPU oracle on
seeded random states, at every qubit position (strides 1, 4, 16, ...).

FP64: max |delta a| <= 1e-14 (single operations; the end-to-end bar is 1e-12).
FP32: max |delta a| <= 2e-7 * 2^-n * ... (relative to the coefficient scale)
and the identity coefficient a_0 = 2^-n stays exact.
"""
import itertools

import numpy as np
import pytest

from oracle import refsim
from dmsim import native, ptm

pytestmark = pytest.mark.gpu


def rand_state(n, rng):
    """Random physical-looking Pauli vector: a_0 = 2^-n, others ~ 2^-n/ sqrt scale."""
    a = rng.normal(size=4 ** n) * (0.3 * 2.0 ** -n)
    a[0] = 2.0 ** -n
    return a


def tol(prec, n):
    return 1e-14 * 2.0 ** -n * 8 if prec == 'fp64' else 3e-7 * 2.0 ** -n * 8


def make(n, prec, a):
    st = native.DeviceState(n, prec)
    st.upload(a)
    return st


PRECS = ['fp64', 'fp32']


@pytest.mark.parametrize('prec', PRECS)
@pytest.mark.parametrize('n', [1, 2, 4, 7])
def test_u3_all_positions(prec, n):
    rng = np.random.default_rng(100 + n)
    for q in range(n):
        a = rand_state(n, rng)
        th, ph, la = rng.uniform(0, 2 * np.pi, 3)
        st = make(n, prec, a)
        st.apply_1q(q, ptm.rows3(ptm.gate_ptm('u3', [th, ph, la])))
        got = st.download()
        want = a.reshape(n * [4]).copy()
        st3 = want.reshape(4 ** q, 4, 4 ** (n - q - 1))
        for axis, ang in refsim.gate_rotations('u3', [th, ph, la]):
            st3 = refsim.rotate_components(st3, axis, ang, [1.0, 0.0])
        assert np.max(np.abs(got - st3.reshape(-1))) <= tol(prec, n), (q, prec)
        assert got[0] == 2.0 ** -n
        st.close()


@pytest.mark.parametrize('prec', PRECS)
def test_1q_affine_reset_and_collapse(prec):
    n = 5
    rng = np.random.default_rng(7)
    for q in range(n):
        a = rand_state(n, rng)
        st = make(n, prec, a)
        st.apply_1q(q, ptm.rows3(ptm.reset()))
        want = refsim.reset_qubit(a.reshape(n * [4]).copy(), n, q).reshape(-1)
        assert np.max(np.abs(st.download() - want)) <= tol(prec, n)
        st.apply_1q(q, ptm.rows3(ptm.collapse_axis(2, 0.9)))
        want, _ = refsim.collapse_axis(want.reshape(n * [4]).copy(), n, q, 2, 0.9)
        assert np.max(np.abs(st.download() - want.reshape(-1))) <= tol(prec, n)
        st.close()


@pytest.mark.parametrize('prec', PRECS)
@pytest.mark.parametrize('n', [2, 3, 5])
def test_cx_all_pairs(prec, n):
    rng = np.random.default_rng(200 + n)
    for c, t in itertools.permutations(range(n), 2):
        a = rand_state(n, rng)
        st = make(n, prec, a)
        st.apply_cx(c, t)
        want = refsim.apply_cx(a.reshape(n * [4]).copy(), c, t, [1.0, 0.0], n).reshape(-1)
        assert np.max(np.abs(st.download() - want)) <= tol(prec, n), (c, t)
        st.close()


@pytest.mark.parametrize('prec', PRECS)
def test_cx_with_depolarizing_scale(prec):
    n = 4
    rng = np.random.default_rng(3)
    for c, t in [(0, 3), (3, 0), (1, 2), (2, 3)]:
        a = rand_state(n, rng)
        st = make(n, prec, a)
        st.apply_cx(c, t, ptm.depolarizing_pair_scale(0.01))
        want = refsim.apply_cx(a.reshape(n * [4]).copy(), c, t, [1.0, 0.0], n)
        want = refsim.apply_gate_depolarizing(want, n, [c, t], 0.01).reshape(-1)
        assert np.max(np.abs(st.download() - want)) <= tol(prec, n)
        st.close()


@pytest.mark.parametrize('prec', PRECS)
def test_tsp_cx_dense(prec):
    n = 4
    rng = np.random.default_rng(4)
    tsp = [0.97, 0.05]
    for c, t in [(0, 1), (1, 0), (0, 3), (3, 1)]:
        a = rand_state(n, rng)
        st = make(n, prec, a)
        st.apply_2q(c, t, ptm.cx_ptm(tsp))
        want = refsim.apply_cx(a.reshape(n * [4]).copy(), c, t, tsp, n).reshape(-1)
        assert np.max(np.abs(st.download() - want)) <= tol(prec, n)
        st.close()


@pytest.mark.parametrize('prec', PRECS)
@pytest.mark.parametrize('fpg', [(0.9, 0.5, 0.99), (0.95, 1.0, 0.97), (0.8, 0.0, 0.9)])
@pytest.mark.parametrize('n', [1, 2, 4, 6, 7, 9, 11])
def test_memory_noise(prec, fpg, n):
    rng = np.random.default_rng(300 + n)
    a = rand_state(n, rng)
    st = make(n, prec, a)
    st.memory_noise(*fpg)
    f, p, g = fpg
    want = refsim.apply_memory_noise(a.reshape(n * [4]).copy(), n, f, p, g).reshape(-1)
    assert np.max(np.abs(st.download() - want)) <= tol(prec, n) * 4
    st.close()


@pytest.mark.parametrize('prec', PRECS)
def test_memory_noise_repeated_with_changing_parameters(prec):
    """The thermal-feed channel's tile passes are kept per handle for the
    last parameters: the same parameters twice, then others, then the first
    again, against the oracle applied in the same order."""
    n = 10
    rng = np.random.default_rng(77)
    a = rand_state(n, rng)
    st = make(n, prec, a)
    want = a.reshape(n * [4]).copy()
    for fpg in ((0.95, 0.8, 0.97), (0.95, 0.8, 0.97), (0.9, 0.3, 0.99), (0.95, 0.8, 0.97)):
        st.memory_noise(*fpg)
        want = refsim.apply_memory_noise(want, n, *fpg)
    assert np.max(np.abs(st.download() - want.reshape(-1))) <= tol(prec, n) * 8
    st.close()


@pytest.mark.parametrize('prec', PRECS)
def test_diag_and_init_product(prec):
    n = 6
    rng = np.random.default_rng(5)
    vecs = rng.normal(size=(n, 4))
    st = native.DeviceState(n, prec)
    st.init_product(vecs.reshape(-1), 0.5 ** n)
    want = np.array([1.0])
    for q in range(n):
        want = np.kron(want, vecs[q])
    want *= 0.5 ** n
    rel = 1e-15 if prec == 'fp64' else 1e-6
    assert np.max(np.abs(st.download() - want)) <= rel * np.max(np.abs(want)) * 4
    tab = rng.uniform(0.5, 1.0, size=(n, 4))
    st.apply_diag(tab.reshape(-1))
    fac = np.array([1.0])
    for q in range(n):
        fac = np.kron(fac, tab[q])
    assert np.max(np.abs(st.download() - want * fac)) <= rel * np.max(np.abs(want)) * 8
    st.close()


@pytest.mark.parametrize('prec', PRECS)
@pytest.mark.parametrize('basis', ['X', 'Y', 'Z'])
def test_probabilities_ensemble_and_partial(prec, basis):
    n = 6
    rng = np.random.default_rng(ord(basis))
    a = rand_state(n, rng)
    st = make(n, prec, a)
    err = 0.93
    got = st.probabilities(list(range(n)), basis, err)
    want = np.array(list(refsim.ensemble_probabilities(a.reshape(n * [4]), n, basis, None, err).values()))
    t = 1e-13 if prec == 'fp64' else 1e-5
    assert np.max(np.abs(got - want)) <= t
    meas = [1, 4]
    got = st.probabilities(meas, basis, err)
    full = want.reshape(n * [2])
    rest = tuple(q for q in range(n) if q not in meas)
    assert np.max(np.abs(got - full.sum(axis=rest).reshape(-1))) <= t
    st.close()


def test_probabilities_n_basis():
    n = 4
    rng = np.random.default_rng(8)
    a = rand_state(n, rng)
    st = make(n, 'fp64', a)
    nv = np.array([1.0, 2.0, 3.0]) / np.sqrt(14)
    got = st.probabilities(list(range(n)), 'N', 0.9, nv)
    want = np.array(list(refsim.ensemble_probabilities(a.reshape(n * [4]), n, 'N', nv, 0.9).values()))
    assert np.max(np.abs(got - want)) <= 1e-14
    st.close()


@pytest.mark.parametrize('prec', PRECS)
def test_trace_purity_overlap(prec):
    n = 7
    rng = np.random.default_rng(9)
    a = rand_state(n, rng)
    b = rand_state(n, rng)
    st = make(n, prec, a)
    tr, pur = st.trace_purity()
    assert tr == 1.0
    t = 1e-12 if prec == 'fp64' else 1e-5
    assert abs(pur - 2 ** n * np.dot(a, a)) <= t * abs(pur)
    assert abs(st.overlap(b) - 2 ** n * np.dot(a, b)) <= t
    st.close()


@pytest.mark.parametrize('n', [1, 3, 5])
def test_to_dense_matches_oracle_bitwise(n):
    rng = np.random.default_rng(10 + n)
    a = rand_state(n, rng)
    st = make(n, 'fp64', a)
    got = st.to_dense()
    want = refsim.pauli_to_matrix(a, n)
    assert np.array_equal(got, want)
    st.close()


def test_download_chop():
    n = 3
    a = np.zeros(64)
    a[0] = 0.125
    a[5] = 3e-16
    a[7] = -2e-15
    st = make(n, 'fp64', a)
    got = st.download(chop=1e-15)
    assert got[5] == 0.0 and got[7] == -2e-15 and got[0] == 0.125
    st.close()


def test_errors_are_reported():
    st = native.DeviceState(3, 'fp64')
    with pytest.raises(native.NativeError, match='out of range'):
        st.apply_1q(3, np.zeros(12))
    with pytest.raises(native.NativeError, match='CX'):
        st.apply_cx(1, 1)
    st.close()


def apply_block_np(a, n, qubits, S):
    """numpy contraction of a k-qubit PTM (component index sum_j i_{qubits[j]} 4^j)."""
    k = len(qubits)
    A = a.reshape(n * [4])
    St = np.asarray(S).reshape([4] * (2 * k))  # [out_{k-1}..out_0, in_{k-1}..in_0]
    in_axes = list(range(k, 2 * k))
    # tensor axis k+m of St is input digit k-1-m -> qubit qubits[k-1-m]
    qaxes = [qubits[k - 1 - m] for m in range(k)]
    out = np.tensordot(St, A, axes=(in_axes, qaxes))  # [out digits..., remaining qubit axes]
    rem = [q for q in range(n) if q not in qaxes]
    # current axes: out_{k-1}..out_0 (-> qubits qaxes), then rem
    cur = qaxes + rem
    perm = [cur.index(q) for q in range(n)]
    return np.transpose(out, perm).reshape(-1)


@pytest.mark.parametrize('prec', PRECS)
@pytest.mark.parametrize('k,n,qubits', [(3, 6, [0, 1, 2]), (3, 7, [6, 2, 4]), (3, 8, [1, 7, 3]),
                                        (4, 6, [5, 0, 3, 2]), (4, 8, [2, 5, 6, 7]), (4, 7, [0, 1, 2, 3])])
def test_block_mfma_general_matrix(prec, k, n, qubits):
    rng = np.random.default_rng(17 * k + n)
    a = rand_state(n, rng)
    S = rng.normal(size=(4 ** k, 4 ** k)) / 4 ** k
    st = make(n, prec, a)
    st.apply_block(qubits, S)
    want = apply_block_np(a, n, qubits, S)
    t = 1e-15 * 2.0 ** -n * 4 ** k if prec == 'fp64' else 2e-6 * 2.0 ** -n * 4
    assert np.max(np.abs(st.download() - want)) <= t
    st.close()


@pytest.mark.parametrize('prec', PRECS)
@pytest.mark.parametrize('mode', ['seq', 'dense'])
@pytest.mark.parametrize('n,kmax', [(6, 4), (7, 3), (8, 4), (9, 2)])
def test_fused_passes_match_plain(prec, mode, n, kmax):
    """Fused tile passes (sequential and matrix-core forms) against one kernel
    per gate on a seeded noisy circuit."""
    from dmsim import fusion, workloads
    gates = fusion.noisy_gates(workloads.local_u3_cx(n, 120, n), 1e-3)
    rng = np.random.default_rng(n)
    a = rand_state(n, rng)
    plain = fusion.plain_program(gates, n)
    fused, blocks = fusion.fused_program(gates, n, kmax, 'greedy', None, mode)
    assert len(fused) < len(plain)
    s1 = make(n, 'fp64', a)
    plain.run(s1)
    want = s1.download()
    s1.close()
    s2 = make(n, prec, a)
    fused.run(s2)
    got = s2.download()
    s2.close()
    t = 1e-14 * 2.0 ** -n * 50 if prec == 'fp64' else 2e-6 * 2.0 ** -n * 50
    assert np.max(np.abs(got - want)) <= t


@pytest.mark.parametrize('prec', PRECS)
@pytest.mark.parametrize('tile', [5, 6, 7])
def test_fused_tile_sizes_match_plain(prec, tile):
    """Sequential tile passes of 5, 6 and 7 digits (7: 512 threads) against
    one kernel per gate."""
    from dmsim import fusion, workloads
    n = 9
    gates = fusion.noisy_gates(workloads.local_u3_cx(n, 150, 40 + tile), 1e-3)
    rng = np.random.default_rng(tile)
    a = rand_state(n, rng)
    plain = fusion.plain_program(gates, n)
    fused, blocks = fusion.fused_program(gates, n, 4, 'greedy', None, 'seq', tile)
    assert any(len(qs) > 1 for qs, _ in blocks)
    s1 = make(n, 'fp64', a)
    plain.run(s1)
    want = s1.download()
    s1.close()
    s2 = make(n, prec, a)
    fused.run(s2)
    got = s2.download()
    s2.close()
    t = 1e-14 * 2.0 ** -n * 50 if prec == 'fp64' else 2e-6 * 2.0 ** -n * 50
    assert np.max(np.abs(got - want)) <= t


@pytest.mark.parametrize('prec', PRECS)
def test_fused_passes_without_parameter_values(prec):
    """Passes of noiseless CX operations carry no parameter values; their
    tile operations still reach the GPU, on a fresh handle and on one whose
    previous program left other operations behind."""
    from dmsim import fusion, workloads
    n = 8
    gates = [g for g in fusion.noisy_gates(workloads.local_u3_cx(n, 160, 5), 0.0) if g.kind == 'cx']
    assert all(g.R is None for g in gates)
    rng = np.random.default_rng(3)
    a = rand_state(n, rng)
    plain = fusion.plain_program(gates, n)
    fused, blocks = fusion.fused_program(gates, n, 4, 'greedy', None, 'seq')
    assert any(len(mem) > 1 for _, mem in blocks)
    s1 = make(n, 'fp64', a)
    plain.run(s1)
    want = s1.download()
    s1.close()
    t = 1e-14 * 2.0 ** -n * 50 if prec == 'fp64' else 2e-6 * 2.0 ** -n * 50
    s2 = make(n, prec, a)
    fused.run(s2)
    assert np.max(np.abs(s2.download() - want)) <= t
    other, _ = fusion.fused_program(fusion.noisy_gates(workloads.local_u3_cx(n, 60, 6), 1e-3), n, 4, 'greedy',
                                    None, 'seq')
    other.run(s2)
    s2.upload(a)
    fused.run(s2)
    assert np.max(np.abs(s2.download() - want)) <= t
    s2.close()


@pytest.mark.parametrize('prec', PRECS)
def test_fused_passes_with_dense_two_qubit_ops(prec):
    """TSP CX gates are dense 16x16 operations inside the tile passes."""
    from dmsim import fusion, workloads
    n = 8
    gates = fusion.noisy_gates(workloads.local_u3_cx(n, 150, 3), 1e-3, tsp=(0.97, 0.05))
    assert any(g.kind == '2q' for g in gates)
    rng = np.random.default_rng(8)
    a = rand_state(n, rng)
    plain = fusion.plain_program(gates, n)
    fused, _ = fusion.fused_program(gates, n, 4, 'greedy', None, 'seq')
    s1 = make(n, 'fp64', a)
    plain.run(s1)
    want = s1.download()
    s1.close()
    s2 = make(n, prec, a)
    fused.run(s2)
    got = s2.download()
    s2.close()
    t = 1e-14 * 2.0 ** -n * 50 if prec == 'fp64' else 2e-6 * 2.0 ** -n * 50
    assert np.max(np.abs(got - want)) <= t


@pytest.mark.parametrize('prec', PRECS)
@pytest.mark.parametrize('m,qs', [(6, [8, 7, 0]), (6, [8, 6, 2, 1]), (7, [8, 7, 3, 0]), (5, [8, 7])])
def test_fused_pass_two_qubit_ops_on_low_digits(prec, m, qs):
    """CX and dense two-qubit operations on the lowest tile digits move those
    digits to the top of the LDS tile (bank-conflict-free layout); results
    must equal one kernel per gate."""
    from dmsim.program import Program
    n = 9
    rng = np.random.default_rng(len(qs) + m)
    a = rand_state(n, rng)
    R = ptm.rows3(ptm.gate_ptm('u3', [0.4, 1.1, -0.3]))
    scale = ptm.depolarizing_pair_scale(2e-3)
    tsp = ptm.cx_ptm((0.97, 0.05))
    sub, plain = [], Program(n)
    for i in range(len(qs)):
        c, t = qs[i], qs[(i + 1) % len(qs)]
        sub += [(native.SUB_1Q, [c], R), (native.SUB_CX, [c, t], scale), (native.SUB_2Q, [t, c], tsp)]
        plain.one_qubit(c, ptm.gate_ptm('u3', [0.4, 1.1, -0.3]))
        plain.cx(c, t, scale)
        plain.two_qubit(t, c, tsp)
    fused = Program(n)
    fused.fused(m, qs, sub)
    s1 = make(n, 'fp64', a)
    plain.run(s1)
    want = s1.download()
    s1.close()
    s2 = make(n, prec, a)
    fused.run(s2)
    got = s2.download()
    s2.close()
    t = 1e-14 * 2.0 ** -n * 50 if prec == 'fp64' else 2e-6 * 2.0 ** -n * 50
    assert np.max(np.abs(got - want)) <= t


@pytest.mark.parametrize('prec', PRECS)
@pytest.mark.parametrize('m,seed', [(5, 1), (6, 2), (6, 3), (7, 4), (7, 5)])
def test_fused_pass_stages_of_random_operations(prec, m, seed):
    """A pass of 48 random operations on 4 qubits of the tile (unital and
    non-unital single-qubit PTMs, scaled and plain CX, dense two-qubit PTMs,
    repeats on one qubit between operations on others): the library
    regroups them into two-digit register stages, moving operations past
    commuting ones and multiplying lone single-qubit runs; results must
    equal one kernel per operation in the given order."""
    from dmsim.program import Program
    n = 9
    rng = np.random.default_rng(100 + seed)
    a = rand_state(n, rng)
    tile_qs = [int(q) for q in rng.choice(n, size=m if m < 7 else 6, replace=False)]
    act = tile_qs[:4]
    sub, plain = [], Program(n)
    for _ in range(48):
        r = rng.random()
        if r < 0.45:
            q = int(rng.choice(act))
            P = ptm.depolarizing1(1e-3) @ ptm.gate_ptm('u3', list(rng.uniform(-3, 3, 3)))
            if r < 0.1:
                # amplitude damping: the Z row has an identity component
                gam = float(rng.uniform(0.01, 0.2))
                P = np.diag([1.0, np.sqrt(1 - gam), np.sqrt(1 - gam), 1 - gam]) @ P
                P[3, 0] = gam
            sub.append((native.SUB_1Q, [q], ptm.rows3(P)))
            plain.one_qubit(q, P)
        else:
            c, t = (int(x) for x in rng.choice(act, size=2, replace=False))
            if r < 0.65:
                S = ptm.depolarizing_pair_scale(float(rng.uniform(1e-4, 1e-2)))
                sub.append((native.SUB_CX, [c, t], S))
                plain.cx(c, t, S)
            elif r < 0.8:
                sub.append((native.SUB_CX, [c, t], None))
                plain.cx(c, t)
            else:
                T = ptm.cx_ptm((0.97, float(rng.uniform(0, 0.1))))
                sub.append((native.SUB_2Q, [c, t], T))
                plain.two_qubit(c, t, T)
    fused = Program(n)
    fused.fused(m, tile_qs, sub)
    s1 = make(n, 'fp64', a)
    plain.run(s1)
    want = s1.download()
    s1.close()
    s2 = make(n, prec, a)
    fused.run(s2)
    got = s2.download()
    s2.close()
    t = 1e-14 * 2.0 ** -n * 50 if prec == 'fp64' else 2e-6 * 2.0 ** -n * 50
    assert np.max(np.abs(got - want)) <= t


def test_fused_pass_rejects_bad_records():
    from dmsim.program import Program
    st = native.DeviceState(6, 'fp64')
    p = Program(6)
    p.fused(6, [0, 1], [(native.SUB_1Q, [9], np.zeros(12))])
    with pytest.raises(native.NativeError, match='not in the fused tile|bad fused'):
        p.run(st)
    p = Program(6)
    p.fused(4, [0], [])
    with pytest.raises(native.NativeError, match='tile size'):
        p.run(st)
    st.close()


def test_compare_states_across_precisions():
    n = 6
    rng = np.random.default_rng(12)
    a = rand_state(n, rng)
    b = a + rng.normal(size=a.size) * 1e-9
    s1 = make(n, 'fp64', a)
    s2 = make(n, 'fp64', b)
    s3 = make(n, 'fp32', a)
    r = s1.compare(s2)
    assert abs(r['max_abs_diff'] - np.max(np.abs(a - b))) <= 1e-18
    assert abs(r['dot'] - np.dot(a, b)) <= 1e-15 * abs(np.dot(a, b))
    r3 = s1.compare(s3)
    assert r3['max_abs_diff'] <= 1e-7 * np.max(np.abs(a))
    for s in (s1, s2, s3):
        s.close()
