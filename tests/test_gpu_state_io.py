"""Chunked state I/O: the range copies and dot products of the C ABI, and the
backend's streamed .npy files (store_densitymatrix, stored_density_matrix,
compare; dm_simulator.py:262-271, 337-343, 1271-1282).

Copies are exact (FP64 <-> FP64, FP32 <-> FP32, FP32 -> FP64; FP64 -> FP32
is the rounding numpy does); dot products agree with numpy's to 1e-13 of
sum |a_i b_i|; a state restored from its file continues bit for bit.
"""
import os

import numpy as np
import pytest

from dmsim import BasicAer, QuantumCircuit, backend as backend_mod, execute, native

pytestmark = pytest.mark.gpu


def backend():
    return BasicAer.get_backend('dm_simulator')


@pytest.fixture(autouse=True)
def clean_flags():
    be = backend()

    def reset():
        be._custom_densitymatrix = None
        be._initial_densitymatrix = None
        be._get_den_mat = True
        for flag in ('STORE_LOCAL', 'COMPARE', 'FILE_EXIST'):
            be.__dict__.pop(flag, None)
    reset()
    yield
    reset()


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_range_copies_and_dots(precision):
    n = 7
    rng = np.random.default_rng(1)
    a = rng.normal(size=4 ** n) * 2.0 ** -n
    st = native.DeviceState(n, precision)
    st.upload(a)
    full = st.download()
    for off, cnt in ((0, 4 ** n), (3, 1000), (4 ** n - 17, 17), (5, 0)):
        out64 = np.full(cnt, np.nan)
        st.download_range(off, out64)
        assert np.array_equal(out64, full[off:off + cnt])
        out32 = np.full(cnt, np.nan, dtype=np.float32)
        st.download_range(off, out32)
        assert np.array_equal(out32, full[off:off + cnt].astype(np.float32))
        b = rng.normal(size=cnt)
        d = st.dot_range(off, b)
        scale = float(np.dot(np.abs(full[off:off + cnt]), np.abs(b))) + 1e-300
        assert abs(d - float(np.dot(full[off:off + cnt], b))) <= 1e-13 * scale
        b32 = b.astype(np.float32)
        d32 = st.dot_range(off, b32)
        assert abs(d32 - float(np.dot(full[off:off + cnt], b32.astype(np.float64)))) <= 1e-13 * scale
    new = rng.normal(size=300)
    st.upload_range(100, new)
    got = st.download()
    want = full.copy()
    want[100:400] = new if precision == 'fp64' else new.astype(np.float32)
    assert np.array_equal(got, want)
    st.upload_range(500, new.astype(np.float32))
    want[500:800] = new.astype(np.float32)
    assert np.array_equal(st.download(), want)
    with pytest.raises(native.NativeError):
        st.download_range(4 ** n - 5, np.zeros(6))
    st.close()


def circuit_a(n):
    qc = QuantumCircuit(n, n)
    for q in range(n):
        qc.u3(0.3 + q, 0.2 * q, 0.1, q)
    for q in range(n - 1):
        qc.cx(q, q + 1)
    return qc


def add_b(qc, n):
    qc.h(n - 1)
    qc.cx(n - 1, 0)
    qc.u3(1.1, 0.4, -0.3, 2)
    qc.measure(0, 0, basis='Ensemble', add_param='Z')
    return qc


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_store_load_and_compare_in_chunks(precision, tmp_path, monkeypatch):
    """Circuit A stored (file in the storage precision, several chunks),
    loaded as the initial state of circuit B: the same state as A then B in
    one run; compare against the stored file gives 2^n <a, b>."""
    monkeypatch.chdir(tmp_path)
    monkeypatch.setattr(backend_mod, 'IO_CHUNK', 1000)
    n = 8
    be = backend()
    qa = circuit_a(n)
    qa.measure(0, 0, basis='Ensemble', add_param='Z')
    ra = execute(qa, be, precision=precision, store_densitymatrix=True, compute_densitymatrix=False).result()
    stored = np.load('stored_coefficients.npy')
    assert stored.dtype == (np.float32 if precision == 'fp32' else np.float64)
    assert np.array_equal(stored.astype(np.float64), ra['results'][0]['data']['coeffmatrix'])
    assert not os.path.exists('stored_coefficients.partial.npy')
    os.replace('stored_coefficients.npy', 'stored_density_matrix.npy')
    # the reference reads the file when initial_densitymatrix is given (any value)
    rb = execute(add_b(QuantumCircuit(n, n), n), be, precision=precision,
                 custom_densitymatrix='stored_density_matrix', initial_densitymatrix='file',
                 store_densitymatrix=False, compute_densitymatrix=False).result()
    be._custom_densitymatrix = None   # sticky on the backend, as in the reference
    be._initial_densitymatrix = None
    both = add_b(circuit_a(n), n)
    rab = execute(both, be, precision=precision, compute_densitymatrix=False).result()
    got = rb['results'][0]['data']['coeffmatrix']
    want = rab['results'][0]['data']['coeffmatrix']
    assert np.array_equal(got, want)
    # compare: stored file of this result, then the same run compared with it
    execute(both, be, precision=precision, store_densitymatrix=True, compute_densitymatrix=False).result()
    rc = execute(both, be, precision=precision, store_densitymatrix=False, compare=True,
                 compute_densitymatrix=False).result()
    a = rc['results'][0]['data']['coeffmatrix']
    assert abs(rc['results'][0]['data']['fidelity'] - 2 ** n * float(np.dot(a, a))) <= 1e-12


def test_checkpoint_14q_roundtrip(tmp_path):
    """A 14-qubit FP64 state (2.1 GB) through the default chunking: stored,
    restored into a second state bit for bit, overlap = purity."""
    n = 14
    st = native.DeviceState(n, 'fp64')
    st.init_product(np.tile([1.0, 0.3, -0.2, 0.5], n), 0.5 ** n)
    for q in range(0, n, 3):
        st.apply_cx(q, (q + 5) % n)
    path = str(tmp_path / 'stored_coefficients.npy')
    backend_mod.store_state(st, path)
    a = backend_mod.load_state(path)
    st2 = native.DeviceState(n, 'fp64')
    backend_mod.upload_state(st2, a)
    cmp = st2.compare(st)
    assert cmp['max_abs_diff'] == 0.0
    tr, pur = st.trace_purity()
    assert abs(backend_mod.state_overlap(st, a) - pur) <= 1e-12 * pur
    st.close()
    st2.close()
