"""Host-side behaviour of the GPU backend that needs no GPU: option
snapshots per job, the densitymatrix size gate, error types, the provider
registration a reference maintainer would use, and argument refusals of the
C ABI that happen before any device call."""
import ctypes
import time
import types

import numpy as np
import pytest

from dmsim import BasicAer, BasicAerError, QuantumCircuit, assemble, native
from dmsim import backend as backend_mod
from dmsim.backend import DmSimulator


def small_qobj(n=2):
    qc = QuantumCircuit(n, n)
    qc.h(0)
    qc.cx(0, 1)
    return assemble(qc)


def test_jobs_keep_the_options_they_were_submitted_with(monkeypatch):
    """Two jobs submitted back to back with different decay factors each run
    with their own options, although the backend object is shared and the
    second run() resets its options before the first job executes."""
    be = DmSimulator()
    seen = []

    def fake_run_experiment(self, experiment):
        time.sleep(0.05)
        seen.append((self._decay_factor, dict(self._rotation_error), self._precision))
        return {'name': 'x', 'number_of_clock_cycles': 0, 'data': {}, 'status': 'DONE', 'success': True,
                'processing_time_taken': 0.0, 'running_time_taken': 0.0, 'header': {}}

    monkeypatch.setattr(DmSimulator, 'run_experiment', fake_run_experiment)
    monkeypatch.setattr(DmSimulator, '_check_dense_output', lambda self, n: None)
    q = small_qobj()
    j1 = be.run(q, backend_options={'decay_factor': 0.9, 'rotation_error': {'rx': [0.9, 0.1]}})
    j2 = be.run(q, backend_options={'decay_factor': 0.5, 'precision': 'fp32'})
    j1.result()
    j2.result()
    assert seen[0][0] == 0.9 and seen[0][1]['rx'] == [0.9, 0.1] and seen[0][2] == 'fp64'
    assert seen[1][0] == 0.5 and seen[1][2] == 'fp32'
    # the reference's leak of rotation_error into later runs (shared
    # defaults dict) is kept on the backend itself
    assert be._rotation_error['rx'] == [0.9, 0.1]
    be.DEFAULT_OPTIONS['rotation_error'].update({'rx': [1.0, 0.0]})


def test_densitymatrix_size_gate(monkeypatch):
    be = DmSimulator()
    be._precision = 'fp64'
    monkeypatch.setattr(native, 'device_info', lambda d=0: {'devices': 1, 'cu_count': 256,
                                                             'hbm_bytes': 288 * 10 ** 9})
    monkeypatch.setattr(backend_mod, '_host_available_bytes', lambda: 512 * 10 ** 9)
    be._check_dense_output(16)             # 68.7 GB dense + 34.4 GB state
    with pytest.raises(BasicAerError, match='compute_densitymatrix=False'):
        be._check_dense_output(17)         # 275 GB dense
    monkeypatch.setattr(backend_mod, '_host_available_bytes', lambda: 10 * 10 ** 9)
    with pytest.raises(BasicAerError, match='densitymatrix of 15 qubits'):
        be._check_dense_output(15)         # 17.2 GB on a 10 GB host
    be._check_dense_output(14)             # 4.3 GB


def test_native_errors_are_basicaer_errors():
    assert issubclass(native.NativeError, BasicAerError)
    assert issubclass(native.NativeError, RuntimeError)


def test_status_and_properties():
    be = BasicAer.get_backend('dm_simulator')
    st = be.status()
    assert st.backend_name == 'dm_simulator' and st.operational is True and st.pending_jobs == 0
    assert set(st.to_dict()) == {'backend_name', 'backend_version', 'operational', 'pending_jobs', 'status_msg'}
    assert be.properties() is None


def test_reference_provider_registration_resolves_the_gpu_backend():
    """The reference's provider keys backends by name() in SIMULATORS order
    (basicaerprovider.py:105-125), so listing the GPU class after
    DmSimulatorPy makes get_backend('dm_simulator') return it.  A minimal
    provider with the reference's resolution steps (instantiate each class,
    key by name(), later entries win, deprecated aliases map to names)."""
    class RefDmSimulatorPy:
        def __init__(self, provider=None):
            pass

        def name(self):
            return 'dm_simulator'

    class RefQasmSimulatorPy(RefDmSimulatorPy):
        def name(self):
            return 'qasm_simulator'

    SIMULATORS = [RefQasmSimulatorPy, RefDmSimulatorPy, DmSimulator]

    def verify_backends():
        ret = {}
        for cls in SIMULATORS:
            inst = cls(provider=None)
            ret[inst.name()] = inst
        return ret

    backends = verify_backends()
    deprecated = {'dm_simulator_py': 'dm_simulator'}

    def get_backend(name):
        return backends[deprecated.get(name, name)]

    assert isinstance(get_backend('dm_simulator'), DmSimulator)
    assert isinstance(get_backend('dm_simulator_py'), DmSimulator)
    assert isinstance(get_backend('qasm_simulator'), RefQasmSimulatorPy)


def test_batch_create_checks_arguments_before_allocating():
    lib = native.load()

    def create(widths, lds_max=7, threads=256, ipp=256):
        n = len(widths)
        nq = (ctypes.c_int * n)(*widths)
        starts = (ctypes.c_int64 * (n + 1))(*([0] * (n + 1)))
        err = (ctypes.c_double * n)(*([1.0] * n))
        ops = (native.DmOp * 1)()
        params = (ctypes.c_double * 1)()
        h = ctypes.c_void_p()
        rc = lib.dm_batch_create(native.FP64, n, nq, starts, ctypes.cast(ops, ctypes.c_void_p), params, 0, err,
                                 lds_max, threads, ipp, ctypes.byref(h))
        return rc, h.value, native.last_error()

    rc, h, msg = create([8, 13])
    assert rc == -1 and not h and '13 qubits' in msg
    rc, h, msg = create([8], lds_max=4)
    assert rc == -1 and not h and 'lds_max_qubits' in msg
    rc, h, msg = create([8], ipp=0)
    assert rc == -1 and not h and 'items_per_pass' in msg
    rc, h, msg = create([8], threads=100)
    assert rc == -1 and not h and 'lds_threads' in msg
