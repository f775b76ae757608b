"""The ``dm_simulator`` backend on the GPU.

Mirrors the reference backend ``DmSimulatorPy``
(qiskit/providers/basicaer/dm_simulator.py): the same name, configuration
shape, ``run(qobj, backend_options)`` -> job -> ``result()`` dict, the same
``backend_options`` keys with the same meaning (retention factors, not error
rates), the same partition-level semantics and the same result keys.  The
coefficient vector lives on the GPU for the whole experiment; every state
update and every reduction is a libdmsim call.  There is no CPU execution
path: without the native library the backend raises.

Options added by this backend (no reference counterpart):
  ``precision``           'fp64' (default) or 'fp32' storage; reductions are FP64.
  ``gate_depolarization`` probability p of a depolarising channel applied on
                          the qubits of every executed u1/u3/cx (default 0).
  ``fusion``              'none' (default), 'greedy' or 'learned': fuse gates
                          into k-qubit Pauli-transfer blocks (dmsim.fusion).
"""
import concurrent.futures
import copy
import itertools
import logging
import math
import os
import time
import types
import uuid

import numpy as np

from . import native, ptm, schedule
from .exceptions import BasicAerError

logger = logging.getLogger(__name__)


# device staging buffer of the dense output's scatter (dm_to_dense)
DENSE_CHUNK_BYTES = 1 << 31
# coefficients per chunk of the streamed .npy state files (512 MB in FP64)
IO_CHUNK = 1 << 26


def store_state(st, path):
    """Write the state's coefficients to the .npy file ``path`` in chunks
    (the reference's _store_density_matrix, dm_simulator.py:1271-1275, holds
    the whole FP64 array).  The file takes the storage precision (float64 or
    float32); it is written under a temporary name and renamed, so a reader
    that mapped the previous file keeps its contents."""
    dtype = np.float32 if st.dtype == native.FP32 else np.float64
    tmp = path[:-4] + '.partial.npy'
    mm = np.lib.format.open_memmap(tmp, mode='w+', dtype=dtype, shape=(st.size,))
    try:
        for i in range(0, st.size, IO_CHUNK):
            st.download_range(i, mm[i:i + IO_CHUNK])
        mm.flush()
    finally:
        del mm
    os.replace(tmp, path)


def load_state(path):
    """The coefficients of a .npy state file, memory-mapped (not read)."""
    a = np.load(path, mmap_mode='r')
    if a.ndim != 1 or a.dtype not in (np.float64, np.float32):
        a = np.asarray(a, dtype=np.float64).reshape(-1)
    return a


def upload_state(st, a):
    for i in range(0, a.size, IO_CHUNK):
        st.upload_range(i, np.ascontiguousarray(a[i:i + IO_CHUNK]))


def state_overlap(st, a):
    """2^n <a_state, a> (_state_overlap, dm_simulator.py:1277-1282), chunked."""
    total = 0.0
    for i in range(0, a.size, IO_CHUNK):
        total += st.dot_range(i, np.ascontiguousarray(a[i:i + IO_CHUNK]))
    return total * 2.0 ** st.n


def _host_available_bytes():
    """MemAvailable of this host, or None where it cannot be read."""
    try:
        with open('/proc/meminfo') as f:
            for line in f:
                if line.startswith('MemAvailable:'):
                    return int(line.split()[1]) * 1024
    except OSError:
        pass
    return None


def _hbm_max_qubits():
    try:
        hbm = native.device_info(0)['hbm_bytes']
    except Exception:  # no device visible (e.g. configuration queried on a CPU host)
        return 17
    n = 1
    while (4 ** (n + 1)) * 8 * 1.1 < hbm:
        n += 1
    return n


class DmSimJob:
    """Future-backed job (the reference's BasicAerJob, basicaerjob.py:40-126).

    Jobs run on one worker thread: a GPU process must not fork workers."""

    _executor = concurrent.futures.ThreadPoolExecutor(max_workers=1)

    def __init__(self, backend, job_id, fn, qobj):
        self._backend = backend
        self._job_id = job_id
        self._fn = fn
        self._qobj = qobj
        self._future = None

    def submit(self):
        if self._future is not None:
            raise BasicAerError('We have already submitted the job!')
        self._future = self._executor.submit(self._fn, self._job_id, self._qobj)

    def result(self, timeout=None):
        return self._future.result(timeout=timeout)

    def cancel(self):
        return self._future.cancel()

    def status(self):
        if self._future.running():
            return 'RUNNING'
        if self._future.cancelled():
            return 'CANCELLED'
        if self._future.done():
            return 'DONE' if self._future.exception() is None else 'ERROR'
        return 'INITIALIZING'

    def job_id(self):
        return self._job_id

    def backend(self):
        return self._backend

    def qobj(self):
        return self._qobj


class DmSimulator:
    """GPU density-matrix simulator in the Pauli basis."""

    DEFAULT_CONFIGURATION = {
        'backend_name': 'dm_simulator',
        'backend_version': '2.0.0',
        'url': 'https://github.com/Qiskit/qiskit-terra',
        'simulator': True,
        'local': True,
        'conditional': True,
        'open_pulse': False,
        'memory': True,
        'max_shots': 1,
        'coupling_map': None,
        'description': 'A GPU (MI355X) simulator for qasm experiments in the Pauli basis',
        'basis_gates': ['u1', 'u2', 'u3', 'cx', 'id', 'unitary'],
        'gates': [],
    }

    # Shared dict: ``rotation_error`` updates it in place, as the reference
    # does (dm_simulator.py:128, 182, 212-213), so an error set in one run
    # persists into later runs of the process.
    DEFAULT_OPTIONS = {
        'initial_densitymatrix': None,
        'chop_threshold': 1e-15,
        'thermal_factor': 1.,
        'decoherence_factor': 1.,
        'depolarization_factor': 1.,
        'bell_depolarization_factor': 1.,
        'decay_factor': 1.,
        'rotation_error': {'rx': [1., 0.], 'ry': [1., 0.], 'rz': [1., 0.]},
        'tsp_model_error': [1., 0.],
    }

    SHOW_FINAL_STATE = True
    PLOTTING = False
    SHOW_PARTITION = False
    STORE_LOCAL = False
    COMPARE = False
    FILE_EXIST = False
    MERGE = True

    def __init__(self, configuration=None, provider=None):
        self._provider = provider
        self._configuration = configuration
        self._custom_densitymatrix = None
        self._initial_densitymatrix = None
        self._chop_threshold = self.DEFAULT_OPTIONS['chop_threshold']
        self._get_den_mat = True
        self._fidelity = None
        self._density_matrix_stored = None
        self._precision = 'fp64'
        self._gate_depolarization = 0.0
        self._fusion = 'none'

    # identity -------------------------------------------------------------------
    def name(self):
        return self.DEFAULT_CONFIGURATION['backend_name']

    def configuration(self):
        if self._configuration is None:
            cfg = dict(self.DEFAULT_CONFIGURATION)
            cfg['n_qubits'] = _hbm_max_qubits()
            self._configuration = types.SimpleNamespace(**cfg)
            self._configuration.to_dict = lambda c=cfg: dict(c)
        return self._configuration

    def provider(self):
        return self._provider

    def properties(self):
        """No calibration data (basebackend.py:61-68: a simulator returns None)."""
        return None

    def status(self):
        """Backend status with the fields of basebackend.py:78-90."""
        st = {'backend_name': self.name(), 'backend_version': self.DEFAULT_CONFIGURATION['backend_version'],
              'operational': True, 'pending_jobs': 0, 'status_msg': ''}
        ns = types.SimpleNamespace(**st)
        ns.to_dict = lambda s=st: dict(s)
        return ns

    def __repr__(self):
        return "<DmSimulator('dm_simulator')>"

    # options (dm_simulator.py:177-271) ---------------------------------------------------
    def _set_options(self, qobj_config=None, backend_options=None):
        d = self.DEFAULT_OPTIONS
        self._initial_densitymatrix = d['initial_densitymatrix']
        self._chop_threshold = d['chop_threshold']
        self._rotation_error = d['rotation_error']
        self._tsp_model_error = d['tsp_model_error']
        self._thermal_factor = d['thermal_factor']
        self._decoherence_factor = d['decoherence_factor']
        self._decay_factor = d['decay_factor']
        self._depolarization_factor = d['depolarization_factor']
        self._bell_depolarization_factor = d['bell_depolarization_factor']
        self._precision = 'fp64'
        self._gate_depolarization = 0.0
        self._fusion = 'none'
        opts = backend_options or {}

        if 'initial_densitymatrix' in opts:
            init = opts['initial_densitymatrix']
            self._initial_densitymatrix = init if isinstance(init, str) else np.array(init, dtype=float)
        elif hasattr(qobj_config, 'initial_densitymatrix'):
            self._initial_densitymatrix = np.array(qobj_config.initial_densitymatrix, dtype=float)
        if 'custom_densitymatrix' in opts:
            self._custom_densitymatrix = opts['custom_densitymatrix']
            if self._custom_densitymatrix in ('binary_string', 'stored_density_matrix'):
                self._initial_densitymatrix = opts.get('initial_densitymatrix')
        if 'rotation_error' in opts:
            re = opts['rotation_error']
            if type(re) != dict or not all(x in ['rx', 'ry', 'rz'] for x in re):
                raise BasicAerError('Error! Incorrect Rotation Error parameters, Expected argument : A dict with '
                                    'rotation gate as key and a list of 2 reals ranging between 0 and 1 both '
                                    'inclusive as their values.')
            for gt, vl in re.items():
                self._rotation_error.update({gt: vl})
        if 'tsp_model_error' in opts:
            te = opts['tsp_model_error']
            if type(te) != list or len(te) != 2 or te[0] > 1 or te[1] > 1:
                raise BasicAerError('Error! Incorrect transition model error parameter, Expected argument : A '
                                    'list of 2 reals ranging between 0 and 1 both inclusive.')
            self._tsp_model_error = te
        if 'thermal_factor' in opts:
            self._thermal_factor = opts['thermal_factor']
        if 'decoherence_factor' in opts:
            self._decoherence_factor = opts['decoherence_factor']
        if 'decay_factor' in opts:
            self._decay_factor = opts['decay_factor']
        if 'depolarization_factor' in opts:
            self._depolarization_factor = opts['depolarization_factor']
        if 'bell_depolarization_factor' in opts:
            # stored on an attribute the measurement never reads, as in the
            # reference (dm_simulator.py:240): Bell measurements stay lossless
            self.bell_depolarization_factor = opts['bell_depolarization_factor']
        if 'chop_threshold' in opts:
            self._chop_threshold = opts['chop_threshold']
        elif hasattr(qobj_config, 'chop_threshold'):
            self._chop_threshold = qobj_config.chop_threshold
        if 'compute_densitymatrix' in opts:
            self._get_den_mat = opts['compute_densitymatrix']
        if 'merge' in opts:
            self.MERGE = opts['merge']
        if 'plot' in opts:
            self.PLOTTING = opts['plot']
        if 'show_partition' in opts:
            self.SHOW_PARTITION = opts['show_partition']
        if 'store_densitymatrix' in opts:
            self.STORE_LOCAL = opts['store_densitymatrix']
        if 'compare' in opts:
            self.COMPARE = opts['compare']
            try:
                self._density_matrix_stored = load_state('stored_coefficients.npy')
                self.FILE_EXIST = True
            except FileNotFoundError:
                print('Stored Coefficient File does not exist')
        if 'precision' in opts:
            if opts['precision'] not in ('fp64', 'fp32'):
                raise BasicAerError("precision must be 'fp64' or 'fp32'")
            self._precision = opts['precision']
        if 'gate_depolarization' in opts:
            p = float(opts['gate_depolarization'])
            if not 0.0 <= p <= 1.0:
                raise BasicAerError('gate_depolarization must be a probability in [0, 1]')
            self._gate_depolarization = p
        if 'fusion' in opts:
            if opts['fusion'] not in ('none', 'greedy', 'learned'):
                raise BasicAerError("fusion must be 'none', 'greedy' or 'learned'")
            self._fusion = opts['fusion']

    # job plumbing -------------------------------------------------------------------------
    def run(self, qobj, backend_options=None):
        """Set the options on this (shared) backend, as the reference does
        (dm_simulator.py:889-919), then run the job on a snapshot of them:
        the reference's process-pool job pickles the backend at submission,
        so a later run() cannot change the options of a queued job."""
        self._set_options(qobj_config=getattr(qobj, 'config', None), backend_options=backend_options)
        job = DmSimJob(self, str(uuid.uuid4()), self._snapshot()._run_job, qobj)
        job.submit()
        return job

    def _snapshot(self):
        """A copy of the backend carrying the options in force now; option
        containers that later runs modify in place (the shared rotation-error
        dict, lists, arrays) are copied too."""
        snap = copy.copy(self)
        for attr in ('_rotation_error', '_tsp_model_error', '_initial_densitymatrix', '_density_matrix_stored'):
            val = getattr(self, attr, None)
            # a memory-mapped stored state is read-only and outlives the
            # rename of a newer file: shared, not copied
            if isinstance(val, (dict, list, np.ndarray)) and not isinstance(val, np.memmap):
                setattr(snap, attr, copy.deepcopy(val))
        return snap

    def _validate(self, qobj):
        n = qobj.config.n_qubits
        mx = self.configuration().n_qubits
        if n > mx:
            raise BasicAerError('Number of qubits {} is greater than maximum ({}) for "{}".'.format(
                n, mx, self.name()))
        for exp in qobj.experiments:
            if 'measure' not in [op.name for op in exp.instructions]:
                logger.warning('No measurements in circuit "%s", classical register will remain all zeros.',
                               exp.header.name)

    def _check_dense_output(self, n):
        """The densitymatrix output of n qubits is a 2^n x 2^n complex128
        array (16 * 4^n bytes) on the host, built on the device in one buffer
        of that size next to the state (dm_to_dense).  Refuse before the
        simulation runs when either does not fit."""
        dense = 16 * 4 ** n
        state = (8 if self._precision == 'fp64' else 4) * 4 ** n
        try:
            hbm = native.device_info(0)['hbm_bytes']
        except Exception:  # no device visible: the run itself will report it
            hbm = None
        host = _host_available_bytes()
        need_dev = dense + state + min(dense, DENSE_CHUNK_BYTES)
        if (hbm is not None and need_dev > 0.95 * hbm) or (host is not None and dense > host):
            raise BasicAerError(
                'The densitymatrix of {} qubits needs {:.1f} GB on the host and {:.1f} GB of device memory '
                '(available: {} host, {} device); run with compute_densitymatrix=False for the Pauli '
                'coefficients only.'.format(n, dense / 1e9, need_dev / 1e9,
                                            'unknown' if host is None else '%.1f GB' % (host / 1e9),
                                            'unknown' if hbm is None else '%.1f GB' % (hbm / 1e9)))

    def _run_job(self, job_id, qobj):
        self._validate(qobj)
        if self._get_den_mat:
            for exp in qobj.experiments:
                self._check_dense_output(exp.config.n_qubits)
        t0 = time.time()
        results = [self.run_experiment(exp) for exp in qobj.experiments]
        t1 = time.time()
        return {'backend_name': self.name(),
                'backend_version': self.DEFAULT_CONFIGURATION['backend_version'],
                'qobj_id': qobj.qobj_id,
                'job_id': job_id,
                'results': results,
                'status': 'COMPLETED',
                'success': True,
                'time_taken': t1 - t0,
                'header': qobj.header.as_dict() if hasattr(qobj.header, 'as_dict') else {}}

    # the experiment -------------------------------------------------------------------------
    def _initial_state(self, st, n):
        """dm_simulator.py:284-365 on the device."""
        init, custom = self._initial_densitymatrix, self._custom_densitymatrix
        if init is None:
            if custom is None:
                one = [1.0, 0.0, 0.0, 1.0]
            elif custom == 'max_mixed':
                one = [1.0, 0.0, 0.0, 0.0]
            elif custom == 'uniform_superpos':
                one = [1.0, 1.0, 0.0, 0.0]
            elif custom == 'thermal_state':
                one = [1.0, 0.0, 0.0, 2 * self._thermal_factor - 1]
            else:
                raise BasicAerError('_custom_densitymatrix value is invalid')
            st.init_product(np.tile(one, n), 0.5 ** n)
            return
        if custom == 'binary_string':
            bits = str(init)
            if len(bits) != n:
                raise BasicAerError('Wrong input binary string length')
            vecs = np.zeros((n, 4))
            for q in range(n):
                # character i of the string is qubit n-1-i (kron(new, old))
                vecs[q] = [1.0, 0.0, 0.0, 1.0 if bits[n - 1 - q] == '0' else -1.0]
            st.init_product(vecs.reshape(-1), 0.5 ** n)
            return
        if custom == 'stored_density_matrix':
            try:
                a = load_state('stored_density_matrix.npy')
            except FileNotFoundError:
                raise BasicAerError('Stored Coefficient File does not exist')
            if len(a) != 4 ** n:
                raise BasicAerError('Wrong input stored density matrix')
            if a[0] != 2.0 ** (-n):
                raise BasicAerError('Trace of initial densitymatrix is not one: {} != {}'.format(a[0], 2.0 ** -n))
            upload_state(st, a)
            return
        raise BasicAerError('_custom_densitymatrix value is invalid')

    def _gate_1q(self, st, op):
        R = ptm.gate_ptm(op.name, getattr(op, 'params', None), self._rotation_error)
        if self._gate_depolarization:
            R = ptm.depolarizing1(self._gate_depolarization) @ R
        st.apply_1q(op.qubits[0], ptm.rows3(R))

    def _gate_cx(self, st, op):
        c, t = op.qubits[0], op.qubits[1]
        if c == t or c >= st.n or t >= st.n:
            raise BasicAerError('Qubit Labels out of bound in CX Gate')
        p = self._gate_depolarization
        if list(self._tsp_model_error) == [1.0, 0.0]:
            st.apply_cx(c, t, ptm.depolarizing_pair_scale(p) if p else None)
        else:
            R = ptm.cx_ptm(self._tsp_model_error)
            if p:
                R = np.diag(ptm.depolarizing_pair_scale(p)) @ R
            st.apply_2q(c, t, R)

    def _collapse(self, st, q, kind, nvec, err):
        if kind == 'N':
            st.apply_1q(q, ptm.rows3(ptm.collapse_direction(nvec, err)))
        else:
            st.apply_1q(q, ptm.rows3(ptm.collapse_axis({'X': 1, 'Y': 2, 'Z': 3}[kind], err)))

    def _probabilities(self, st, qubits, basis, add_param, err):
        """Probabilities keyed by '0'/'1' strings, qubit order ascending;
        side effects of the reference's ensemble measure (store / compare)."""
        qs = sorted(qubits)
        nv = np.asarray(add_param, dtype=float) if basis == 'N' else None
        vals = st.probabilities(qs, basis, err, nv)
        keys = [''.join(s) for s in itertools.product('01', repeat=len(qs))]
        if self.STORE_LOCAL:
            store_state(st, 'stored_coefficients.npy')
        if self.COMPARE and self.FILE_EXIST:
            if len(self._density_matrix_stored) != st.size:
                raise BasicAerError('Stored coefficients of a different register size')
            self._fidelity = state_overlap(st, self._density_matrix_stored)
        return dict(zip(keys, [float(v) for v in vals]))

    def _bell(self, st, n, qa, qb, err):
        """Bell measurement on the mirrored axes the reference addresses
        (dm_simulator.py:721-777)."""
        q1, q2 = min(qa, qb), max(qa, qb)
        ua, ub = n - 1 - q2, n - 1 - q1  # qubits whose stride is 4^q2 and 4^q1
        idx = [i * 4 ** q2 + j * 4 ** q1 for i in range(4) for j in range(4)]
        red = st.read(idx).reshape(4, 4) * 2.0 ** (n - 2)
        R = np.zeros((16, 16))
        R[0, 0] = 1.0
        for i in (1, 2, 3):
            R[5 * i, 5 * i] = err
        st.apply_2q(ua, ub, R)
        k = st.read([i * 4 ** q2 + i * 4 ** q1 for i in range(4)]) * 2.0 ** n
        probs = [0.25 * (k[0] + k[1] - k[2] + k[3]), 0.25 * (k[0] - k[1] + k[2] + k[3]),
                 0.25 * (k[0] + k[1] + k[2] - k[3]), 0.25 * (k[0] - k[1] - k[2] - k[3])]
        return dict(zip(['Bell_1', 'Bell_2', 'Bell_3', 'Bell_4'], [float(x) for x in probs])), red

    def _expect(self, st, n, pstring, err):
        """Collapse onto a Pauli string and read its coefficient
        (dm_simulator.py:540-572)."""
        letter = {'I': 0, 'X': 1, 'Y': 2, 'Z': 3}
        tab = np.ones((n, 4))
        for q in range(n):
            k = letter[pstring[q]]
            if k:
                tab[q] = 0.0
                tab[q, 0] = 1.0
                tab[q, k] = err
        st.apply_diag(tab.reshape(-1))
        idx = sum(letter[pstring[q]] * 4 ** (n - 1 - q) for q in range(n))
        return float(st.read([idx])[0] * 2.0 ** n)

    @staticmethod
    def _unit(v):
        v = np.asarray(v, dtype=float)
        nrm = np.linalg.norm(v)
        if nrm == 1:
            return v
        logger.warning('Given direction for the measurement was not normalised. It has been normalised to be '
                       'unit vector!!')
        return v / nrm

    def run_experiment(self, experiment):
        """dm_simulator.py:950-1196 with the state on the GPU."""
        t_start = time.time()
        n = experiment.config.n_qubits
        insts = copy.deepcopy(list(experiment.instructions))
        st = native.DeviceState(n, self._precision)
        try:
            self._initial_state(st, n)
            insts = schedule.merge_single_gates(insts, n, self.MERGE)
            levels, count = schedule.partition(insts, n)
            if self.SHOW_PARTITION:
                describe_partition(levels[:count])
            t_proc = time.time()
            data = {}
            self._fidelity = None
            f, p, g = self._decoherence_factor, self._thermal_factor, self._decay_factor
            meas_err = self._depolarization_factor
            bell_err = self._bell_depolarization_factor
            if self._fusion != 'none':
                from . import fusion
                plan = fusion.plan_levels(levels, count, n, self)
            else:
                plan = None
            self._run_levels(st, n, levels, count, data, f, p, g, meas_err, bell_err, plan)
            if self._get_den_mat:
                data['densitymatrix'] = st.to_dense()
            data['coeffmatrix'] = st.download(chop=self._chop_threshold)
            if self._fidelity is not None:
                data['fidelity'] = self._fidelity
            t_end = time.time()
        finally:
            st.close()
        return {'name': experiment.header.name,
                'number_of_clock_cycles': count,
                'data': data,
                'status': 'DONE',
                'success': True,
                'processing_time_taken': t_proc - t_start,
                'running_time_taken': t_end - t_proc,
                'header': experiment.header.as_dict() if hasattr(experiment.header, 'as_dict') else {}}

    def _run_levels(self, st, n, levels, count, data, f, p, g, meas_err, bell_err, plan):
        creg = cmem = 0
        part_measure = False
        clock = 0
        while clock < count:
            if plan is not None and clock in plan:
                seg_end, runner = plan[clock]
                runner(st)
                clock = seg_end
                continue
            level = levels[clock]
            for op in level:
                if op.name == 'measure':
                    basis = str(getattr(level[0], 'params', ['Z'])[0])
                    prm = getattr(op, 'params', ['Z'])
                    if str(prm[0]) in ('X', 'Y', 'Z', 'N'):
                        if str(prm[0]) == basis:
                            part_measure = True
                        else:
                            part_measure = False
                            break
                    else:
                        part_measure = False
                        continue
            # The reference iterates the live level list and removes single
            # measures from it (dm_simulator.py:1020, 1100), so the op after a
            # removed one is skipped; an index walk reproduces that.
            i = 0
            while i < len(level):
                op = level[i]
                i += 1
                cond = getattr(op, 'conditional', None)
                if cond is not None:
                    raise BasicAerError('conditional operations are not supported by dm_simulator')
                name = op.name
                if name in ('u1', 'u3'):
                    self._gate_1q(st, op)
                elif name == 'cx':
                    self._gate_cx(st, op)
                elif name == 'reset':
                    st.apply_1q(op.qubits[0], ptm.rows3(ptm.reset()))
                elif name == 'barrier':
                    pass
                elif name == 'measure':
                    stop = self._measure(st, n, level, op, part_measure, data, meas_err, bell_err)
                    if stop:
                        break
                else:
                    raise BasicAerError('{0} encountered unrecognized operation "{1}"'.format(self.name(), name))
            if not (f == 1 and g == 1):
                st.memory_noise(f, p, g)
            clock += 1

    def _measure(self, st, n, level, op, part_measure, data, meas_err, bell_err):
        """One measure op of a level; returns True when the level is done
        (dm_simulator.py:1054-1136)."""
        prm = getattr(op, 'params', None)
        if prm is None:
            prm = ['Z']
            op.params = prm
        q = op.qubits[0]
        kind = str(prm[0])
        single = expect = ens = False
        if kind == 'Ensemble':
            ens = True
        elif kind == 'Expect':
            expect = True
        elif len(level) == 1:
            single = True
        elif part_measure is not True:
            single = True
        if len(prm) == 1:
            prm.append(None)
        if single:
            if kind in ('X', 'Y'):
                self._collapse(st, q, kind, None, meas_err)
            elif kind == 'N':
                prm[1] = self._unit(prm[1])
                self._collapse(st, q, 'N', prm[1], meas_err)
            elif kind == 'Bell':
                pair = str(prm[1])
                bp, red = self._bell(st, n, int(pair[0]), int(pair[1]), bell_err)
                data['bell_probabilities%s%s' % (pair[0], pair[1])] = bp
                data['reduced_bell_densitymatrix%s%s' % (pair[0], pair[1])] = red
            else:
                self._collapse(st, q, 'Z', None, meas_err)
            del level[next(j for j, x in enumerate(level) if x is op)]
            return False
        if part_measure:
            mqs = [x.qubits[0] for x in level]
            addp = prm[1] if kind == 'N' else None
            probs = self._probabilities(st, mqs, kind, addp, meas_err)
            data['partial_probability'] = probs
            for mq in mqs:
                if kind == 'N' and addp is not None:
                    self._collapse(st, mq, 'N', np.asarray(addp, dtype=float), meas_err)
                else:
                    self._collapse(st, mq, kind, None, meas_err)
            return True
        if expect:
            data['Pauli_string_expectation'] = self._expect(st, n, str(prm[1]), meas_err)
            return True
        if ens:
            spec = prm[1]
            if isinstance(spec, (tuple, list)) and len(spec) == 2 and str(spec[0]) == 'N':
                basis, addp = 'N', self._unit(spec[1])
            else:
                basis, addp = str(spec), None
            data['ensemble_probability'] = self._probabilities(st, list(range(n)), basis, addp, meas_err)
            return True
        return False


def describe_partition(levels):
    """Print the level table (dm_simulator.py:1284-1313)."""
    print('\nPARTITIONED CIRCUIT')
    for i, lv in enumerate(levels):
        print('\nPartition ', i)
        for inst in lv:
            name, qubit = inst.name, inst.qubits
            if name in ('u3', 'u1'):
                print(name.upper(), '   qubit', qubit, '    ', [round(float(x), 6) for x in inst.params])
            elif name == 'cx':
                print('C-NOT', '   qubit', qubit)
            elif name == 'measure':
                prm = getattr(inst, 'params', None) or ['Z']
                if str(prm[0]) == 'Bell':
                    print('Bell Measure', '   qubit', [int(x) for x in str(prm[1])])
                else:
                    print(name, '   qubit', qubit, '    ', prm)
