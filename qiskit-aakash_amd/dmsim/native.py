"""ctypes binding of libdmsim.so (the C ABI declared in include/dmsim.h).

The binding is the only way the Python host touches device state; there is
no CPU execution path.  Loading fails loudly when the library is missing.
"""
import ctypes
import os
import threading

import numpy as np

from .exceptions import BasicAerError

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libdmsim.so')
_lib = None
_lock = threading.Lock()

FP64, FP32 = 0, 1
# return codes (DM_ERR_* of include/dmsim.h)
ERR_ARG, ERR_HIP, ERR_STATE = -1, -2, -3

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_u64_p = ctypes.POINTER(ctypes.c_uint64)
c_float_p = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes); every entry point of include/dmsim.h
SIGNATURES = {
    'dm_last_error': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    'dm_device_info': (ctypes.c_int, [ctypes.c_int, c_int_p, c_int_p, c_u64_p]),
    'dm_create': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    'dm_destroy': (ctypes.c_int, [ctypes.c_void_p]),
    'dm_num_qubits': (ctypes.c_int, [ctypes.c_void_p, c_int_p]),
    'dm_dtype': (ctypes.c_int, [ctypes.c_void_p, c_int_p]),
    'dm_device_ptr': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    'dm_init_product': (ctypes.c_int, [ctypes.c_void_p, c_double_p, ctypes.c_double]),
    'dm_upload': (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    'dm_download': (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    'dm_download_chop': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, c_double_p]),
    'dm_read_coeffs': (ctypes.c_int, [ctypes.c_void_p, c_u64_p, ctypes.c_int, c_double_p]),
    'dm_apply_1q': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_double_p]),
    'dm_apply_cx': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_double_p]),
    'dm_apply_2q': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_double_p]),
    'dm_memory_noise': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    'dm_apply_diag': (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    'dm_apply_block': (ctypes.c_int, [ctypes.c_void_p, c_int_p, ctypes.c_int, c_double_p]),
    'dm_probabilities': (ctypes.c_int, [ctypes.c_void_p, c_int_p, ctypes.c_int, ctypes.c_char,
                                        c_double_p, ctypes.c_double, c_double_p]),
    'dm_trace_purity': (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p]),
    'dm_overlap': (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p]),
    'dm_to_dense': (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    'dm_compare': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_double_p, c_double_p, c_double_p,
                                  c_double_p]),
    'dm_download_range': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                         ctypes.c_void_p]),
    'dm_upload_range': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                       ctypes.c_void_p]),
    'dm_dot_range': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.c_void_p, c_double_p]),
    'dm_merge_rotations': (ctypes.c_int, [ctypes.c_int, c_int_p, c_int_p, c_double_p, ctypes.c_int, c_int_p, c_int_p,
                                          c_double_p, c_int_p]),
    'dm_partition': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                    c_int_p]),
    'dm_plan_fusion': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_int_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      c_int_p, ctypes.c_int, c_float_p, ctypes.c_int, c_int_p, c_int_p, c_int_p,
                                      c_int_p, c_int_p]),
    'dm_mlp_predict_host': (ctypes.c_int, [c_float_p, ctypes.c_int, ctypes.c_int, c_float_p, ctypes.c_int, c_float_p]),
    'dm_cost_block_seconds': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_int_p, ctypes.c_int, c_int_p, c_int_p,
                                             c_int_p, c_int_p, c_float_p, ctypes.c_int, c_double_p]),
    'dm_run_program': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, c_double_p,
                                      ctypes.c_int64]),
    'dm_run_program_noalloc': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, c_double_p,
                                              ctypes.c_int64]),
    'dm_mlp_param_count': (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    'dm_mlp_init': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, c_float_p]),
    'dm_mlp_predict': (ctypes.c_int, [c_float_p, ctypes.c_int, ctypes.c_int, c_float_p, ctypes.c_int, c_float_p]),
    'dm_mlp_loss_grad': (ctypes.c_int, [c_float_p, ctypes.c_int, ctypes.c_int, c_float_p, c_float_p, ctypes.c_int,
                                        c_float_p, c_float_p]),
    'dm_mlp_train': (ctypes.c_int, [c_float_p, ctypes.c_int, ctypes.c_int, c_float_p, c_float_p, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, c_float_p]),
    'dm_batch_create': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_int_p, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.c_void_p, c_double_p, ctypes.c_int64, c_double_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    'dm_batch_info': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    'dm_batch_run': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_double_p, c_double_p, c_double_p]),
    'dm_batch_circuit_program': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_int_p, c_int_p, c_int_p, c_double_p,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, c_double_p,
                                                ctypes.c_int64, c_int_p, ctypes.POINTER(ctypes.c_int64)]),
    'dm_batch_destroy': (ctypes.c_int, [ctypes.c_void_p]),
    'dm_synchronize': (ctypes.c_int, [ctypes.c_void_p]),
    'dm_profile': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    'dm_profile_read': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_u64_p, c_double_p]),
    'dm_profile_reset': (ctypes.c_int, [ctypes.c_void_p]),
}

class DmOp(ctypes.Structure):
    """struct dm_op of include/dmsim.h."""
    _fields_ = [('kind', ctypes.c_int32), ('k', ctypes.c_int32), ('qubits', ctypes.c_int32 * 4),
                ('param', ctypes.c_int64)]


# the same record as a numpy structured type (views of DmOp arrays)
DMOP_DTYPE = np.dtype({'names': ['kind', 'k', 'qubits', 'param'], 'formats': ['<i4', '<i4', ('<i4', (4,)), '<i8'],
                       'offsets': [0, 4, 8, 24], 'itemsize': 32})
assert ctypes.sizeof(DmOp) == DMOP_DTYPE.itemsize

OP_1Q, OP_CX, OP_CX_SCALED, OP_2Q, OP_BLOCK, OP_MEMORY_NOISE, OP_DIAG, OP_FUSED = 1, 2, 3, 4, 5, 6, 7, 8
# sub-operation kinds of a fused pass
SUB_1Q, SUB_CX, SUB_2Q, SUB_DENSE = 1, 2, 3, 4

FAMILIES = {'1q': 0, 'cx': 1, '2q': 2, 'memory_noise': 3, 'diag': 4, 'block3': 5, 'reduce': 6, 'init': 7,
            'block4': 8, 'fused': 9}


class NativeError(BasicAerError, RuntimeError):
    """A failure reported by libdmsim (message from dm_last_error).  It is a
    BasicAerError, the error type the reference backend raises, so callers
    catching the reference's errors catch device failures (out of HBM, bad
    arguments) too."""


def library_path():
    return _LIB_PATH


def load():
    """Load libdmsim.so and declare its signatures (idempotent)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(_LIB_PATH):
            raise NativeError('libdmsim.so not built (expected at %s); build it with dmsim._build.build()' % _LIB_PATH)
        lib = ctypes.CDLL(_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error():
    buf = ctypes.create_string_buffer(512)
    load().dm_last_error(buf, len(buf))
    return buf.value.decode(errors='replace')


def check(rc, what):
    if rc != 0:
        raise NativeError('%s failed (%d): %s' % (what, rc, last_error()))


def _dp(arr):
    return arr.ctypes.data_as(c_double_p)


def _f64(x, size=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if size is not None and a.size != size:
        raise ValueError('expected %d values, got %d' % (size, a.size))
    return a


def device_info(device=0):
    lib = load()
    nd, cu, hbm = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_uint64(0)
    check(lib.dm_device_info(device, ctypes.byref(nd), ctypes.byref(cu), ctypes.byref(hbm)), 'dm_device_info')
    return {'devices': nd.value, 'cu_count': cu.value, 'hbm_bytes': hbm.value}


class DeviceState:
    """An n-qubit Pauli-coefficient vector resident on the GPU."""

    def __init__(self, n_qubits, precision='fp64'):
        self.lib = load()
        self.n = int(n_qubits)
        self.dtype = FP32 if precision in ('fp32', FP32, 'float32') else FP64
        h = ctypes.c_void_p()
        check(self.lib.dm_create(self.n, self.dtype, ctypes.byref(h)), 'dm_create(%d qubits)' % self.n)
        self.h = h

    # lifecycle ------------------------------------------------------------
    def close(self):
        if getattr(self, 'h', None):
            self.lib.dm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def size(self):
        return 4 ** self.n

    # state I/O --------------------------------------------------------------
    def init_product(self, per_qubit4, scale=1.0):
        v = _f64(per_qubit4, 4 * self.n)
        check(self.lib.dm_init_product(self.h, _dp(v), float(scale)), 'dm_init_product')

    def upload(self, coeffs):
        a = _f64(coeffs, self.size)
        check(self.lib.dm_upload(self.h, _dp(a)), 'dm_upload')

    def download(self, chop=None, out=None):
        if out is None:
            out = np.empty(self.size, dtype=np.float64)
        if chop:
            check(self.lib.dm_download_chop(self.h, float(chop), _dp(out)), 'dm_download_chop')
        else:
            check(self.lib.dm_download(self.h, _dp(out)), 'dm_download')
        return out

    def _host_range(self, arr, offset):
        a = np.asarray(arr)
        if a.dtype not in (np.float64, np.float32) or a.ndim != 1 or not a.flags.c_contiguous:
            raise ValueError('host range must be a contiguous 1-d float64 or float32 array')
        return a, (FP64 if a.dtype == np.float64 else FP32), ctypes.c_uint64(int(offset)), ctypes.c_uint64(a.size)

    def download_range(self, offset, out):
        """Coefficients [offset, offset + out.size) into ``out`` (float64 or
        float32, converted on the device)."""
        a, hd, off, cnt = self._host_range(out, offset)
        if not a.flags.writeable:
            raise ValueError('output array is read-only')
        check(self.lib.dm_download_range(self.h, off, cnt, hd, a.ctypes.data_as(ctypes.c_void_p)),
              'dm_download_range')
        return out

    def upload_range(self, offset, values):
        a, hd, off, cnt = self._host_range(values, offset)
        check(self.lib.dm_upload_range(self.h, off, cnt, hd, a.ctypes.data_as(ctypes.c_void_p)), 'dm_upload_range')

    def dot_range(self, offset, values):
        """FP64 sum of a_i * values[i - offset] over the range."""
        a, hd, off, cnt = self._host_range(values, offset)
        out = ctypes.c_double(0)
        check(self.lib.dm_dot_range(self.h, off, cnt, hd, a.ctypes.data_as(ctypes.c_void_p), ctypes.byref(out)),
              'dm_dot_range')
        return out.value

    def read(self, indices):
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64))
        out = np.empty(idx.size, dtype=np.float64)
        check(self.lib.dm_read_coeffs(self.h, idx.ctypes.data_as(c_u64_p), idx.size, _dp(out)), 'dm_read_coeffs')
        return out

    # updates ----------------------------------------------------------------
    def apply_1q(self, q, rows3x4):
        R = _f64(rows3x4, 12)
        check(self.lib.dm_apply_1q(self.h, int(q), _dp(R)), 'dm_apply_1q')

    def apply_cx(self, control, target, scale16=None):
        if scale16 is None:
            check(self.lib.dm_apply_cx(self.h, int(control), int(target), None), 'dm_apply_cx')
        else:
            s = _f64(scale16, 16)
            check(self.lib.dm_apply_cx(self.h, int(control), int(target), _dp(s)), 'dm_apply_cx')

    def apply_2q(self, qa, qb, R16x16):
        R = _f64(R16x16, 256)
        check(self.lib.dm_apply_2q(self.h, int(qa), int(qb), _dp(R)), 'dm_apply_2q')

    def memory_noise(self, f, p, g):
        check(self.lib.dm_memory_noise(self.h, float(f), float(p), float(g)), 'dm_memory_noise')

    def apply_diag(self, tab_nx4):
        t = _f64(tab_nx4, 4 * self.n)
        check(self.lib.dm_apply_diag(self.h, _dp(t)), 'dm_apply_diag')

    def apply_block(self, qubits, S):
        qs = np.ascontiguousarray(np.asarray(qubits, dtype=np.int32))
        k = qs.size
        M = _f64(S, 4 ** (2 * k))
        check(self.lib.dm_apply_block(self.h, qs.ctypes.data_as(c_int_p), k, _dp(M)), 'dm_apply_block')

    # reductions -------------------------------------------------------------
    def probabilities(self, qubits, basis='Z', err=1.0, nvec=None):
        qs = np.ascontiguousarray(np.asarray(qubits, dtype=np.int32))
        out = np.empty(2 ** qs.size, dtype=np.float64)
        nv = _f64(nvec if nvec is not None else [0.0, 0.0, 1.0], 3)
        check(self.lib.dm_probabilities(self.h, qs.ctypes.data_as(c_int_p), qs.size, basis.encode()[:1],
                                        _dp(nv), float(err), _dp(out)), 'dm_probabilities')
        return out

    def trace_purity(self):
        tr, pur = ctypes.c_double(0), ctypes.c_double(0)
        check(self.lib.dm_trace_purity(self.h, ctypes.byref(tr), ctypes.byref(pur)), 'dm_trace_purity')
        return tr.value, pur.value

    def overlap(self, other_coeffs):
        b = _f64(other_coeffs, self.size)
        out = ctypes.c_double(0)
        check(self.lib.dm_overlap(self.h, _dp(b), ctypes.byref(out)), 'dm_overlap')
        return out.value

    def compare(self, other):
        """FP64 device comparison with another state of the same register."""
        m, d, a2, b2 = (ctypes.c_double(0) for _ in range(4))
        check(self.lib.dm_compare(self.h, other.h, ctypes.byref(m), ctypes.byref(d), ctypes.byref(a2),
                                  ctypes.byref(b2)), 'dm_compare')
        return {'max_abs_diff': m.value, 'dot': d.value, 'norm2_self': a2.value, 'norm2_other': b2.value}

    def to_dense(self):
        dim = 2 ** self.n
        out = np.empty((dim, dim), dtype=np.complex128)
        check(self.lib.dm_to_dense(self.h, out.ctypes.data_as(c_double_p)), 'dm_to_dense')
        return out

    def run_program(self, ops, params):
        """``ops``: ctypes array of DmOp; ``params``: float64 array."""
        pv = _f64(params) if params is not None and len(params) else np.zeros(1)
        check(self.lib.dm_run_program(self.h, ctypes.cast(ops, ctypes.c_void_p), len(ops), _dp(pv),
                                      int(len(params) if params is not None else 0)), 'dm_run_program')

    def synchronize(self):
        check(self.lib.dm_synchronize(self.h), 'dm_synchronize')

    # timing -----------------------------------------------------------------
    def profile(self, enable=True):
        check(self.lib.dm_profile(self.h, 1 if enable else 0), 'dm_profile')

    def profile_reset(self):
        check(self.lib.dm_profile_reset(self.h), 'dm_profile_reset')

    def profile_read(self, family):
        n, ms = ctypes.c_uint64(0), ctypes.c_double(0)
        check(self.lib.dm_profile_read(self.h, FAMILIES[family], ctypes.byref(n), ctypes.byref(ms)),
              'dm_profile_read')
        return n.value, ms.value


# ---------------------------------------------------------------------------
# cost-model perceptron
# ---------------------------------------------------------------------------

def _fp(a):
    return a.ctypes.data_as(c_float_p)


def _f32(x, size=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    if size is not None and a.size != size:
        raise ValueError('expected %d values, got %d' % (size, a.size))
    return a


def mlp_param_count(n_features, hidden):
    c = load().dm_mlp_param_count(int(n_features), int(hidden))
    check(c if c < 0 else 0, 'dm_mlp_param_count')
    return c


def mlp_init(n_features, hidden, seed):
    p = np.zeros(mlp_param_count(n_features, hidden), dtype=np.float32)
    check(load().dm_mlp_init(int(n_features), int(hidden), int(seed), _fp(p)), 'dm_mlp_init')
    return p


def mlp_predict(params, n_features, hidden, X):
    X = _f32(X)
    n = X.shape[0] if X.ndim == 2 else 0
    out = np.zeros(max(n, 1), dtype=np.float32)
    check(load().dm_mlp_predict(_fp(_f32(params)), int(n_features), int(hidden), _fp(X), n, _fp(out)),
          'dm_mlp_predict')
    return out[:n]


def mlp_loss_grad(params, n_features, hidden, X, y):
    p = _f32(params)
    X = _f32(X)
    y = _f32(y)
    loss = np.zeros(1, dtype=np.float32)
    grad = np.zeros(p.size, dtype=np.float32)
    check(load().dm_mlp_loss_grad(_fp(p), int(n_features), int(hidden), _fp(X), _fp(y), y.size, _fp(loss),
                                  _fp(grad)), 'dm_mlp_loss_grad')
    return float(loss[0]), grad


def mlp_train(params, n_features, hidden, X, y, epochs, batch, lr, seed):
    p = _f32(params).copy()
    X = _f32(X)
    y = _f32(y)
    losses = np.zeros(max(1, epochs), dtype=np.float32)
    check(load().dm_mlp_train(_fp(p), int(n_features), int(hidden), _fp(X), _fp(y), y.size, int(epochs),
                              int(batch), float(lr), int(seed), _fp(losses)), 'dm_mlp_train')
    return p, losses[:epochs]


def mlp_predict_host(params, n_features, hidden, X):
    """The perceptron's forward pass on the host (the planner's predictions)."""
    X = _f32(X)
    n = X.shape[0] if X.ndim == 2 else 0
    out = np.zeros(max(n, 1), dtype=np.float32)
    check(load().dm_mlp_predict_host(_fp(_f32(params)), int(n_features), int(hidden), _fp(X), n, _fp(out)),
          'dm_mlp_predict_host')
    return out[:n]


# ---------------------------------------------------------------------------
# host front end: merge, partition, fusion planning
# ---------------------------------------------------------------------------

INST_U1, INST_U3, INST_ID, INST_SEPARATOR = 1, 2, 3, 4
INST_ROT, INST_CX, INST_MEASURE, INST_RESET, INST_BARRIER, INST_OWN_LEVEL = 5, 6, 7, 8, 9, 10
PLAN_METHODS = {'greedy': 0, 'search': 1, 'learned': 2}


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _ip(a):
    return a.ctypes.data_as(c_int_p)


def merge_rotations(kind, qubit, params, merge=True):
    """(src, kind, params[k, 3]) of the merged instruction list."""
    kind, qubit = _i32(kind), _i32(qubit)
    p = _f64(np.asarray(params, dtype=np.float64).reshape(-1))
    m = kind.size
    src = np.zeros(max(m, 1), dtype=np.int32)
    okind = np.zeros(max(m, 1), dtype=np.int32)
    op = np.zeros(3 * max(m, 1))
    nout = ctypes.c_int(0)
    check(load().dm_merge_rotations(m, _ip(kind), _ip(qubit), _dp(p), 1 if merge else 0, _ip(src), _ip(okind),
                                    _dp(op), ctypes.byref(nout)), 'dm_merge_rotations')
    k = nout.value
    return src[:k], okind[:k], op[:3 * k].reshape(k, 3)


def partition_levels(n, kind, q0, q1):
    """(levels as lists of instruction indices, count)."""
    kind, q0, q1 = _i32(kind), _i32(q0), _i32(q1)
    m = kind.size
    ops = np.zeros(max(m, 1), dtype=np.int32)
    starts = np.zeros(4 * m + 4, dtype=np.int32)
    nl, cnt = ctypes.c_int(0), ctypes.c_int(0)
    check(load().dm_partition(int(n), m, _ip(kind), _ip(q0), _ip(q1), _ip(ops), _ip(starts), ctypes.byref(nl),
                              ctypes.byref(cnt)), 'dm_partition')
    levels = [ops[starts[i]:starts[i + 1]].tolist() for i in range(nl.value)]
    return levels, cnt.value


def plan_fusion(n, gate_qubits, method, kmax, tiles, lookahead=(), params=None, precision='fp32'):
    """Blocks (qubits, members) of the fusion planner."""
    gq = _i32(gate_qubits).reshape(-1)
    ng = gq.size // 2
    la = _i32(list(lookahead) or [0])
    p = _f32(params) if params is not None else None
    members = np.zeros(max(ng, 1), dtype=np.int32)
    bstart = np.zeros(ng + 1, dtype=np.int32)
    qubits = np.zeros(max(2 * ng, 1), dtype=np.int32)
    qstart = np.zeros(ng + 1, dtype=np.int32)
    nb = ctypes.c_int(0)
    check(load().dm_plan_fusion(int(n), ng, _ip(gq), PLAN_METHODS[method], int(kmax), 1 if tiles else 0, _ip(la),
                                len(lookahead), _fp(p) if p is not None else None,
                                FP64 if precision == 'fp64' else FP32, _ip(members), _ip(bstart), _ip(qubits),
                                _ip(qstart), ctypes.byref(nb)), 'dm_plan_fusion')
    return [(qubits[qstart[b]:qstart[b + 1]].tolist(), members[bstart[b]:bstart[b + 1]].tolist())
            for b in range(nb.value)]


def batch_circuit_program(n, gates, decay_factor=1.0, decoherence_factor=1.0, thermal_factor=1.0,
                          gate_depolarization=0.0, fused=False, precision='fp64'):
    """(DmOp array, values) of one batch circuit from dm_batch_circuit_program:
    gates (name, qubits, params) of u1 / u2 / u3 / cx, each followed by a
    memory-noise level; ``fused`` gives the tile-pass form."""
    m = len(gates)
    kind, q0, q1, ang = [], [], [], []
    nprm = {'cx': 0, 'u3': 3, 'U': 3, 'u2': 2, 'u1': 1}
    for name, qs, prm in gates:
        if name not in nprm:
            raise ValueError('batch circuits hold u1/u2/u3/cx gates, got %s' % name)
        if len(prm) != nprm[name] or len(qs) != (2 if name == 'cx' else 1):
            raise ValueError('%s takes %d qubit(s) and %d parameter(s), got %r, %r'
                             % (name, 2 if name == 'cx' else 1, nprm[name], qs, prm))
        q0.append(qs[0])
        if name == 'cx':
            kind.append(INST_CX)
            q1.append(qs[1])
            ang += (0.0, 0.0, 0.0)
            continue
        q1.append(0)
        if name in ('u3', 'U'):
            kind.append(INST_U3)
            ang += prm[:3]
        elif name == 'u2':
            kind.append(INST_U3)
            ang += (np.pi / 2, prm[0], prm[1])
        else:
            kind.append(INST_U1)
            ang += (prm[0], 0.0, 0.0)
    kind, q0, q1 = _i32(kind or [0]), _i32(q0 or [0]), _i32(q1 or [0])
    ang = np.asarray(ang or [0.0], dtype=np.float64)
    lib = load()
    nops, nvals = ctypes.c_int(0), ctypes.c_int64(0)
    head = (int(n), m, _ip(kind), _ip(q0), _ip(q1), _dp(ang), float(decay_factor), float(decoherence_factor),
            float(thermal_factor), float(gate_depolarization), 1 if fused else 0,
            FP32 if precision == 'fp32' else FP64)
    # capacity for the largest possible program (one record per noisy gate,
    # each with its 7-value header and 16 values, plus block headers); the
    # call reports the sizes, and a second call follows if they do not fit
    cap_ops = max(1, m * (n + 1 if fused else 2))
    cap_vals = max(1, cap_ops * 30)
    for _ in range(2):
        ops = (DmOp * cap_ops)()
        vals = np.empty(cap_vals)
        rc = lib.dm_batch_circuit_program(*head, ctypes.cast(ops, ctypes.c_void_p), cap_ops, _dp(vals), cap_vals,
                                          ctypes.byref(nops), ctypes.byref(nvals))
        if rc == 0 or (nops.value <= cap_ops and nvals.value <= cap_vals):
            break
        cap_ops, cap_vals = max(1, nops.value), max(1, nvals.value)
    check(rc, 'dm_batch_circuit_program')
    # exact-size copies: the capacity bound is many times the usual program
    k = nops.value
    exact = (DmOp * max(1, k))()
    ctypes.memmove(exact, ops, k * ctypes.sizeof(DmOp))
    return exact, k, vals[:nvals.value].copy()


def cost_block_seconds(n, gate_qubits, blocks, params, precision='fp32'):
    """The cost model's predicted seconds of each (qubits, members) block."""
    gq = _i32(gate_qubits).reshape(-1)
    members = _i32([i for _, mem in blocks for i in mem] or [0])
    qubits = _i32([q for qs, _ in blocks for q in qs] or [0])
    bstart = _i32(np.cumsum([0] + [len(mem) for _, mem in blocks]))
    qstart = _i32(np.cumsum([0] + [len(qs) for qs, _ in blocks]))
    out = np.zeros(max(len(blocks), 1))
    check(load().dm_cost_block_seconds(int(n), gq.size // 2, _ip(gq), len(blocks), _ip(members), _ip(bstart),
                                       _ip(qubits), _ip(qstart), _fp(_f32(params)),
                                       FP64 if precision == 'fp64' else FP32, _dp(out)), 'dm_cost_block_seconds')
    return out[:len(blocks)]
