"""Learned cost model for gate fusion.

A 3-layer perceptron (64 hidden units; forward, backward and Adam kernels in
csrc/mlp.hip) predicts the GPU time of one pass over the state -- a plain
gate kernel, a sequential tile pass or a dense (matrix-core) tile pass --
from the pass's features:

  kind (one-hot: plain 1q, plain cx, tile sequence, tile dense), number of
  block qubits k, counts of 1- and 2-qubit gates in the pass, the length of
  the contiguous run the tile keeps in HBM (4^w coefficients), storage
  precision, register size, lowest / highest block position, a bitmask of
  the block's positions (their strides 4^p), the number of 2-qubit gates on
  a qubit at the three lowest positions (they move the tile's LDS layout)
  and the number of distinct qubits the 2-qubit gates use.

The target is ln(time / state bytes), so one network covers register sizes.
Training data are GPU timings of passes drawn from the candidate blocks of
seeded random circuits (``collect``).  The fusion planner (dmsim.fusion,
method 'learned') asks ``block_times`` for the candidate blocks of each step
and keeps the one with the most gates per predicted second; ``best_mode``
picks the cheaper tile form.
"""
import json
import math
import os
import time

import numpy as np

from . import fusion, native, workloads

N_FEATURES = 34
HIDDEN = 64
MAX_POS = 20
TILE = 6
PARAMS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'costmodel_params.npz')

KINDS = ('plain1q', 'plaincx', 'seq', 'dense')


def tile_run_digits(n, positions, precision='fp32'):
    """w of the tile a sequential pass over these positions uses."""
    return fusion.run_digits(n, positions, fusion.tile_for(n, [n - 1 - p for p in positions], precision))


def gate_stats(n, gates):
    """(n1, n2, n2low, d2) of a pass's gates: 1- and 2-qubit gate counts, the
    2-qubit gates with a qubit at the three lowest positions, and the distinct
    qubits of the 2-qubit gates."""
    low = fusion.low_qubits(n)
    two = [g for g in gates if len(g.qubits) == 2]
    return (len(gates) - len(two), len(two), sum(1 for g in two if low.intersection(g.qubits)),
            len({q for g in two for q in g.qubits}))


def features(kind, n, qubits, n1, n2, precision, n2low=0, d2=0):
    """Feature vector (length N_FEATURES) of one pass."""
    pos = sorted(n - 1 - q for q in qubits)
    x = np.zeros(N_FEATURES, dtype=np.float32)
    x[KINDS.index(kind)] = 1.0
    x[4] = len(pos) / 6.0
    x[5] = math.log2(1 + n1) / 4.0
    x[6] = math.log2(1 + n2) / 4.0
    if kind == 'seq':
        w = tile_run_digits(n, pos, precision)
    elif kind == 'dense':
        w = fusion.run_digits(n, pos, 6)
    else:
        w = 1 if min(pos) > 0 else 0
    x[7] = w / 6.0
    x[8] = 1.0 if precision == 'fp64' else 0.0
    x[9] = n / 20.0
    x[10] = pos[0] / n
    x[11] = pos[-1] / n
    for p in pos:
        if p < MAX_POS:
            x[12 + p] = 1.0
    x[32] = math.log2(1 + n2low) / 4.0
    x[33] = d2 / 7.0
    return x


def state_gib(n, precision):
    return 4 ** n * (8 if precision == 'fp64' else 4) / 2 ** 30


def pass_kind(qs, gates, mode):
    """Kind of pass a block runs as: one qubit -> the plain 1q kernel (gates
    composed); one two-qubit gate -> the plain CX kernel; else ``mode``."""
    if len(qs) == 1:
        return 'plain1q'
    if len(gates) == 1:
        return 'plaincx'
    return mode


class CostModel:
    """Trained perceptron + a cache of predictions."""

    def __init__(self, params, precision='fp32', meta=None):
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        self.precision = precision
        self.meta = meta or {}
        self._cache = {}

    @classmethod
    def load(cls, path=PARAMS_FILE, precision='fp32'):
        with np.load(path, allow_pickle=False) as z:
            params = z['params']
            meta = json.loads(str(z['meta'])) if 'meta' in z else {}
        return cls(params, precision, meta)

    def save(self, path=PARAMS_FILE):
        np.savez(path, params=self.params, meta=json.dumps(self.meta))

    def predict_many(self, rows, n, precision):
        """Seconds for each (kind, qubits, n1, n2[, n2low, d2]) row: the host
        forward pass and the C library's exp, as the native planner
        evaluates them (dm_plan_fusion), so both plan alike."""
        X = np.stack([features(r[0], n, r[1], r[2], r[3], precision, *r[4:]) for r in rows])
        y = native.mlp_predict_host(self.params, N_FEATURES, HIDDEN, X)
        gib = state_gib(n, precision)
        return np.array([math.exp(float(v)) * gib for v in y])

    def _time(self, kind, n, qs, st):
        key = (kind, n, tuple(sorted(qs))) + tuple(st) + (self.precision,)
        t = self._cache.get(key)
        if t is None:
            t = float(self.predict_many([(kind, qs) + tuple(st)], n, self.precision)[0])
            self._cache[key] = t
        return t

    def best_mode(self, n, qs, gates):
        st = gate_stats(n, gates)
        if len(qs) < 2 or n < TILE:
            return 'seq'
        ts = self._time('seq', n, qs, st)
        td = self._time('dense', n, qs, st) if len(qs) <= 4 else float('inf')
        return 'seq' if ts <= td else 'dense'

    def block_times(self, n, blocks):
        """Predicted seconds of each (qubits, gates) block (one batched
        prediction for the rows not cached yet)."""
        rows, keys = [], []
        for qs, gates in blocks:
            st = gate_stats(n, gates)
            kind = pass_kind(qs, gates, 'seq')
            want = [kind] if kind != 'seq' else (['seq', 'dense'] if 2 <= len(qs) <= 4 else ['seq'])
            ks = []
            for kd in want:
                key = (kd, n, tuple(sorted(qs))) + st + (self.precision,)
                ks.append(key)
                if key not in self._cache:
                    rows.append((kd, qs) + st)
            keys.append(ks)
        if rows:
            uniq = {}
            for r in rows:
                uniq.setdefault((r[0], n, tuple(sorted(r[1]))) + tuple(r[2:]) + (self.precision,), r)
            t = self.predict_many(list(uniq.values()), n, self.precision)
            for key, tv in zip(uniq.keys(), t):
                self._cache[key] = float(tv)
        return [min(self._cache[k] for k in ks) for ks in keys]

    def block_time(self, n, qs, gates):
        st = gate_stats(n, gates)
        kind = pass_kind(qs, gates, 'seq')
        if kind != 'seq':
            return self._time(kind, n, qs, st)
        t = self._time('seq', n, qs, st)
        if 2 <= len(qs) <= 4:
            t = min(t, self._time('dense', n, qs, st))
        return t


_default = None


def default_model():
    """The trained model shipped with the package (raises if absent)."""
    global _default
    if _default is None:
        if not os.path.exists(PARAMS_FILE):
            raise FileNotFoundError('no trained cost model at %s' % PARAMS_FILE)
        _default = CostModel.load()
    return _default


# ---------------------------------------------------------------------------
# data collection and training (run on the GPU)
# ---------------------------------------------------------------------------

def candidate_passes(seeds, n, gates_per_circuit=200, kmax=4, precision='fp32'):
    """Distinct pass specs (kind, qubits, gate list) drawn from the candidate
    blocks the planners consider (fusion.candidate_blocks) on seeded random
    circuits; each circuit advances by a randomly chosen candidate."""
    specs = {}
    for seed in seeds:
        noisy = fusion.noisy_gates(workloads.local_u3_cx(n, gates_per_circuit, seed), 1e-3)
        done = [False] * len(noisy)
        first = 0
        rng = np.random.default_rng(seed)
        while first < len(noisy):
            while first < len(noisy) and done[first]:
                first += 1
            if first >= len(noisy):
                break
            options = fusion.candidate_blocks(noisy, done, first, kmax, n, n >= fusion.MIN_TILE_QUBITS,
                                              precision=precision)
            for qs, mem in options:
                gs = [noisy[i] for i in mem]
                modes = ('seq', 'dense') if 2 <= len(qs) <= 4 and len(mem) > 1 else ('seq',)
                for mode in modes:
                    kind = pass_kind(qs, gs, mode)
                    key = (kind, tuple(sorted(qs))) + gate_stats(n, gs)
                    specs.setdefault(key, (kind, qs, gs))
            qs, mem = options[int(rng.integers(len(options)))]
            for i in mem:
                done[i] = True
    return list(specs.values())


def time_pass(st, kind, qs, gates, reps=3):
    """GPU seconds of one pass (HIP events on the state's stream)."""
    from .program import Program
    prog = Program(st.n)
    mode = 'dense' if kind == 'dense' else 'seq'
    idx = list(range(len(gates)))
    for _ in range(reps):
        fusion.add_block(prog, gates, qs, idx, mode, precision='fp64' if st.dtype == native.FP64 else 'fp32')
    prog.run(st)  # warm-up (and upload of the pass's matrices)
    st.synchronize()
    st.profile(True)
    st.profile_reset()
    prog.run(st)
    total = 0.0
    for fam in native.FAMILIES:
        launches, ms = st.profile_read(fam)
        total += ms
    st.profile(False)
    return total / 1e3 / reps


def collect(n=16, seeds=range(1000, 3000), precisions=('fp32', 'fp64'), max_specs=1200, seed=0, log=None):
    """Time a sample of candidate passes; returns (X, y, rows)."""
    specs = candidate_passes(seeds, n)
    rng = np.random.default_rng(seed)
    if len(specs) > max_specs:
        # stratified by (kind, block size): rare kinds (plain passes) are kept whole
        groups = {}
        for sp in specs:
            groups.setdefault((sp[0], len(sp[1])), []).append(sp)
        keys = sorted(groups, key=lambda k: len(groups[k]))
        left, out = max_specs, []
        for i, k in enumerate(keys):
            quota = left // (len(keys) - i)
            g = groups[k]
            take = min(quota, len(g))
            pick = rng.choice(len(g), take, replace=False)
            out += [g[j] for j in sorted(pick)]
            left -= take
        specs = out
    X, y, rows = [], [], []
    for prec in precisions:
        st = native.DeviceState(n, prec)
        st.init_product(np.tile([1.0, 0.0, 0.0, 1.0], n), 0.5 ** n)
        t0 = time.perf_counter()
        for i, (kind, qs, gs) in enumerate(specs):
            t = time_pass(st, kind, qs, gs)
            n1, n2, n2low, d2 = gate_stats(n, gs)
            X.append(features(kind, n, qs, n1, n2, prec, n2low, d2))
            y.append(math.log(t / state_gib(n, prec)))
            rows.append({'kind': kind, 'qubits': list(qs), 'n1': n1, 'n2': n2, 'n2low': n2low, 'd2': d2,
                         'precision': prec, 'n': n, 'seconds': t})
            if log and (i + 1) % 200 == 0:
                log('collected %d/%d %s passes (%.0f s)' % (i + 1, len(specs), prec, time.perf_counter() - t0))
        st.close()
    return np.stack(X), np.asarray(y, dtype=np.float32), rows


def train(X, y, epochs=400, batch=64, lr=2e-3, seed=1, holdout=0.15):
    """Train on the GPU; returns (params, report)."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(y))
    nh = max(1, int(holdout * len(y)))
    te, tr = idx[:nh], idx[nh:]
    p0 = native.mlp_init(N_FEATURES, HIDDEN, seed)
    p, curve = native.mlp_train(p0, N_FEATURES, HIDDEN, X[tr], y[tr], epochs, batch, lr, seed)
    pred = native.mlp_predict(p, N_FEATURES, HIDDEN, X[te])
    resid = pred - y[te]
    r2 = 1.0 - float(np.sum(resid ** 2) / np.sum((y[te] - y[te].mean()) ** 2))
    report = {'train_loss_first': float(curve[0]), 'train_loss_last': float(curve[-1]),
              'holdout_rmse_log': float(np.sqrt(np.mean(resid ** 2))), 'holdout_r2': r2,
              'holdout_median_abs_rel_err': float(np.median(np.abs(np.expm1(resid)))),
              'samples': int(len(y)), 'holdout': int(nh)}
    return p, report
