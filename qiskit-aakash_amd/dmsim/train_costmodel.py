"""Train the fusion cost model on GPU timings and compare its plans.

    python -m dmsim.train_costmodel [--qubits 16] [--max-specs 1200] [--out PATH]

1. draws candidate passes from seeded random circuits (seeds 1000-2999, the
   local U3/CX generator), times a sample of them on the GPU in FP32 and FP64;
2. trains the perceptron on the GPU (dmsim.costmodel.train) and saves it;
3. A/B: on held-out circuits (seeds 5000+) times the learned plan, the greedy
   plan (k <= 4, sequential tile passes), the same candidate search scored by
   gate count alone (the learned planner without its network) and plain
   per-gate application.
Writes a JSON report next to the parameters.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import costmodel, fusion, native, workloads


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_program(prog, n, precision, reps=2):
    st = native.DeviceState(n, precision)
    st.init_product(np.tile([1.0, 0.0, 0.0, 1.0], n), 0.5 ** n)
    prog.run(st)
    st.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        prog.run(st)
    st.synchronize()
    dt = (time.perf_counter() - t0) / reps
    st.close()
    return dt


def ab_compare(model, n=16, seeds=(5000, 5001, 5002, 5003, 5004), gates=400, precision='fp32'):
    """Measured seconds of learned vs greedy vs plain plans on held-out circuits."""
    out = []
    for s in seeds:
        noisy = fusion.noisy_gates(workloads.local_u3_cx(n, gates, s), 1e-3)
        model.precision = precision
        learned, lb = fusion.fused_program(noisy, n, fusion.KMAX['learned'], 'learned', model)
        greedy, gb = fusion.fused_program(noisy, n, fusion.KMAX['greedy'], 'greedy', None, 'seq')
        search, sb = fusion.fused_program(noisy, n, fusion.KMAX['search'], 'search', None, 'seq')
        plain = fusion.plain_program(noisy, n)
        row = {'seed': s, 'gates': gates, 'passes_learned': len(learned), 'passes_greedy': len(greedy),
               'passes_search': len(search),
               't_learned': time_program(learned, n, precision), 't_greedy': time_program(greedy, n, precision),
               't_search': time_program(search, n, precision), 't_plain': time_program(plain, n, precision, reps=1)}
        row['gain_vs_greedy'] = row['t_greedy'] / row['t_learned']
        row['gain_vs_search'] = row['t_search'] / row['t_learned']
        row['gain_vs_plain'] = row['t_plain'] / row['t_learned']
        log('A/B seed %d: %s' % (s, json.dumps(row)))
        out.append(row)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--qubits', type=int, default=16)
    ap.add_argument('--max-specs', type=int, default=1200)
    ap.add_argument('--epochs', type=int, default=400)
    ap.add_argument('--out', default=costmodel.PARAMS_FILE)
    ap.add_argument('--data', default=None, help='reuse a collected dataset (.npz)')
    ap.add_argument('--save-data', default=None)
    ap.add_argument('--ab-only', action='store_true', help='re-run the A/B comparison with the saved model')
    args = ap.parse_args()
    n = args.qubits
    if args.ab_only:
        model = costmodel.CostModel.load(args.out)
        meta = dict(model.meta)
        meta['ab'] = ab_compare(model, n)
        meta['time_weight'] = fusion.TIME_WEIGHT
        model.meta = meta
        model.save(args.out)
        with open(os.path.splitext(args.out)[0] + '_report.json', 'w') as f:
            json.dump(meta, f, indent=1)
        print(json.dumps(meta['ab']), flush=True)
        return
    t0 = time.perf_counter()
    if args.data:
        with np.load(args.data, allow_pickle=False) as z:
            X, y = z['X'], z['y']
        rows = []
    else:
        X, y, rows = costmodel.collect(n, range(1000, 3000), ('fp32', 'fp64'), args.max_specs, log=log)
        if args.save_data:
            np.savez(args.save_data, X=X, y=y)
    t_collect = time.perf_counter() - t0
    log('dataset: %d timed passes in %.0f s' % (len(y), t_collect))
    params, report = costmodel.train(X, y, epochs=args.epochs)
    log('training: %s' % json.dumps(report))
    meta = dict(report, n=n, features=costmodel.N_FEATURES, hidden=costmodel.HIDDEN,
                seeds='1000-2999', collect_seconds=t_collect, time_weight=fusion.TIME_WEIGHT)
    model = costmodel.CostModel(params, 'fp32', meta)
    model.save(args.out)
    ab = ab_compare(model, n)
    meta['ab'] = ab
    model.meta = meta
    model.save(args.out)
    with open(os.path.splitext(args.out)[0] + '_report.json', 'w') as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta), flush=True)


if __name__ == '__main__':
    main()
