"""A C host of the library, on the CPU: a C99 program that includes
include/dmsim.h and links -ldmsim builds a batch circuit program through the
C ABI (size query, then the records), reads an error message through
dm_last_error, and plans a gate list with dm_plan_fusion.  Its output must
equal what the Python binding gets from the same entry points (host-only
calls: no device is used)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from dmsim import batch, fusion, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

C_SRC = r'''
#include <stdio.h>
#include "dmsim.h"

int main(void) {
  /* u3(0.3, 0.2, 0.1) on qubit 0, cx 0 -> 1, u1(0.7) on qubit 1; 6 qubits */
  int32_t kind[3] = {DM_INST_U3, DM_INST_CX, DM_INST_U1};
  int32_t q0[3] = {0, 0, 1}, q1[3] = {0, 1, 0};
  double ang[9] = {0.3, 0.2, 0.1, 0, 0, 0, 0.7, 0, 0};
  int nops;
  int64_t nv;
  if (dm_batch_circuit_program(6, 3, kind, q0, q1, ang, 0.99, 1.0, 1.0, 0.0, 1, DM_FP64, NULL, 0, NULL, 0,
                               &nops, &nv)) return 1;
  static dm_op ops[64];
  static double vals[4096];
  if (nops > 64 || nv > 4096) return 2;
  if (dm_batch_circuit_program(6, 3, kind, q0, q1, ang, 0.99, 1.0, 1.0, 0.0, 1, DM_FP64, ops, 64, vals, 4096,
                               &nops, &nv)) return 3;
  printf("%d %lld\n", nops, (long long)nv);
  for (int i = 0; i < nops; ++i)
    printf("%d %d %d %d %d %d %lld\n", ops[i].kind, ops[i].k, ops[i].qubits[0], ops[i].qubits[1], ops[i].qubits[2],
           ops[i].qubits[3], (long long)ops[i].param);
  for (int64_t j = 0; j < nv; ++j) printf("%.17g\n", vals[j]);

  /* an error: fused programs need 6 qubits or more */
  char msg[256];
  if (dm_batch_circuit_program(5, 3, kind, q0, q1, ang, 1.0, 1.0, 1.0, 0.0, 1, DM_FP64, NULL, 0, NULL, 0, &nops,
                               &nv) != DM_ERR_ARG) return 4;
  dm_last_error(msg, sizeof msg);
  printf("%s\n", msg);

  /* greedy plan of: cx(0,1) u3(2) cx(1,2) u3(0) on 6 qubits, blocks of <= 2 qubits */
  int32_t gq[8] = {0, 1, 2, -1, 1, 2, 0, -1};
  int32_t members[4], bstart[5], bq[8], qstart[5];
  int nb;
  if (dm_plan_fusion(6, 4, gq, DM_PLAN_GREEDY, 2, 0, NULL, 0, NULL, DM_FP64, members, bstart, bq, qstart, &nb))
    return 5;
  printf("%d\n", nb);
  for (int b = 0; b < nb; ++b) {
    for (int j = bstart[b]; j < bstart[b + 1]; ++j) printf("%d ", members[j]);
    printf("|");
    for (int j = qstart[b]; j < qstart[b + 1]; ++j) printf(" %d", bq[j]);
    printf("\n");
  }
  return 0;
}
'''


@pytest.fixture(scope='module')
def host_output(tmp_path_factory):
    if shutil.which('gcc') is None:
        pytest.skip('no C compiler')
    d = tmp_path_factory.mktemp('c_host')
    src = d / 'host.c'
    src.write_text(C_SRC)
    libdir = os.path.dirname(native.library_path())
    exe = d / 'host'
    subprocess.run(['gcc', '-std=c99', '-Wall', '-Wextra', '-Werror', '-I', os.path.join(ROOT, 'include'),
                    str(src), '-L', libdir, '-ldmsim', '-Wl,-rpath,' + libdir, '-o', str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, (res.returncode, res.stderr)
    return res.stdout.splitlines()


def test_c_host_builds_the_same_batch_program(host_output):
    gates = [('u3', (0,), (0.3, 0.2, 0.1)), ('cx', (0, 1), ()), ('u1', (1,), (0.7,))]
    prog = batch.native_program(6, gates, decay_factor=0.99, fused=True)
    ops, vals = prog.compile()
    nops, nv = (int(x) for x in host_output[0].split())
    assert (nops, nv) == (len(prog), vals.size)
    for i in range(nops):
        got = [int(x) for x in host_output[1 + i].split()]
        o = ops[i]
        assert got == [o.kind, o.k] + list(o.qubits) + [o.param]
    got_vals = np.array([float(x) for x in host_output[1 + nops:1 + nops + nv]])
    assert np.array_equal(got_vals, vals)   # %.17g round-trips a double exactly


def test_c_host_reads_errors_and_plans(host_output):
    gates = [fusion.NoisyGate((0, 1), None, 'cx'), fusion.NoisyGate((2,), np.eye(4), 'u3'),
             fusion.NoisyGate((1, 2), None, 'cx'), fusion.NoisyGate((0,), np.eye(4), 'u3')]
    rest = host_output[-1 - len(fusion.plan(gates, 6, 2, 'greedy', tiles=False)) - 1:]
    assert 'n >= 6' in rest[0]
    blocks = fusion.plan(gates, 6, 2, 'greedy', tiles=False)
    assert int(rest[1]) == len(blocks)
    for line, (qs, mem) in zip(rest[2:], blocks):
        m, q = line.split('|')
        assert [int(x) for x in m.split()] == mem and [int(x) for x in q.split()] == qs
