"""dm_run_program_noalloc against dm_run_program, through the C ABI.

Both entry points enqueue the same kernels on the same program image, so the
states must agree bit for bit.  Two calls in a row on the shared scratch
array, the second with a smaller register and program (the array keeps the
first call's image; the second must not see it), then an image above the
array's size, which must be refused before any work.
"""
import ctypes

import numpy as np
import pytest

from dmsim import fusion, native, ptm, workloads
from dmsim.program import Program

pytestmark = pytest.mark.gpu


def program(n, gates, seed):
    """A fused program of n qubits plus the other op kinds a program holds:
    a dense block, a 16x16 PTM, memory noise with a thermal feed."""
    rng = np.random.default_rng(seed)
    noisy = fusion.noisy_gates(gates, 1e-3)
    prog, _ = fusion.fused_program(noisy, n, 4, 'greedy', None, 'seq')
    S = np.eye(256)
    S[1:, 1:] = rng.normal(size=(255, 255)) / 256
    prog.block([0, 2, 3, 5], S)
    prog.two_qubit(1, 4, ptm.cx_ptm((0.97, 0.02)))
    prog.memory_noise(0.99, 0.7, 0.95)
    prog.one_qubit(n - 1, ptm.gate_ptm('u3', [0.3, 1.1, -0.4]))
    return prog


def run(prog, n, precision, entry):
    lib = native.load()
    st = native.DeviceState(n, precision)
    st.init_product(np.tile([1.0, 0.0, 0.0, 1.0], n), 0.5 ** n)
    arr, params = prog.compile()
    ops = (native.DmOp * len(prog)).from_buffer(arr)
    pv = np.ascontiguousarray(params if params.size else np.zeros(1), dtype=np.float64)
    rc = getattr(lib, entry)(st.h, ctypes.cast(ops, ctypes.c_void_p), len(prog),
                             pv.ctypes.data_as(native.c_double_p), int(params.size))
    native.check(rc, entry)
    out = st.download()
    st.close()
    return out


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_noalloc_matches_run_program_on_two_calls(precision):
    for n, G, seed in ((10, 160, 1), (7, 40, 2)):
        prog = program(n, workloads.local_u3_cx(n, G, seed), seed)
        want = run(prog, n, precision, 'dm_run_program')
        got = run(prog, n, precision, 'dm_run_program_noalloc')
        assert np.array_equal(got, want)
        assert got[0] == 2.0 ** -n


def test_noalloc_refuses_an_image_above_its_scratch():
    n = 8
    prog = Program(n)
    S = np.eye(256)
    for _ in range(17):   # 17 x 512 KiB of FP64 block values > 8 MiB
        prog.block([0, 1, 2, 3], S)
    lib = native.load()
    st = native.DeviceState(n, 'fp64')
    st.init_product(np.tile([1.0, 0.0, 0.0, 1.0], n), 0.5 ** n)
    before = st.download()
    arr, params = prog.compile()
    ops = (native.DmOp * len(prog)).from_buffer(arr)
    rc = lib.dm_run_program_noalloc(st.h, ctypes.cast(ops, ctypes.c_void_p), len(prog),
                                    params.ctypes.data_as(native.c_double_p), int(params.size))
    assert rc == -1
    assert 'exceeds' in native.last_error()
    assert np.array_equal(st.download(), before)
    st.close()
