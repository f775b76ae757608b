"""Build recipe for the native library (libdmsim.so, gfx950)."""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(PKG_DIR))
CSRC = os.path.join(os.path.dirname(PKG_DIR), 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
LIB = os.path.join(PKG_DIR, 'libdmsim.so')

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    # packed FP32 (v_pk_*_f32) stays off for every HIP translation unit
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result',
    # no FMA contraction in host code: the planner's scores and the host MLP
    # forward round the same way wherever they are evaluated
    '-ffp-contract=off',
]


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError('hipcc not found')


def build(verbose=False):
    """Compile csrc/*.hip (device + host) and csrc/*.cpp (host) into
    dmsim/libdmsim.so (always recompiles)."""
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(('.hip', '.cpp')))
    tmp = LIB + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-I', INCLUDE] + srcs + ['-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB
