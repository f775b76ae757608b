"""bench.py command line (CPU only): the one-GPU contract."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    """The path has one GPU's form only: --gpus above 1 exits with a message
    before any device work."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr and r.stdout == ''


def test_driver_command_plans_bounded_work():
    """The driver's `--steps 20 --warmup 5` headline run applies the
    1000-gate program a bounded number of times: at the rates measured on
    MI355X (plain FP64 ~11 s, fused FP32 ~0.9 s per application) plus the
    20 s CPU baseline it stays far inside a 600 s limit, and the optional
    paths only run when asked for."""
    sys.path.insert(0, ROOT)
    import bench
    args = bench.parse_args(['--gpus', '1', '--steps', '20', '--warmup', '5'])
    plan = bench.planned_applications(args)
    assert plan['other_paths'] == 0
    assert plan['plain_fp64'] <= 1.2
    assert plan['fused_fp32'] == 26
    est_s = 11.0 * plan['plain_fp64'] + 0.9 * plan['fused_fp32'] + 20.0
    assert est_s < 120.0
    assert bench.planned_applications(bench.parse_args(['--all-paths']))['other_paths'] > 0


def test_batch_choice_parsing():
    """--batch-choice fixes the batch configuration per precision (no
    autotuning for the precisions it names) and refuses malformed values."""
    import pytest
    sys.path.insert(0, ROOT)
    import bench
    assert bench.parse_batch_choice('') == {}
    assert bench.parse_batch_choice('fp64=6/256/256,fp32=5/256/64') == {'fp64': (6, 256, 256), 'fp32': (5, 256, 64)}
    with pytest.raises(ValueError):
        bench.parse_batch_choice('fp16=6/256/256')
    with pytest.raises(ValueError):
        bench.parse_batch_choice('fp32=6/256')
