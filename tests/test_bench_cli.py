"""bench.py's command line on the CPU: it measures one GPU and says so for any other count."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in ("2", "8"):
        r = run_bench("--gpus", n)
        assert r.returncode == 2
        assert "one GPU" in r.stdout + r.stderr and f"--gpus {n}" in r.stdout + r.stderr


def test_help_lists_the_legs():
    r = run_bench("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--gpus", "--batched", "--adjoint", "--pipe", "--voxel-form", "--frames",
                 "--small", "--tune"):
        assert flag in r.stdout
