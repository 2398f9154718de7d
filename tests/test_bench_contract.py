"""bench.py's command-line contract that needs no GPU: more than one GPU is refused with a message
and a non-zero status before anything touches a device, and the workload constants match the
benchmark's configuration (4096 CCC updates, nb = 512, ranks log-uniform in [8, 192], decay 0.85,
max_rank 384, accuracy 1e-8)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_workload_constants():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)  # defines constants and functions only; main() is not run
    assert (b.NB, b.N_UPDATES, b.LO, b.HI, b.DECAY, b.MAX_RANK, b.ACC) == (512, 4096, 8, 192, 0.85, 384, 1e-8)
    assert b.PEAK_FP64_TFLOPS == 78.6
