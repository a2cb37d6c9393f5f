"""Host-side checks of bench.py (no GPU): the one-GPU contract and the
per-op byte accounting behind the roofline figures."""

import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert p.returncode == 2 and "one GPU" in p.stderr


def test_algorithmic_bytes_are_inputs_plus_outputs():
    b = load_bench()
    for name, pix, nbytes in b.ALL_OPS:
        assert b.input_bytes(name) + b.out_bytes(name) == nbytes, name
        assert name in b.OP_KERNELS
