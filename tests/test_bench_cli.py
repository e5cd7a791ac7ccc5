"""bench.py's one-GPU contract: more than one GPU is refused with a message
(the local matching path is measured on one MI355X); no GPU is touched."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_roofline_unit_counts():
    """The per-subset figures the bench's roofline is built from: 12 s^2 +
    4 (s+8)^2 + 32 algorithmic bytes (72,048 B at s = 65, 17,648 B at 31) and
    120 flop per pixel per evaluation + 46 per pixel once."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("aldic_bench_mod", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    assert b.icgn_bytes(65) == 72048.0 and b.icgn_bytes(31) == 17648.0
    assert b.icgn_flops(65, 4) == (120.0 * 4 + 46.0) * 65 * 65
