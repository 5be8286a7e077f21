"""bench.py's command line parses without a GPU: every leg has its switch,
and more than one GPU is refused with a message (the batched QPs are
independent; only one GPU is measured)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_help_lists_every_leg_switch():
    r = run("--help")
    assert r.returncode == 0, r.stderr
    for flag in ("--gpus", "--steps", "--warmup", "--no-cpu", "--no-plain", "--no-fp64-batch", "--no-serving",
                 "--no-dropin", "--no-ica", "--no-coupled", "--no-step", "--no-traffic", "--no-spmv", "--no-tune"):
        assert flag in r.stdout, flag


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode != 0
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
