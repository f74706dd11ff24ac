"""bench.py's command line without a GPU: every leg can be switched off,
and a multi-GPU request is refused with a message before any device work
(this build measures one MI355X)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_multi_gpu_request_is_refused():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_every_leg_has_a_switch():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--no-cpu", "--no-regcor", "--no-energy", "--no-irr",
                 "--no-predict-pot", "--no-hermite", "--no-ensemble",
                 "--no-learned", "--steps", "--warmup", "--gpus"):
        assert flag in r.stdout, flag
