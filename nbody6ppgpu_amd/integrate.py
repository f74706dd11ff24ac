"""Runs the host Hermite/Ahmad-Cohen driver (bin/hermite_ac, built from
host/hermite_ac.cpp) on a cluster and returns its energy-check records."""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

import numpy as np

from .build import HERMITE


def write_ics(path, m, x, v):
    n = len(m)
    with open(path, "wb") as f:
        np.array([n], dtype=np.int32).tofile(f)
        for a in (m, x, v):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)


def run(m, x, v, t_end=1.0, etai=0.05, etar=0.1, dtadj=0.125, nnbopt=48,
        send=False, prec=0, timeout=1200, extra=(), progress=None):
    """Energy-check records of one run.  With ``progress`` (a callable), the
    driver prints a progress line every 30 s and every output line is passed
    to it as it arrives."""
    if not os.path.exists(HERMITE):
        raise FileNotFoundError(f"{HERMITE}: build it first (build())")
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "ics.bin")
        write_ics(p, m, x, v)
        cmd = [HERMITE, p, str(t_end), "--etai", str(etai), "--etar",
               str(etar), "--dtadj", str(dtadj), "--nnbopt", str(nnbopt)]
        if send:
            cmd.append("--send")
        if prec:
            cmd += ["--prec", str(prec)]
        cmd += list(extra)
        env = dict(os.environ, GPUNB_QUIET="1")
        if progress is None:
            r = subprocess.run(cmd, capture_output=True, text=True, env=env,
                               timeout=timeout)
            if r.returncode != 0:
                raise RuntimeError(f"hermite_ac failed ({r.returncode}): "
                                   f"{r.stderr[-2000:]}")
            out = r.stdout.splitlines()
        else:
            out = _run_streaming(cmd + ["--progress", "30"], env, timeout,
                                 progress)
    return [json.loads(l) for l in out if l.startswith("{")]


def _run_streaming(cmd, env, timeout, progress):
    import time
    t0 = time.monotonic()
    pr = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, env=env, bufsize=1)
    lines = []
    try:
        for line in pr.stdout:
            line = line.rstrip("\n")
            lines.append(line)
            progress(line)
            if time.monotonic() - t0 > timeout:
                raise subprocess.TimeoutExpired(cmd, timeout)
        rc = pr.wait(timeout=max(timeout - (time.monotonic() - t0), 1))
    except BaseException:
        pr.kill()
        pr.wait()
        raise
    if rc != 0:
        raise RuntimeError(f"hermite_ac failed ({rc}): " + "\n".join(lines[-20:]))
    return lines
