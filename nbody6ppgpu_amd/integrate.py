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
        send=False, prec=0, timeout=1200, extra=()):
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
        r = subprocess.run(cmd, capture_output=True, text=True, env=env,
                           timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"hermite_ac failed ({r.returncode}): {r.stderr[-2000:]}")
    return [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]
