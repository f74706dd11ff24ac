"""The host driver's on-disk outputs: conf.3 snapshots (the reference's
output.F:488-535 format and string_left.f file names) and checkpoints.

* Every energy check writes conf.3_<t>; its potential energy
  1/2 sum BODY*PHI (PHI = -PHIDBL) and kinetic energy reproduce the
  driver's EP and EK to REAL*4 accuracy, the bodies are 1..N, and the
  Fortran reader with output.F's I/O list reads it.
* A run restarted from the checkpoint at t = 0.25 writes, at t = 0.375,
  the same checkpoint as the uninterrupted run, byte for byte: restart
  parity of the whole integration state.
"""
import os
import subprocess

import numpy as np
import pytest

from nbody6ppgpu_amd import ics, integrate, snapshot

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F_PROG = os.path.join(ROOT, "tests", "fortran", "_build", "conf3_io")


def test_conf3_snapshots(tmp_path):
    c = ics.plummer(2048, seed=501)
    d = str(tmp_path)
    rec = integrate.run(c.m, c.x, c.v, t_end=0.25, etai=0.05, etar=0.1,
                        dtadj=0.125, extra=("--conf3", d))
    names = sorted(os.listdir(d))
    assert names == ["conf.3_0.000", "conf.3_0.125", "conf.3_0.250"]
    for r, name in zip(rec, names):
        s = snapshot.read_conf3(os.path.join(d, name))
        assert s["NTOT"] == 2048 and s["NK"] == 20
        assert s["TTOT"] == pytest.approx(r["t"])
        assert np.array_equal(s["name"], np.arange(1, 2049))
        assert np.array_equal(s["body"], c.m.astype(np.float32))
        m = s["body"].astype(np.float64)
        ep = 0.5 * np.sum(m * s["phi"].astype(np.float64))
        ek = 0.5 * np.sum(m * np.sum(s["v"].astype(np.float64) ** 2, axis=1))
        assert ep == pytest.approx(r["EP"], rel=1e-5)
        assert ek == pytest.approx(r["EK"], rel=1e-5)
        assert s["RSCALE"] > 0 and np.all(s["rho"] >= 0)
    if os.path.exists(F_PROG):
        out = subprocess.run([F_PROG, "read", os.path.join(d, names[-1])],
                             check=True, capture_output=True,
                             text=True).stdout.split()
        assert [int(w) for w in out[:4]] == [2048, 1, 1, 20]


def test_restart_continues_bit_for_bit(tmp_path):
    c = ics.king(4096, 6.0, seed=502, m=ics.kroupa(4096, seed=503))
    a, b, r = (str(tmp_path / k) for k in ("uninterrupted", "first", "restart"))
    kw = dict(etai=0.05, etar=0.1, dtadj=0.125)
    ra = integrate.run(c.m, c.x, c.v, t_end=0.375, extra=("--dump", a), **kw)
    integrate.run(c.m, c.x, c.v, t_end=0.25, extra=("--dump", b), **kw)
    rr = integrate.run(c.m, c.x, c.v, t_end=0.375,
                       extra=("--restart", b, "--dump", r), **kw)
    assert rr[0].get("restart") and rr[0]["t"] == pytest.approx(0.25)
    assert rr[-1]["t"] == pytest.approx(0.375)
    assert rr[-1]["E"] == ra[-1]["E"]
    assert rr[-1]["blocks"] == ra[-1]["blocks"]
    assert open(a, "rb").read() == open(r, "rb").read()
    # a checkpoint for other options is refused
    with pytest.raises(RuntimeError, match="other integration options"):
        integrate.run(c.m, c.x, c.v, t_end=0.375, etai=0.02, etar=0.1,
                      dtadj=0.125, extra=("--restart", b))
