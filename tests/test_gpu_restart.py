"""Restart: the library holds no state that changes results beyond what a
restarted host gives it again.

A host restarted from its dump (mydump.F writes the bodies' X0, X0DOT, F,
FDOT, BODY, T0 and radii: src/Main/mydump.F:100-113) opens the library
afresh, sets the predictor records of every body and carries on.  Its next
regular forces must equal, bit for bit, those of the run that never stopped,
whatever that run's library had cached from earlier calls: the sorted j view,
the i order and its buffers, the work tables, the list-size estimate of the
host-array path, the scratch counts."""
import numpy as np
import pytest

from nbody6ppgpu_amd import gpunb, ics

pytestmark = pytest.mark.gpu

LMAX, NNBMAX = 301, 300


def blocks(n, rng):
    """The regular blocks of one step: every body (Hilbert-sorted mixed
    kernel), 1000 scattered bodies (FP64 kernel), and 257 scattered bodies
    with mass-weighted lists."""
    return [(np.arange(n), 0), (np.sort(rng.choice(n, 1000, replace=False)), 0),
            (rng.choice(n, 257, replace=False), 1)]


def step(t, n, h2, dtr, mbar, blist):
    gpunb.gpunb_predict(t)
    x, v = gpunb.gpunb_jpget(n)
    out = []
    for idx, m_flag in blist:
        hh = h2[idx] / mbar if m_flag else h2[idx]
        out.append(gpunb.gpunb_regf(hh, dtr[idx], x[idx], v[idx], LMAX,
                                    NNBMAX, m_flag, nearest=True))
    return out


def test_restart_reproduces_uninterrupted_calls():
    n = 20000
    c = ics.king(n, 6.0, seed=61, m=ics.kroupa(n, seed=62))
    rng = np.random.default_rng(63)
    rs, dtr = ics.neighbour_radii(c, k=48)
    h2 = rs * rs
    mbar = float(np.mean(c.m))
    f = rng.normal(0, 0.5, (n, 3))
    fd = rng.normal(0, 0.5, (n, 3))
    t0 = rng.uniform(0.0, 1.0 / 64, n)
    t1, t2 = 1.0 / 32, 3.0 / 64
    addr = np.arange(1, n + 1)
    b1, b2 = blocks(n, rng), blocks(n, rng)
    # The state a corrector leaves at t1 for the bodies of a block.
    upd = np.sort(rng.choice(n, 4000, replace=False))
    x1, v1 = c.x.copy(), c.v.copy()
    f1, fd1, t01 = f.copy(), fd.copy(), t0.copy()
    x1[upd] += rng.normal(0, 1e-4, (len(upd), 3))
    v1[upd] += rng.normal(0, 1e-4, (len(upd), 3))
    f1[upd] = rng.normal(0, 0.5, (len(upd), 3))
    fd1[upd] = rng.normal(0, 0.5, (len(upd), 3))
    t01[upd] = t1

    print("restart: uninterrupted run", flush=True)
    gpunb.gpunb_open(n + 16)
    try:
        gpunb.gpunb_jpset(addr, c.x, c.v, f, fd, c.m, t0)
        step(t1, n, h2, dtr, mbar, b1)
        gpunb.gpunb_jpset(addr[upd], x1[upd], v1[upd], f1[upd], fd1[upd],
                          c.m[upd], t01[upd])
        ref = step(t2, n, h2, dtr, mbar, b2)
    finally:
        gpunb.gpunb_close()

    print("restart: run restarted at t1", flush=True)
    gpunb.gpunb_open(n + 16)
    try:
        gpunb.gpunb_jpset(addr, x1, v1, f1, fd1, c.m, t01)
        got = step(t2, n, h2, dtr, mbar, b2)
    finally:
        gpunb.gpunb_close()

    for k, (a, b) in enumerate(zip(got, ref)):
        for name, u, w in zip(("acc", "jrk", "pot", "list", "nnid"), a, b):
            assert np.array_equal(u, w), (k, name)
    assert np.all(ref[0][3][:, 0] > 0), "every body lists itself"
