"""Ensemble path: many independent clusters in one launch pair
(gpunb_ens_open/send/regf_dev), optionally captured into a hipGraph.

Bar: every cluster's neighbour lists and nearest neighbours identical to the
single-cluster ABI's and the oracle's; forces within the oracle/parity.py bar;
graph replays identical to the eager launch, with eager work between replays.
"""
import numpy as np
import pytest

from nbody6ppgpu_amd import gpunb, ics
from nbody6ppgpu_amd.hiprt import DeviceArray
from oracle import oracle
from oracle.parity import assert_parity

pytestmark = pytest.mark.gpu


def make_clusters(sizes, seed):
    out = []
    for k, n in enumerate(sizes):
        m = ics.kroupa(n, seed=seed + 7 * k) if k % 2 else None
        c = ics.plummer(n, seed=seed + 7 * k + 1, m=m)
        rs, dtr = ics.neighbour_radii(c, k=min(40, max(n - 2, 1)))
        out.append((c, rs * rs, dtr))
    return out


class EnsRun:
    """Device buffers and one ensemble call over the first ni[c] bodies."""

    def __init__(self, cl, ni, nnbmax):
        self.cl, self.ni, self.nnbmax = cl, list(ni), nnbmax
        cat = lambda f: np.concatenate([f(c, n) for c, n in zip(cl, ni)])
        self.x = DeviceArray.from_numpy(cat(lambda c, n: c[0].x[:n]))
        self.v = DeviceArray.from_numpy(cat(lambda c, n: c[0].v[:n]))
        self.h2 = DeviceArray.from_numpy(cat(lambda c, n: c[1][:n]))
        self.dt = DeviceArray.from_numpy(cat(lambda c, n: c[2][:n]))
        tot = int(sum(ni))
        self.tot = tot
        self.acc = DeviceArray((tot, 3), np.float64)
        self.jrk = DeviceArray((tot, 3), np.float64)
        self.pot = DeviceArray((tot,), np.float64)
        self.nnb = DeviceArray((tot,), np.int32)
        self.nnid = DeviceArray((tot,), np.int32)
        self.off = DeviceArray((tot,), np.int32)
        self.cap = max(tot * nnbmax, 1)
        self.lst = DeviceArray((self.cap,), np.int32)
        self.total = DeviceArray((1,), np.int32)

    def call(self):
        niv = np.ascontiguousarray(self.ni, dtype=np.int32)
        rc = gpunb.load().gpunb_ens_regf_dev(
            len(self.ni), niv.ctypes.data_as(gpunb._ip), self.h2.ptr,
            self.dt.ptr, self.x.ptr, self.v.ptr, self.acc.ptr, self.jrk.ptr,
            self.pot.ptr, self.nnbmax, 0, self.nnb.ptr, self.nnid.ptr,
            self.lst.ptr, self.off.ptr, self.total.ptr, self.cap)
        if rc != 0:
            gpunb._check()
        gpunb.gpunb_sync()

    def results(self):
        """Per cluster (acc, jrk, pot, list rows (ni, nnbmax+1), nnid)."""
        acc, jrk, pot = self.acc.numpy(), self.jrk.numpy(), self.pot.numpy()
        nnb, nnid, off = self.nnb.numpy(), self.nnid.numpy(), self.off.numpy()
        L = self.lst.numpy()
        out, i0 = [], 0
        for n in self.ni:
            rows = np.zeros((n, self.nnbmax + 1), dtype=np.int32)
            for i in range(n):
                k = nnb[i0 + i]
                rows[i, 0] = k
                if k > 0:
                    rows[i, 1:1 + k] = L[off[i0 + i]:off[i0 + i] + k]
            s = slice(i0, i0 + n)
            out.append((acc[s], jrk[s], pot[s], rows, nnid[s]))
            i0 += n
        return out


SIZES = [2048, 37, 1000, 64, 5000, 130, 3071, 700]


def test_ensemble_matches_single_cluster_and_oracle():
    cl = make_clusters(SIZES, 501)
    nnbmax = 300
    ni = [c[0].n for c in cl]
    ni[2] = 333                                   # a partial active block
    gpunb.gpunb_ens_open([c[0].n + 10 for c in cl])
    try:
        gpunb.gpunb_ens_send([(c[0].m, c[0].x, c[0].v) for c in cl])
        run = EnsRun(cl, ni, nnbmax)
        run.call()
        res = run.results()
        assert gpunb.load().gpunb_ens_pairs() == sum(
            n * c[0].n for n, c in zip(ni, cl))
    finally:
        gpunb.gpunb_ens_close()
    for k, ((c, h2, dtr), n, r) in enumerate(zip(cl, ni, res)):
        o = oracle.regf(c.m, c.x, c.v, h2[:n], dtr[:n], c.x[:n], c.v[:n],
                        nnbmax + 1, nnbmax, 0)
        assert_parity(r, o, label=f"ensemble cluster {k} (n={c.n}, ni={n})",
                      inputs=(c.m, c.x, c.v, c.x[:n], c.v[:n]))


def test_graph_replays_match_eager():
    cl = make_clusters([4096, 1500, 256, 8000], 601)
    nnbmax = 300
    ni = [c[0].n for c in cl]
    gpunb.gpunb_ens_open([c[0].n for c in cl])
    try:
        gpunb.gpunb_ens_send([(c[0].m, c[0].x, c[0].v) for c in cl])
        run = EnsRun(cl, ni, nnbmax)
        gpunb.gpunb_ens_graph(False)
        run.call()
        eager = run.results()
        gpunb.gpunb_ens_graph(True)
        # Unrelated eager work between replays (single-cluster ABI calls).
        c0 = cl[0][0]
        gpunb.gpunb_open(c0.n + 10)
        for rep in range(4):
            run.acc.zero()
            run.nnb.zero()
            run.call()
            got = run.results()
            for a, b in zip(got, eager):
                for u, w in zip(a, b):
                    assert np.array_equal(u, w), f"replay {rep}"
            gpunb.gpunb_send(c0.m, c0.x, c0.v)
            gpunb.gpunb_regf(cl[0][1][:100], cl[0][2][:100], c0.x[:100],
                             c0.v[:100], 301, 300, 0)
        gpunb.gpunb_close()
    finally:
        gpunb.gpunb_ens_graph(False)
        gpunb.gpunb_ens_close()


def test_bucket_tiles_change_decomposition_not_results():
    cl = make_clusters([3000, 1100, 9000], 701)
    nnbmax = 300
    ni = [c[0].n for c in cl]
    gpunb.gpunb_ens_open([c[0].n for c in cl])
    try:
        gpunb.gpunb_ens_send([(c[0].m, c[0].x, c[0].v) for c in cl])
        run = EnsRun(cl, ni, nnbmax)
        run.call()
        base = run.results()
        for tiles in ([1] * 24, [3] * 24, [1000] * 24):
            gpunb.gpunb_ens_set_bucket_tiles(tiles)
            run.call()
            got = run.results()
            for a, b in zip(got, base):
                assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
                for u, w in zip(a[:3], b[:3]):
                    assert np.max(np.abs(u - w)) <= 1e-6 * np.max(np.abs(w))
    finally:
        gpunb.gpunb_ens_set_bucket_tiles([])
        gpunb.gpunb_ens_close()


@pytest.mark.parametrize("prec", [0, 65])
def test_persistent_launch_matches_plain(prec):
    """The persistent pass 1 (per-XCD task queues, largest tasks first)
    computes every (i-wavefront, slice) partial exactly as the one-workgroup-
    per-task launch does: results bit-identical, in the default (mixed,
    Hilbert-sorted i) and the FP64 kernel."""
    cl = make_clusters([6000, 300, 2500, 64, 12000, 900], 801)
    nnbmax = 300
    ni = [c[0].n for c in cl]
    ni[3] = 17
    gpunb.gpunb_ens_open([c[0].n for c in cl])
    try:
        gpunb.gpunb_set_precision(prec)
        gpunb.gpunb_ens_send([(c[0].m, c[0].x, c[0].v) for c in cl])
        run = EnsRun(cl, ni, nnbmax)
        gpunb.gpunb_ens_persistent(False)
        run.call()
        plain = run.results()
        gpunb.gpunb_ens_persistent(True)
        for rep in range(2):
            run.acc.zero()
            run.nnb.zero()
            run.call()
            got = run.results()
            for k, (a, b) in enumerate(zip(got, plain)):
                for u, w in zip(a, b):
                    assert np.array_equal(u, w), (prec, rep, k)
    finally:
        gpunb.gpunb_ens_persistent(True)
        gpunb.gpunb_set_precision(0)
        gpunb.gpunb_ens_close()
    o = cl[4]
    c, h2, dtr = o
    sub = np.arange(0, c.n, 97)
    r = plain[4]
    g = (r[0][sub], r[1][sub], r[2][sub], r[3][sub], r[4][sub])
    ref = oracle.regf(c.m, c.x, c.v, h2[sub], dtr[sub], c.x[sub], c.v[sub],
                      nnbmax + 1, nnbmax, 0)
    assert_parity(g, ref, label=f"persistent ensemble cluster 4 (prec {prec})",
                  inputs=(c.m, c.x, c.v, c.x[sub], c.v[sub]))
