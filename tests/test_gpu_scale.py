"""GPU parity at full size: the 64-cluster ensemble (about 845k bodies) and
the N = 1,048,576 mass-segregated Kroupa cluster.

The oracle is exact FP64 O(ni N): it checks a seeded sample of i from every
cluster (lists bit-exact, nearest neighbours, forces at the plain 1e-6 bar);
size-independent properties cover every i of the 1M cluster:

* h2 = 0: Newton's third law (sum_i m_i a_i = 0, sum_i m_i jerk_i = 0 to
  1e-6 of sum_i m_i |a_i|) and the regular potential equal to gpupot_'s;
* a uniform h2 and dtr: exactly symmetric neighbour lists (the FP64
  predicate is symmetric in (i, j)).
"""
import numpy as np
import pytest

from nbody6ppgpu_amd import ensemble as ens
from nbody6ppgpu_amd import gpunb, ics
from nbody6ppgpu_amd.hiprt import DeviceArray
from oracle import oracle
from oracle.parity import FORCE_TOL, assert_parity

pytestmark = pytest.mark.gpu


def rows_of(nnb, off, lst, idx, width):
    """List rows (count, members...) of the i in idx from compact lists."""
    out = np.zeros((len(idx), width), dtype=np.int32)
    for r, i in enumerate(idx):
        k = nnb[i]
        out[r, 0] = k
        if k > 0:
            out[r, 1:1 + k] = lst[off[i]:off[i] + k]
    return out


def test_ensemble_64_clusters_full_size():
    """64 Plummer clusters, N log-uniform in [1024, 65536], equal-mass and
    Kroupa alternating, every body active, through the batched launch
    captured into a hipGraph with a per-bucket tile table (the autotuner's
    kind of output), replayed twice; 24 seeded i per cluster against the
    oracle."""
    cl = ens.make_ensemble()
    nnbmax = 300
    sizes = [c[0].n for c in cl]
    print(f"ensemble: {len(cl)} clusters, {sum(sizes)} bodies", flush=True)
    gpunb.gpunb_ens_open(sizes)
    try:
        gpunb.gpunb_ens_send([(c[0].m, c[0].x, c[0].v) for c in cl])
        b = ens.upload(cl, nnbmax)
        tiles = [0] * 24
        for bk in range(10, 17):
            tiles[bk] = 8 if bk < 14 else 32
        gpunb.gpunb_ens_set_bucket_tiles(tiles)
        gpunb.gpunb_ens_graph(True)
        for _ in range(2):
            b.nnb.zero()
            ens.call(b, b.ni, nnbmax)
            gpunb.gpunb_sync()
        nnb, off, nnid = b.nnb.numpy(), b.off.numpy(), b.nnid.numpy()
        lst = b.lst.numpy()
        acc, jrk, pot = b.acc.numpy(), b.jrk.numpy(), b.pot.numpy()
    finally:
        gpunb.gpunb_ens_graph(False)
        gpunb.gpunb_ens_set_bucket_tiles([])
        gpunb.gpunb_ens_close()
    rng = np.random.default_rng(404)
    i0 = 0
    for k, (c, h2, dtr) in enumerate(cl):
        n = c.n
        sub = np.sort(rng.choice(n, 24, replace=False))
        g = (acc[i0 + sub], jrk[i0 + sub], pot[i0 + sub],
             rows_of(nnb, off, lst, i0 + sub, nnbmax + 1), nnid[i0 + sub])
        o = oracle.regf(c.m, c.x, c.v, h2[sub], dtr[sub], c.x[sub], c.v[sub],
                        nnbmax + 1, nnbmax, 0)
        assert_parity(g, o, label=f"ensemble cluster {k} (n={n})",
                      inputs=(c.m, c.x, c.v, c.x[sub], c.v[sub]))
        i0 += n


def test_segregated_kroupa_1m():
    """Mass-segregated Plummer N = 1,048,576, Kroupa IMF: one regular force
    of all bodies through the device entry point (h2 = RS^2 from the 48th
    neighbour distance), a seeded oracle sample of 192 i, and the full-size
    invariants."""
    n = 1 << 20
    c = ics.plummer_segregated(n, seed=505)
    print("segregated 1M: cluster built", flush=True)
    rs, dtr = ics.neighbour_radii(c, k=48)
    h2 = rs * rs
    nnbmax = 550
    print("segregated 1M: radii set", flush=True)
    gpunb.gpunb_open(n + 10)
    try:
        gpunb.gpunb_send(c.m, c.x, c.v)
        d = {k: DeviceArray.from_numpy(a) for k, a in
             (("x", c.x), ("v", c.v), ("h2", h2), ("dtr", dtr))}
        acc = DeviceArray((n, 3), np.float64)
        jrk = DeviceArray((n, 3), np.float64)
        pot = DeviceArray((n,), np.float64)
        nnb = DeviceArray((n,), np.int32)
        nnid = DeviceArray((n,), np.int32)
        off = DeviceArray((n,), np.int32)
        cap = n * 160
        lst = DeviceArray((cap,), np.int32)
        tot = DeviceArray((1,), np.int32)
        lib = gpunb.load()

        def regf(h2dev, dtdev, m_flag):
            rc = lib.gpunb_regf_dev(n, h2dev.ptr, dtdev.ptr, d["x"].ptr,
                                    d["v"].ptr, acc.ptr, jrk.ptr, pot.ptr,
                                    nnbmax, m_flag, nnb.ptr, nnid.ptr, lst.ptr,
                                    off.ptr, tot.ptr, cap)
            if rc != 0:
                gpunb._check()
            gpunb.gpunb_sync()

        regf(d["h2"], d["dtr"], 0)
        print("segregated 1M: regular force done", flush=True)
        rng = np.random.default_rng(506)
        sub = np.sort(rng.choice(n, 192, replace=False))
        nb, of, nid = nnb.numpy(), off.numpy(), nnid.numpy()
        L = lst.numpy()
        g = (acc.numpy()[sub], jrk.numpy()[sub], pot.numpy()[sub],
             rows_of(nb, of, L, sub, nnbmax + 1), nid[sub])
        assert np.all(nb >= 1), "every body lists itself"
        o = oracle.regf(c.m, c.x, c.v, h2[sub], dtr[sub], c.x[sub], c.v[sub],
                        nnbmax + 1, nnbmax, 0)
        assert_parity(g, o, label="segregated Kroupa 1M",
                      inputs=(c.m, c.x, c.v, c.x[sub], c.v[sub]))

        # h2 = 0: Newton's third law and the potential against gpupot_.
        z = DeviceArray.from_numpy(np.zeros(n))
        regf(z, z, 0)
        a0, j0, p0 = acc.numpy(), jrk.numpy(), pot.numpy()
        assert np.all(nnb.numpy() == 0)
        m = c.m[:, None]
        for name, a in (("acc", a0), ("jrk", j0)):
            t = np.abs(np.sum(m * a, axis=0))
            s = np.sum(m * np.abs(a), axis=0)
            assert np.all(t <= 1e-6 * s), (name, t / s)
        phi = gpunb.gpupot(1, n, c.m, c.x)
        assert np.max(np.abs(p0 - phi) / phi) < FORCE_TOL
        print("segregated 1M: h2 = 0 invariants hold", flush=True)

        # Uniform radius and step: symmetric lists (0.8 of the median radius
        # keeps the densest core body at about 300 neighbours).
        hu = DeviceArray.from_numpy(np.full(n, (0.8 * np.median(rs)) ** 2))
        du = DeviceArray.from_numpy(np.full(n, 1.0 / 64))
        regf(hu, du, 0)
        cnt = nnb.numpy()
        assert np.all((cnt >= 1) & (cnt <= nnbmax))
        of = off.numpy()
        L = lst.numpy()
        rows = np.repeat(np.arange(n, dtype=np.int64), cnt)
        cols = np.concatenate([L[of[i]:of[i] + cnt[i]] for i in range(n)])
        keep = rows != cols
        a = rows[keep] * n + cols[keep]
        b = cols[keep].astype(np.int64) * n + rows[keep]
        assert np.array_equal(np.sort(a), np.sort(b)), "lists not symmetric"
        print("segregated 1M: uniform-radius lists symmetric", flush=True)
    finally:
        gpunb.gpunb_close()
