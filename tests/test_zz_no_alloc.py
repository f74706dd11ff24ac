"""gpunb_regf_dev_noalloc (the regular force with its scratch in the
library's static device array) against gpunb_regf_dev on the same inputs:
two calls in a row, the second smaller; lists, counts and nearest
neighbours bit-exact, forces within the 1e-6 relative bar per particle (the
two may split the j set into different slices, so FP64 partial sums can be
added in another order).  A call too large for the array is refused."""
import numpy as np
import pytest

from nbody6ppgpu_amd import gpunb, ics
from nbody6ppgpu_amd.hiprt import DeviceArray
from oracle.parity import FORCE_TOL

pytestmark = pytest.mark.gpu


class Outputs:
    def __init__(self, ni, nnbmax):
        self.acc = DeviceArray((ni, 3), np.float64).zero()
        self.jrk = DeviceArray((ni, 3), np.float64).zero()
        self.pot = DeviceArray((ni,), np.float64).zero()
        self.nnb = DeviceArray((ni,), np.int32).zero()
        self.nnid = DeviceArray((ni,), np.int32).zero()
        self.off = DeviceArray((ni,), np.int32).zero()
        self.cap = ni * max(nnbmax, 1)
        self.lst = DeviceArray((self.cap,), np.int32).zero()
        self.tot = DeviceArray((1,), np.int32).zero()

    def args(self):
        return (self.acc.ptr, self.jrk.ptr, self.pot.ptr)

    def list_args(self):
        return (self.nnb.ptr, self.nnid.ptr, self.lst.ptr, self.off.ptr,
                self.tot.ptr, self.cap)

    def rows(self):
        nnb, off, L = self.nnb.numpy(), self.off.numpy(), self.lst.numpy()
        return [L[off[i]:off[i] + nnb[i]].copy() if nnb[i] > 0 else None
                for i in range(len(nnb))]


def call(fn, ni, d, out, nnbmax, m_flag):
    rc = fn(ni, d["h2"].ptr, d["dtr"].ptr, d["x"].ptr, d["v"].ptr,
            *out.args(), nnbmax, m_flag, *out.list_args())
    if rc != 0:
        gpunb._check()
    gpunb.gpunb_sync()


def rel_err(a, b):
    a = a.reshape(len(a), -1)
    b = b.reshape(len(b), -1)
    den = np.linalg.norm(b, axis=1)
    return np.max(np.linalg.norm(a - b, axis=1) / np.where(den > 0, den, 1))


def test_noalloc_matches_regf_dev_two_calls():
    n = 20000
    c = ics.plummer(n, seed=71, m=ics.kroupa(n, seed=72))
    rs, dtr = ics.neighbour_radii(c, k=48)
    lib = gpunb.load()
    gpunb.gpunb_open(n + 16)
    try:
        gpunb.gpunb_send(c.m, c.x, c.v)
        rng = np.random.default_rng(73)
        for ni, nnbmax, m_flag in ((6000, 200, 0), (2500, 1200, 1)):
            idx = rng.choice(n, ni, replace=False)
            h2 = rs[idx] ** 2
            if m_flag:
                h2 = h2 / np.mean(c.m)
            print(f"noalloc: ni={ni} nnbmax={nnbmax} m_flag={m_flag}",
                  flush=True)
            d = {"x": DeviceArray.from_numpy(c.x[idx]),
                 "v": DeviceArray.from_numpy(c.v[idx]),
                 "h2": DeviceArray.from_numpy(h2),
                 "dtr": DeviceArray.from_numpy(dtr[idx])}
            ref, got = Outputs(ni, nnbmax), Outputs(ni, nnbmax)
            call(lib.gpunb_regf_dev, ni, d, ref, nnbmax, m_flag)
            call(lib.gpunb_regf_dev_noalloc, ni, d, got, nnbmax, m_flag)
            nnb = ref.nnb.numpy()
            assert np.all(nnb > 0), "every i lists itself, no overflow"
            assert np.array_equal(got.nnb.numpy(), nnb)
            assert np.array_equal(got.nnid.numpy(), ref.nnid.numpy())
            assert int(got.tot.numpy()[0]) == int(ref.tot.numpy()[0])
            for a, b in zip(got.rows(), ref.rows()):
                assert np.array_equal(a, b)
            for name in ("acc", "jrk", "pot"):
                e = rel_err(getattr(got, name).numpy(),
                            getattr(ref, name).numpy())
                print(f"  {name}: max relative difference {e:.2e}")
                assert e < FORCE_TOL, name
    finally:
        gpunb.gpunb_close()


def test_noalloc_refuses_oversize_call():
    n = 8000
    c = ics.plummer(n, seed=74)
    rs, dtr = ics.neighbour_radii(c, k=48)
    lib = gpunb.load()
    gpunb.gpunb_open(n + 16)
    try:
        gpunb.gpunb_send(c.m, c.x, c.v)
        ni, nnbmax = 6000, 3000     # 6016 * 3000 * 4 bytes of lists > 48 MiB
        d = {"x": DeviceArray.from_numpy(c.x[:ni]),
             "v": DeviceArray.from_numpy(c.v[:ni]),
             "h2": DeviceArray.from_numpy(rs[:ni] ** 2),
             "dtr": DeviceArray.from_numpy(dtr[:ni])}
        out = Outputs(ni, nnbmax)
        with pytest.raises(gpunb.GpunbError, match="scratch"):
            call(lib.gpunb_regf_dev_noalloc, ni, d, out, nnbmax, 0)
        # The library stays usable.
        ref, got = Outputs(ni, 100), Outputs(ni, 100)
        call(lib.gpunb_regf_dev, ni, d, ref, 100, 0)
        call(lib.gpunb_regf_dev_noalloc, ni, d, got, 100, 0)
        assert np.array_equal(got.nnb.numpy(), ref.nnb.numpy())
    finally:
        gpunb.gpunb_close()
