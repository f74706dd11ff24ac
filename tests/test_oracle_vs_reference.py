"""The oracle against the reference's own CPU path.

Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py
from src/Main/reg.avx.cpp, reg.sse.cpp and pot.avx.cpp compiled from the
reference sources) are always checked; the live reference libraries under
oracle/_ref/ are checked too when they have been built.

The reference CPU path computes in FP32 (FP32-rounded coordinates, rsqrtps
plus one Newton step, FP32 sums), so forces agree to FP32 accuracy and the
neighbour sets agree except for pairs within FP32 rounding of the boundary.
"""
import glob
import os

import numpy as np
import pytest

from nbody6ppgpu_amd import ics
from oracle import oracle, ref

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def boundary_pairs(z, i, mflag):
    """j whose FP64 min(r2, r2p) is within 1e-5 relative of the threshold."""
    d = z["x"] - z["x"][i]
    p = d + z["dtr"][i] * (z["v"] - z["v"][i])
    mn = np.minimum(np.sum(d * d, 1), np.sum(p * p, 1))
    thr = z["h2"][i] * (z["m"] if mflag else 1.0)
    return set(np.nonzero(np.abs(mn / thr - 1.0) < 1e-5)[0])


def check_against(z, acc, jrk, pot, lst, tag, tol):
    mflag = int(z["m_flag"])
    rl = z[f"list_{tag}"]
    nb_ok = 0
    for i in range(lst.shape[0]):
        a, b = lst[i, 0], rl[i, 0]
        if a < 0 or b < 0:
            assert a == b, f"overflow code of i={i}: {a} vs {b}"
            continue
        sa, sb = set(lst[i, 1:1 + a]), set(rl[i, 1:1 + b])
        assert (sa ^ sb) <= boundary_pairs(z, i, mflag), f"list of i={i}"
        nb_ok += 1
    assert nb_ok > 0 or np.all(lst[:, 0] < 0)
    # Reference outputs are FP32 sums; compare normwise per particle against
    # the scale of the sum (sum of |terms| ~ the magnitude of the result).
    ra, rj, rp = z[f"acc_{tag}"], z[f"jrk_{tag}"], z[f"pot_{tag}"]
    ok = lst[:, 0] >= 0
    ea = np.linalg.norm(acc - ra, axis=1) / np.linalg.norm(acc, axis=1)
    ej = np.linalg.norm(jrk - rj, axis=1) / np.linalg.norm(jrk, axis=1)
    ep = np.abs(pot - rp) / np.abs(pot)
    if ok.any():
        assert np.max(ea[ok]) < tol
        assert np.max(ej[ok]) < 20 * tol
        assert np.max(ep[ok]) < tol


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "*.npz"))))
def test_oracle_matches_golden(path):
    z = np.load(path, allow_pickle=False)
    if "pot_avx" in z.files and "h2" not in z.files:
        p = oracle.pot(1, z["m"].shape[0], z["m"], z["x"])
        for tag in ("avx", "sse"):
            assert np.max(np.abs(p - z[f"pot_{tag}"]) / p) < 1e-6
        return
    acc, jrk, pot, lst, _ = oracle.regf(z["m"], z["x"], z["v"], z["h2"],
                                        z["dtr"], z["x"], z["v"],
                                        int(z["lmax"]), int(z["nnbmax"]),
                                        int(z["m_flag"]))
    for tag in ("avx", "sse"):
        check_against(z, acc, jrk, pot, lst, tag, 1e-5)


@pytest.mark.skipif(not ref.available("reg_avx"),
                    reason="reference CPU libraries not built (oracle/_ref)")
@pytest.mark.parametrize("lib", ["reg_avx", "reg_sse"])
@pytest.mark.parametrize("m_flag", [0, 1])
def test_oracle_matches_live_reference(lib, m_flag):
    c = ics.with_binaries(ics.king(1500, 6.0, seed=31, m=ics.kroupa(1500, seed=32)),
                          0.1, seed=33)
    rs, stepr = ics.neighbour_radii(c, k=50)
    h2 = rs * rs / (np.mean(c.m) if m_flag else 1.0)
    lmax, nnbmax = 401, 400
    R = ref.RegLib(lib, c.n + 10)
    R.send(c.m, c.x, c.v)
    ra, rj, rp, rl = R.regf(h2, stepr, c.x, c.v, lmax, nnbmax, m_flag)
    R.close()
    tag = lib.split("_")[1]
    z = {"m": c.m, "x": c.x, "v": c.v, "h2": h2, "dtr": stepr,
         "m_flag": m_flag, f"acc_{tag}": ra, f"jrk_{tag}": rj,
         f"pot_{tag}": rp, f"list_{tag}": rl}
    acc, jrk, pot, lst, _ = oracle.regf(c.m, c.x, c.v, h2, stepr, c.x, c.v,
                                        lmax, nnbmax, m_flag)
    check_against(z, acc, jrk, pot, lst, tag, 1e-5)


@pytest.mark.skipif(not ref.available("pot_avx"),
                    reason="reference CPU libraries not built (oracle/_ref)")
def test_oracle_pot_matches_live_reference():
    c = ics.plummer(2000, seed=34, m=ics.kroupa(2000, seed=35))
    p = oracle.pot(1, c.n, c.m, c.x)
    for lib in ("pot_avx", "pot_sse"):
        assert np.max(np.abs(ref.pot(lib, c.m, c.x) - p) / p) < 1e-6
