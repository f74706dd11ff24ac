"""GPU parity of the regular force (gpunb_regf_ / gpunb_regfnn_) against the
FP64 oracle, through the C-ABI.  The bar is oracle/parity.py's: lists and
nearest neighbours bit-exact, acc/pot within 1e-6 relative per particle, jerk
within 1e-6 of max(|jerk|, J/2) (J = sum of |jerk terms|).
"""
import numpy as np
import pytest

from nbody6ppgpu_amd import gpunb, ics
from oracle import oracle
from oracle.parity import FORCE_TOL, assert_parity, jerk_scale, rel_err  # noqa: F401

pytestmark = pytest.mark.gpu


def gpu_regf(m, xj, vj, h2, dtr, xi, vi, lmax, nnbmax, m_flag=0, nbmax=None):
    gpunb.gpunb_open(nbmax or len(m) + 10)
    try:
        gpunb.gpunb_send(m, xj, vj)
        return gpunb.gpunb_regf(h2, dtr, xi, vi, lmax, nnbmax, m_flag,
                                nearest=True)
    finally:
        gpunb.gpunb_close()


def setup(c, k=48, m_flag=0):
    rs, stepr = ics.neighbour_radii(c, k=k)
    h2 = rs * rs
    if m_flag:
        h2 = h2 / np.mean(c.m)
    return h2, stepr


@pytest.mark.parametrize("m_flag", [0, 1])
def test_plummer_2048_one_call(m_flag):
    """Plummer N=2048, one gpunb_regf_ of all 2048 i, <nnb> ~ 50, nnbmax 500."""
    c = ics.plummer(2048, seed=101)
    if m_flag:
        c = ics.plummer(2048, seed=101, m=ics.kroupa(2048, seed=102))
    h2, dtr = setup(c, 48, m_flag)
    args = (c.m, c.x, c.v, h2, dtr, c.x, c.v, 501, 500, m_flag)
    g = gpu_regf(*args)
    o = oracle.regf(*args)
    assert 40 < o[3][:, 0].mean() < (70 if m_flag == 0 else 120)
    assert_parity(g, o, label=f"N2048 m_flag={m_flag}", inputs=args[:3] + args[5:7])


@pytest.mark.parametrize("nj,ni", [(1, 1), (63, 65), (64, 64), (65, 63),
                                   (1000, 1), (1000, 1000), (130, 257)])
def test_ragged_sizes(nj, ni):
    c = ics.plummer(max(nj, ni), seed=103 + nj)
    h2, dtr = setup(c, min(8, max(nj - 2, 1)))
    sl = slice(0, nj)
    args = (c.m[sl], c.x[sl], c.v[sl], h2[:ni], dtr[:ni], c.x[:ni], c.v[:ni],
            64, 40, 0)
    assert_parity(gpu_regf(*args), oracle.regf(*args), label=f"{nj}x{ni}",
                  inputs=args[:3] + args[5:7])


def test_overflow_codes_and_all_neighbours():
    c = ics.plummer(700, seed=104)
    h2, dtr = setup(c, 48)
    # nnbmax below the typical count: many overflows (-nnb), some not.
    args = (c.m, c.x, c.v, h2, dtr, c.x, c.v, 61, 50, 0)
    g = gpu_regf(*args)
    o = oracle.regf(*args)
    assert (o[3][:, 0] < 0).any() and (o[3][:, 0] >= 0).any()
    assert_parity(g, o, label="overflow", inputs=args[:3] + args[5:7])
    # Huge radius: every j is a neighbour, the far field is empty.  The
    # lists (700 x 700 entries) outgrow the speculative first copy of the
    # compact list, so the rest comes back in a second copy.
    big = np.full(c.n, 1e6)
    args = (c.m, c.x, c.v, big, dtr, c.x, c.v, 1001, 1000, 0)
    g = gpu_regf(*args)
    o = oracle.regf(*args)
    assert np.array_equal(g[3][:, 0], o[3][:, 0])
    assert np.array_equal(g[3][:, 1:c.n + 1], o[3][:, 1:c.n + 1])
    assert np.all(g[3][:, 0] == c.n)
    assert np.all(g[2] == 0.0) and np.all(g[0] == 0.0)


def test_zero_radius_and_lmax_stride():
    c = ics.plummer(500, seed=105)
    z = np.zeros(c.n)
    args = (c.m, c.x, c.v, z, z, c.x, c.v, 7, 3, 0)
    g = gpu_regf(*args)
    assert np.all(g[3][:, 0] == 0) and np.all(np.isfinite(g[0]))
    # h2 = 0: no neighbour sphere anchors the FP64 near band, the closest
    # pairs are in the far field: jerk at the condition-scaled bar.
    assert_parity(g, oracle.regf(*args), label="h2=0", inputs=args[:3] + args[5:7],
                  jerk_plain=False)
    h2, dtr = setup(c, 20)
    args = (c.m, c.x, c.v, h2, dtr, c.x, c.v, 200, 60, 0)
    assert_parity(gpu_regf(*args), oracle.regf(*args), label="stride",
                  inputs=args[:3] + args[5:7])


def test_zero_masses_and_coincident_particles():
    c = ics.plummer(400, seed=106)
    m = c.m.copy()
    m[::7] = 0.0                       # massless ghosts (merged-binary c.m.)
    x = c.x.copy()
    x[5] = x[6]                        # two bodies at the same point
    v = c.v.copy()
    for m_flag in (0, 1):
        h2, dtr = setup(ics.Cluster(m, x, v), 24, 0)
        if m_flag:
            h2 = h2 / np.mean(m)
        args = (m, x, v, h2, dtr, x, v, 200, 150, m_flag)
        g = gpu_regf(*args)
        assert np.all(np.isfinite(g[0])) and np.all(np.isfinite(g[1]))
        assert_parity(g, oracle.regf(*args), label=f"ghosts m_flag={m_flag}",
                      inputs=args[:3] + args[5:7])


def test_resend_and_repeated_calls():
    c = ics.plummer(3000, seed=107)
    h2, dtr = setup(c, 40)
    gpunb.gpunb_open(c.n + 10)
    try:
        for nj in (3000, 1500, 2999, 77):
            gpunb.gpunb_send(c.m[:nj], c.x[:nj], c.v[:nj])
            for ni in (512, 100):
                g = gpunb.gpunb_regf(h2[:ni], dtr[:ni], c.x[:ni], c.v[:ni],
                                     301, 300, 0, nearest=True)
                o = oracle.regf(c.m[:nj], c.x[:nj], c.v[:nj], h2[:ni], dtr[:ni],
                                c.x[:ni], c.v[:ni], 301, 300, 0)
                assert_parity(g, o, label=f"resend nj={nj} ni={ni}",
                              inputs=(c.m[:nj], c.x[:nj], c.v[:nj], c.x[:ni],
                                      c.v[:ni]))
    finally:
        gpunb.gpunb_close()


def test_group_variants_agree():
    c = ics.king(5000, 6.0, seed=108)
    h2, dtr = setup(c, 48)
    args = (c.m, c.x, c.v, h2, dtr, c.x, c.v, 501, 500, 0)
    o = oracle.regf(*args)
    try:
        for grp in (2, 4):
            gpunb.gpunb_set_tuning(grp, 0.0)
            assert_parity(gpu_regf(*args), o, label=f"group {grp}",
                          inputs=args[:3] + args[5:7])
        for rounds in (0.5, 1.0, 4.0):
            gpunb.gpunb_set_tuning(0, rounds)
            assert_parity(gpu_regf(*args), o, label=f"rounds {rounds}",
                          inputs=args[:3] + args[5:7])
    finally:
        gpunb.gpunb_set_tuning(0, 0.0)


@pytest.mark.parametrize("prec", [32, 64, 65])
@pytest.mark.parametrize("m_flag", [0, 1])
def test_precision_modes(prec, m_flag):
    """Every pair-arithmetic mode meets the bar; the FP64 + Newton mode (65)
    matches the oracle to ~1e-13 on every output of every particle."""
    c = ics.king(6000, 6.0, seed=110, m=ics.kroupa(6000, seed=111))
    h2, dtr = setup(c, 48, m_flag)
    args = (c.m, c.x, c.v, h2, dtr, c.x, c.v, 501, 500, m_flag)
    o = oracle.regf(*args)
    try:
        gpunb.gpunb_set_precision(prec)
        g = gpu_regf(*args)
    finally:
        gpunb.gpunb_set_precision(0)
    assert_parity(g, o, label=f"prec {prec} m_flag={m_flag}",
                  inputs=args[:3] + args[5:7])
    if prec == 65:
        ok = o[3][:, 0] >= 0
        worst = max(rel_err(g[k][ok], o[k][ok]).max() for k in range(3))
        print(f"prec {prec}: max per-particle relative error {worst:.2e}")
        assert worst < 1e-11


def test_nan_input_is_an_error():
    c = ics.plummer(256, seed=109)
    h2, dtr = setup(c, 16)
    x = c.x.copy()
    x[3, 1] = np.nan
    gpunb.gpunb_open(c.n + 10)
    try:
        gpunb.gpunb_send(c.m, c.x, c.v)
        with pytest.raises(gpunb.GpunbError, match="NaN"):
            gpunb.gpunb_regf(h2, dtr, x, c.v, 64, 60, 0)
        gpunb.gpunb_send(c.m, x, c.v)
        with pytest.raises(gpunb.GpunbError, match="NaN"):
            gpunb.gpunb_regf(h2, dtr, c.x, c.v, 64, 60, 0)
        # The library stays usable after the error.
        gpunb.gpunb_send(c.m, c.x, c.v)
        g = gpunb.gpunb_regf(h2, dtr, c.x, c.v, 64, 60, 0, nearest=True)
        assert_parity(g, oracle.regf(c.m, c.x, c.v, h2, dtr, c.x, c.v, 64, 60),
                      label="after NaN", inputs=(c.m, c.x, c.v, c.x, c.v))
    finally:
        gpunb.gpunb_close()


def test_precondition_errors():
    gpunb.gpunb_open(100)
    try:
        with pytest.raises(gpunb.GpunbError, match="nbmax"):
            gpunb.gpunb_send(np.ones(200), np.zeros((200, 3)), np.zeros((200, 3)))
        with pytest.raises(gpunb.GpunbError, match="lmax"):
            gpunb.gpunb_regf(np.ones(4), np.ones(4), np.zeros((4, 3)),
                             np.zeros((4, 3)), 10, 10, 0)
    finally:
        gpunb.gpunb_close()
    with pytest.raises(gpunb.GpunbError, match="not open"):
        gpunb.gpunb_regf(np.ones(4), np.ones(4), np.zeros((4, 3)),
                         np.zeros((4, 3)), 11, 10, 0)
