"""Pins of the FP64 oracle by closed forms, invariants and brute force.

Each check is chosen so that a plausible mistake in oracle.c (a dropped term,
a wrong sign, a swapped i/j, a wrong list index or overflow code) fails it;
none retypes the oracle's own formula.
"""
import numpy as np
import pytest

from nbody6ppgpu_amd import ics
from oracle import oracle


def two_body(d=0.5, m=0.3, vj=(0.0, 0.0, 0.0)):
    xj = np.array([[d, 0.0, 0.0]])
    vj = np.array([vj], dtype=float)
    xi = np.zeros((1, 3))
    vi = np.zeros((1, 3))
    return np.array([m]), xj, vj, xi, vi


def test_two_body_closed_form():
    # a = m/d^2 along +x, phi = m/d; jerk = m (v_perp - 2 v_par)/d^3.
    d, m = 0.5, 0.3
    vj = (0.4, -0.7, 0.2)
    mj, xj, vjj, xi, vi = two_body(d, m, vj)
    acc, jrk, pot, lst, nnid = oracle.regf(mj, xj, vjj, [0.0], [0.0], xi, vi,
                                           4, 2)
    assert np.allclose(acc[0], [m / d**2, 0, 0], rtol=1e-15, atol=0)
    assert pot[0] == pytest.approx(m / d, rel=1e-15)
    expect = m / d**3 * np.array([-2 * vj[0], vj[1], vj[2]])
    assert np.allclose(jrk[0], expect, rtol=1e-14, atol=1e-15)
    assert lst[0, 0] == 0 and nnid[0] == -1


def test_jerk_is_time_derivative_of_acc():
    c = ics.plummer(64, seed=3)
    h2 = np.zeros(c.n)
    eps = 1e-6
    a_p = oracle.regf(c.m, c.x + eps * c.v, c.v, h2, h2, c.x + eps * c.v, c.v,
                      2, 1)[0]
    a_m = oracle.regf(c.m, c.x - eps * c.v, c.v, h2, h2, c.x - eps * c.v, c.v,
                      2, 1)[0]
    _, jrk, _, _, _ = oracle.regf(c.m, c.x, c.v, h2, h2, c.x, c.v, 2, 1)
    fd = (a_p - a_m) / (2 * eps)
    err = np.linalg.norm(fd - jrk, axis=1) / np.linalg.norm(jrk, axis=1)
    assert np.max(err) < 1e-5


def test_acc_is_gradient_of_potential():
    # acc_i = grad_i sum_j m_j/|x_j - x_i| (sign and scale of the pair force).
    c = ics.plummer(48, seed=4)
    h2 = np.zeros(c.n)
    acc = oracle.regf(c.m, c.x, c.v, h2, h2, c.x, c.v, 2, 1)[0]
    eps = 1e-6
    for i in (0, 7, 31):
        g = np.zeros(3)
        keep = np.arange(c.n) != i  # the probe moves i: drop it from the j set
        mj, xj, vj = c.m[keep], c.x[keep], c.v[keep]
        for k in range(3):
            xp = c.x[i].copy()
            xp[k] += eps
            xm = c.x[i].copy()
            xm[k] -= eps
            pp = oracle.regf(mj, xj, vj, [0.0], [0.0], xp[None], c.v[i:i+1],
                             2, 1)[2][0]
            pm = oracle.regf(mj, xj, vj, [0.0], [0.0], xm[None], c.v[i:i+1],
                             2, 1)[2][0]
            g[k] = (pp - pm) / (2 * eps)
        assert np.allclose(g, acc[i], rtol=1e-6, atol=1e-9)


def test_newton_third_law_without_neighbours():
    c = ics.plummer(200, seed=5, m=ics.kroupa(200, seed=6))
    h2 = np.zeros(c.n)
    acc, jrk, pot, lst, _ = oracle.regf(c.m, c.x, c.v, h2, h2, c.x, c.v, 2, 1)
    scale = np.sum(c.m[:, None] * np.abs(acc))
    assert np.all(np.abs(np.sum(c.m[:, None] * acc, 0)) < 1e-14 * scale)
    scale = np.sum(c.m[:, None] * np.abs(jrk))
    assert np.all(np.abs(np.sum(c.m[:, None] * jrk, 0)) < 1e-14 * scale)
    # Symmetric pair sum of the potential energy.
    assert np.sum(c.m * pot) == pytest.approx(
        np.sum(c.m * oracle.pot(1, c.n, c.m, c.x)), rel=1e-14)


def test_long_double_brute_force():
    c = ics.plummer(300, seed=7)
    rs, stepr = ics.neighbour_radii(c, k=20)
    h2 = rs * rs
    acc, jrk, pot, lst, _ = oracle.regf(c.m, c.x, c.v, h2, stepr, c.x, c.v,
                                        64, 60)
    mask = np.zeros((c.n, c.n), dtype=np.uint8)
    for i in range(c.n):
        mask[i, lst[i, 1:1 + lst[i, 0]]] = 1
    la, lj, lp = oracle.regf_longdouble(c.m, c.x, c.v, c.x, c.v, mask)
    assert np.max(np.linalg.norm(acc - la, axis=1) / np.linalg.norm(la, axis=1)) < 1e-13
    assert np.max(np.linalg.norm(jrk - lj, axis=1) / np.linalg.norm(lj, axis=1)) < 1e-12
    assert np.max(np.abs(pot - lp) / lp) < 1e-13


def test_neighbour_predicate_strict_and_predicted_distance():
    # Dyadic values make r2 and r2p exact: the boundary is strict <.
    xi = np.zeros(3)
    vi = np.zeros(3)
    xj = np.array([0.5, 0.0, 0.0])
    nb, r2 = oracle.pair_is_nb(xi, vi, 0.25, 0.0, xj, vi, 1.0)
    assert r2 == 0.25 and not nb
    nb, _ = oracle.pair_is_nb(xi, vi, np.nextafter(0.25, 1), 0.0, xj, vi, 1.0)
    assert nb
    # Approaching j: r2p = (0.5 - 0.25*1)^2 < h2 = 0.09 < r2.
    nb, _ = oracle.pair_is_nb(xi, vi, 0.09, 0.25, xj, [-1.0, 0, 0], 1.0)
    assert nb
    # Receding j: both distances above h2.
    nb, _ = oracle.pair_is_nb(xi, vi, 0.09, 0.25, xj, [1.0, 0, 0], 1.0)
    assert not nb
    # Mass-weighted criterion: r2 < m_j * h2.
    assert oracle.pair_is_nb(xi, vi, 0.2, 0.0, xj, vi, 2.0, 1)[0]
    assert not oracle.pair_is_nb(xi, vi, 0.2, 0.0, xj, vi, 1.0, 1)[0]


def test_lists_sorted_self_included_and_brute_force_membership():
    c = ics.plummer(256, seed=8)
    rs, stepr = ics.neighbour_radii(c, k=16)
    h2 = rs * rs
    _, _, _, lst, nnid = oracle.regf(c.m, c.x, c.v, h2, stepr, c.x, c.v, 101, 100)
    for i in range(c.n):
        row = lst[i, 1:1 + lst[i, 0]]
        assert np.all(np.diff(row) > 0)
        assert i in row
        d = c.x - c.x[i]
        p = d + stepr[i] * (c.v - c.v[i])
        mn = np.minimum(np.sum(d * d, 1), np.sum(p * p, 1))
        assert set(row) == set(np.nonzero(mn < h2[i])[0])
        # Nearest neighbour: argmin over the list without self.
        others = [j for j in row if j != i]
        if others:
            r2 = np.sum((c.x[others] - c.x[i]) ** 2, 1)
            assert nnid[i] == others[int(np.argmin(r2))]
        else:
            assert nnid[i] == -1
        assert oracle.nearest(c.x[i], row, c.x) == nnid[i]


def test_overflow_code_is_minus_count():
    c = ics.plummer(128, seed=9)
    h2 = np.full(c.n, 100.0)  # everybody is a neighbour of everybody
    _, _, pot, lst, nnid = oracle.regf(c.m, c.x, c.v, h2, h2 * 0, c.x, c.v, 11, 10)
    assert np.all(lst[:, 0] == -c.n)
    assert np.all(nnid == -1)
    assert np.all(pot == 0.0)


def test_zero_radius_excludes_self_from_force():
    c = ics.plummer(64, seed=10)
    z = np.zeros(c.n)
    acc, _, pot, lst, _ = oracle.regf(c.m, c.x, c.v, z, z, c.x, c.v, 2, 1)
    assert np.all(lst[:, 0] == 0) and np.all(np.isfinite(acc))
    assert np.allclose(pot, oracle.pot(1, c.n, c.m, c.x), rtol=1e-14)


def test_predictor_exact_on_cubics():
    # Dyadic coefficients and times: every operation is exact in FP64, so the
    # prediction must equal the cubic x0 + v0 s + f s^2 + fd s^3 exactly.
    rng = np.random.default_rng(1)
    n = 50
    q = lambda *s: rng.integers(-64, 64, size=s) / 16.0
    x0, v0, f, fd = q(n, 3), q(n, 3), q(n, 3), q(n, 3)
    t0 = rng.integers(0, 8, n) / 8.0
    t = 1.25
    x, v = oracle.predict(t, t0, x0, v0, f, fd)
    s = (t - t0)[:, None]
    assert np.array_equal(x, x0 + v0 * s + f * s**2 + fd * s**3)
    assert np.array_equal(v, v0 + 2 * f * s + 3 * fd * s**2)


def test_potential_closed_forms():
    # Two bodies; a regular tetrahedron of edge 1.
    x = np.array([[0.0, 0, 0], [0.0, 2.0, 0]])
    m = np.array([0.25, 0.75])
    p = oracle.pot(1, 2, m, x)
    assert p[0] == 0.75 / 2 and p[1] == 0.25 / 2
    t = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(8)
    p = oracle.pot(1, 4, np.ones(4), t)
    assert np.allclose(p, 3.0, rtol=1e-15)
    # istart/ni window selects rows.
    c = ics.plummer(40, seed=2)
    full = oracle.pot(1, c.n, c.m, c.x)
    assert np.array_equal(oracle.pot(11, 7, c.m, c.x), full[10:17])
