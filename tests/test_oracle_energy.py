"""Pins of the energy-sum restatement (oracle/energy.c: energy.F:37-79 with
phicor.f:1-112) against physics, not against a retyping of its formulas:

* without KS pairs, POT is the pairwise potential energy and the
  corrections vanish;
* with KS pairs, complete lists and an unlimited resolution radius, the
  corrected POT is the potential energy of singles and components with each
  pair's bond left out (the c.m. approximation fully undone);
* with a finite radius, a c.m. neighbour inside it is resolved and one
  outside it is not, each by the exact potential difference;
* ZKIN = 1/2 sum m v^2 over IFIRST..NTOT.
"""
import numpy as np
import pytest

from oracle import energy as orc
from tests import energy_cases as ec


def run(s, cmsep2_rmin2, phi=None):
    if phi is None:
        phi = ec.gpupot_phidbl(s)
    return orc.energy(s["n"], s["ntot"], s["lmax"], cmsep2_rmin2, s["body"],
                      s["x"], s["xdot"], s["list"], phi)


def test_no_pairs_is_the_pairwise_potential():
    s = ec.build(nsingle=500, npairs=0, seed=5)
    phi0 = ec.gpupot_phidbl(s)
    phi, pot, zkin = run(s, 1e30, phi0)
    assert np.array_equal(phi, phi0)
    assert pot == pytest.approx(ec.resolved_potential(s), rel=1e-13)
    m, v = s["body"], s["xdot"]
    assert zkin == pytest.approx(0.5 * np.sum(m * np.sum(v * v, axis=1)),
                                 rel=1e-14)


def test_complete_lists_undo_the_cm_approximation():
    s = ec.build(nsingle=300, npairs=30, seed=6, complete=True)
    phi, pot, zkin = run(s, 1e30)
    assert pot == pytest.approx(ec.resolved_potential(s), rel=1e-12)
    # the pair terms carry the components' own potentials
    N, npairs = s["n"], s["npairs"]
    x, m = s["x"], s["body"]
    for p in (0, npairs - 1):
        for c in (2 * p, 2 * p + 1):
            others = [j for j in range(N) if j not in (2 * p, 2 * p + 1)]
            d = x[others] - x[c]
            want = np.sum(m[others] / np.sqrt(np.sum(d * d, axis=1)))
            assert phi[c] == pytest.approx(want, rel=1e-12)
    # the c.m. approximation is not exact: the correction is visible
    assert abs(pot - 0.5 * sum(m[i] * ec.gpupot_phidbl(s)[i]
                               for i in range(s["ifirst"] - 1, s["ntot"]))) > 0


def test_resolution_radius_cuts_where_it_should():
    s = ec.build(nsingle=400, npairs=20, seed=7, rnb=0.6)
    N, x, m = s["n"], s["x"], s["body"]
    # a single with at least one c.m. neighbour
    for i in range(s["ifirst"] - 1, N):
        row = s["list"][i, 1:1 + s["list"][i, 0]]
        cms = [j for j in row if j > N]
        if len(cms) >= 2:
            break
    r2 = sorted((np.sum((x[i] - x[j - 1]) ** 2), j) for j in cms)
    cut = 0.5 * (r2[0][0] + r2[1][0])          # first inside, second outside
    phi0 = ec.gpupot_phidbl(s)
    phi, _, _ = run(s, cut, phi0)
    j = r2[0][1]
    j1, j2 = 2 * (j - N) - 2, 2 * (j - N) - 1   # 0-based components
    want = (-m[j - 1] / np.sqrt(r2[0][0]) +
            m[j1] / np.linalg.norm(x[i] - x[j1]) +
            m[j2] / np.linalg.norm(x[i] - x[j2]))
    assert phi[i] - phi0[i] == pytest.approx(want, rel=1e-9, abs=1e-14)


def test_pair_with_one_single_perturber():
    s = ec.build(nsingle=50, npairs=1, seed=8)
    N, x, m = s["n"], s["x"], s["body"]
    J = s["ifirst"] + 3                          # a single (1-based)
    s["list"][0, 0] = 1
    s["list"][0, 1] = J
    s["list"][2:N, 0] = 0                        # no single lists a c.m.
    phi0 = ec.gpupot_phidbl(s)
    phi, _, _ = run(s, 1e30, phi0)
    icm = N                                      # 0-based c.m. of pair 1
    rcm = np.linalg.norm(x[icm] - x[J - 1])
    for c in (0, 1):
        want = phi0[icm] - m[J - 1] / rcm + m[J - 1] / np.linalg.norm(x[c] - x[J - 1])
        assert phi[c] == pytest.approx(want, rel=1e-13)


def test_invalid_row_is_an_error():
    s = ec.build(nsingle=50, npairs=2, seed=9)
    s["list"][10, 0] = s["lmax"] + 5
    with pytest.raises(ValueError):
        run(s, 1.0)
