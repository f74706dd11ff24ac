"""Pins of the irregular-force oracle (oracle.irr, FP64 restatement of
irr.avx.cpp:406-438): a closed form, Newton's third law over symmetric lists,
and the reference's own AVX library compiled from its sources (FP32 pair
arithmetic on double-single positions, so agreement is to FP32 accuracy and
the nearest neighbour agrees wherever the two closest members are separated
by more than the FP32 resolution of r^2)."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from nbody6ppgpu_amd import ics
from oracle import oracle, ref


def state(n, seed):
    rng = np.random.default_rng(seed)
    c = ics.plummer(n, seed=seed, m=ics.kroupa(n, seed=seed + 1))
    # Predictor state at slightly different times (t0 <= t).
    t0 = -rng.uniform(0, 1 / 256, n)
    f = rng.normal(0, 0.5, (n, 3))          # a/2
    fd = rng.normal(0, 2.0, (n, 3))         # jerk/6
    return c.m, c.x, c.v, f, fd, t0


def knn_lists(x, k):
    _, idx = cKDTree(x).query(x, k=k + 1)
    return [list(np.sort(r[1:])) for r in idx]


def test_closed_form_single_neighbour():
    # i at the origin, j on the x axis at d, relative velocity (u, w, 0).
    d, mj, u, w = 0.05, 3e-4, 0.3, -0.7
    x0 = np.array([[0.0, 0, 0], [d, 0, 0]])
    v0 = np.array([[0.0, 0, 0], [u, w, 0]])
    z = np.zeros((2, 3))
    m = np.array([1e-4, mj])
    acc, jrk, nn = oracle.irr(0.0, [0], np.zeros(2), x0, v0, z, z, m, [[1], []])
    assert np.allclose(acc[0], [mj / d ** 2, 0, 0], rtol=1e-14, atol=0)
    # jerk = m (v_perp - 2 v_par) / d^3
    assert np.allclose(jrk[0], [mj * (-2 * u) / d ** 3, mj * w / d ** 3, 0],
                       rtol=1e-14, atol=0)
    assert nn[0] == 2
    acc, jrk, nn = oracle.irr(0.0, [1], np.zeros(2), x0, v0, z, z, m, [[1], []])
    assert nn[0] == 0 and not acc.any() and not jrk.any()


def test_newton_third_law_on_symmetric_lists():
    m, x, v, f, fd, t0 = state(600, 3)
    L = knn_lists(x, 30)
    sym = [set(l) for l in L]
    for i, l in enumerate(L):
        for j in l:
            sym[j].add(i)
    lists = [sorted(s) for s in sym]
    acc, jrk, _ = oracle.irr(0.01, np.arange(600), t0, x, v, f, fd, m, lists)
    scale = np.sum(m[:, None] * np.abs(acc))
    assert np.all(np.abs(np.sum(m[:, None] * acc, axis=0)) < 1e-13 * scale)
    jscale = np.sum(m[:, None] * np.abs(jrk))
    assert np.all(np.abs(np.sum(m[:, None] * jrk, axis=0)) < 1e-13 * jscale)


@pytest.mark.skipif(not ref.available("irr_avx"),
                    reason="reference AVX irregular library not built")
def test_matches_reference_avx_library():
    n = 3000
    m, x, v, f, fd, t0 = state(n, 5)
    lists = knn_lists(x, 48)
    lists[17] = []                          # an empty list
    R = ref.IrrLib("irr_avx", n, 601)
    try:
        R.set_jp(t0, x, v, f, fd, m)
        for i, l in enumerate(lists):
            R.set_list(i, l)
        idx = np.arange(0, n, 3)
        ra, rj, rn = R.firr(0.002, idx)
    finally:
        R.close()
    oa, oj, on = oracle.irr(0.002, idx, t0, x, v, f, fd, m, lists)
    # FP32 pair sums: per-particle relative error against the FP64 oracle.
    rel_a = np.linalg.norm(ra - oa, axis=1) / np.maximum(np.linalg.norm(oa, axis=1), 1e-300)
    assert np.median(rel_a) < 1e-6 and np.percentile(rel_a, 99) < 1e-4
    aj = np.linalg.norm(oj, axis=1)
    rel_j = np.linalg.norm(rj - oj, axis=1) / np.maximum(aj, 1e-300)
    assert np.median(rel_j) < 1e-5
    # Nearest neighbour: identical unless the two closest are FP32-tied.
    xp, _ = oracle.predict(0.002, t0, x, v, f, fd)
    for q, i in enumerate(idx):
        if not lists[i]:
            assert rn[q] == 0 and on[q] == 0
            continue
        r2 = np.sum((xp[lists[i]] - xp[i]) ** 2, axis=1)
        s = np.sort(r2)
        if len(s) > 1 and (s[1] - s[0]) < 1e-5 * s[0]:
            continue
        assert rn[q] == on[q], (i, rn[q], on[q])
