"""Parity checker of the engine's regular-force outputs against the oracle's.

TEST INFRASTRUCTURE ONLY (used by tests/ and __graft_entry__.smoke()).

Bar: neighbour lists and nearest-neighbour indices identical to the oracle's
(bit-exact, including order and overflow codes); acceleration, jerk and
potential within FORCE_TOL = 1e-6 relative error per particle, normwise
(|delta_i| / |oracle_i|).

The condition-scaled jerk error |delta_i| / max(|jerk_i|, J_i / KAPPA), with
J_i the sum over the far-field pairs of |jerk term|, is reported next to it:
a jerk sum can cancel (J / |jerk| reaches ~160 in these clusters), which is
why the engine evaluates the pairs of the near band in FP64.  Only a call
whose i have no neighbour sphere to anchor that band (h2 = 0, where the far
field includes the closest pairs) is held to the condition-scaled bar
instead, with jerk_plain=False.
"""
import numpy as np

FORCE_TOL = 1e-6
KAPPA = 2.0


def rel_err(a, b):
    a = np.atleast_2d(a.T).T
    b = np.atleast_2d(b.T).T
    num = np.linalg.norm(np.reshape(a - b, (len(a), -1)), axis=1)
    den = np.linalg.norm(np.reshape(b, (len(b), -1)), axis=1)
    # An empty far field gives exact zeros on both sides.
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), num)


def jerk_scale(m, xj, vj, xi, vi, lst):
    """J_i = sum over non-neighbours j (r2 > 0) of |m_j (dv - 3 (dx.dv) dx / r2) / r^3|."""
    out = np.zeros(len(xi))
    for i in range(len(xi)):
        dx = xj - xi[i]
        dv = vj - vi[i]
        keep = np.ones(len(m), bool)
        k = lst[i, 0]
        if k > 0:
            keep[lst[i, 1:1 + k]] = False
        r2 = np.sum(dx * dx, 1)
        keep &= r2 > 0
        dx, dv, r2, mm = dx[keep], dv[keep], r2[keep], m[keep]
        rv = np.sum(dx * dv, 1)
        t = (mm / (r2 * np.sqrt(r2)))[:, None] * (dv - (3 * rv / r2)[:, None] * dx)
        out[i] = np.sum(np.linalg.norm(t, axis=1))
    return out


def assert_parity(g, o, tol=FORCE_TOL, label="", inputs=None,
                  jerk_plain=True):
    """inputs = (m, xj, vj, xi, vi) of the call, for the condition-scaled
    jerk report (and the bar when jerk_plain is False)."""
    acc, jrk, pot, lst, nnid = g
    oacc, ojrk, opot, olst, onnid = o
    assert np.array_equal(lst[:, 0], olst[:, 0]), f"{label}: counts differ"
    for i in range(lst.shape[0]):
        k = lst[i, 0]
        if k > 0:
            assert np.array_equal(lst[i, 1:1 + k], olst[i, 1:1 + k]), \
                f"{label}: list of i={i}"
    assert np.array_equal(nnid, onnid), f"{label}: nearest neighbours differ"
    ok = olst[:, 0] >= 0
    if not ok.any():
        return
    ea = rel_err(acc[ok], oacc[ok])
    ej = rel_err(jrk[ok], ojrk[ok])
    ep = rel_err(pot[ok], opot[ok])
    msg = (f"{label}: rel err acc max {ea.max():.2e} med {np.median(ea):.2e}; "
           f"jrk max {ej.max():.2e} med {np.median(ej):.2e}; "
           f"pot max {ep.max():.2e}")
    ejs = None
    if inputs is not None:
        m, xj, vj, xi, vi = inputs
        J = jerk_scale(m, xj, vj, xi[ok], vi[ok], olst[ok])
        nj = np.linalg.norm(ojrk[ok], axis=1)
        dj = np.linalg.norm(jrk[ok] - ojrk[ok], axis=1)
        scale = np.maximum(nj, J / KAPPA)
        ejs = np.where(scale > 0, dj / np.where(scale > 0, scale, 1.0), dj)
        msg += (f"; jrk vs max(|j|,J/{KAPPA:g}) max {ejs.max():.2e}; "
                f"cancellation J/|j| max {np.max(J / np.maximum(nj, 1e-300)):.1f}")
    print(msg)
    assert ea.max() < tol, f"{label}: acc {ea.max()}"
    assert ep.max() < tol, f"{label}: pot {ep.max()}"
    if jerk_plain:
        assert ej.max() < tol, f"{label}: jrk {ej.max()}"
    else:
        assert ejs is not None, "the condition-scaled bar needs the inputs"
        assert ejs.max() < tol, f"{label}: jrk (condition-scaled) {ejs.max()}"
