"""Synthetic star-cluster inputs for the regular-force path.

Everything is seeded and FP64, in Henon N-body units (G = 1, M = 1,
E = -1/4 for the virialised models):

* ``plummer``  -- Plummer sphere by the Aarseth-Henon-Wielen recipe (the same
  recipe as the reference's src/Main/setup.F:55-104), scaled analytically to
  E = -1/4 (scale radius 3*pi/16).
* ``plummer_segregated`` -- Plummer sphere with full radial mass segregation
  (masses ranked by binding energy), virial ratio 1/2, E = -1/4.
* ``king``     -- King (1966) model of central potential W0: the Poisson
  equation of the lowered isothermal distribution is integrated numerically,
  radii are drawn from the cumulative mass, speeds from the local distribution
  f ~ exp(W - v^2/2) - 1; scaled to E = -1/4 with the model's own potential
  energy (the reference reads King models from a file, data.F:31-38, and has
  no generator).
* ``kroupa``   -- Kroupa (2001) IMF, slopes 1.3 below and 2.3 above 0.5 Msun,
  between given bounds, rescaled to total mass 1.
* ``with_binaries`` -- a fraction of the stars paired into binaries that
  enter the regular force as centre-of-mass particles (intgrt.F:917-918 sends
  only c.m. bodies); semi-major axes log-uniform.
* ``neighbour_radii`` -- RS per particle such that the distance criterion
  alone holds ``k`` neighbours (midway between the k-th and (k+1)-th nearest
  distances, so no pair sits on the boundary), and STEPR as FPOLY0 sets it
  (fpoly0.F:53-55).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Cluster:
    m: np.ndarray            # (n,)
    x: np.ndarray            # (n, 3)
    v: np.ndarray            # (n, 3)
    info: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.m.shape[0]


def _isotropic(rng, n):
    z = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(1.0 - z * z)
    return np.stack([s * np.cos(ph), s * np.sin(ph), z], axis=1)


def plummer(n: int, seed: int = 1, m=None) -> Cluster:
    """Plummer sphere in Henon units (equal masses unless ``m`` is given)."""
    rng = np.random.default_rng(seed)
    if m is None:
        m = np.full(n, 1.0 / n)
    m = np.asarray(m, dtype=np.float64)
    m = m / m.sum()
    # Radii from the cumulative mass, truncated at 10 scale radii as
    # setup.F does (X1 < 0.9 mass fraction cut there; here r < 10).
    r = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        u = rng.uniform(1e-10, 1.0, todo.size)
        rr = 1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        ok = rr < 10.0
        r[todo[ok]] = rr[ok]
        todo = todo[~ok]
    x = r[:, None] * _isotropic(rng, n)
    # Speeds: q = v / v_esc from g(q) = q^2 (1 - q^2)^3.5 by rejection.
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        qq = rng.uniform(0.0, 1.0, todo.size)
        yy = rng.uniform(0.0, 0.1, todo.size)
        ok = yy < qq * qq * (1.0 - qq * qq) ** 3.5
        q[todo[ok]] = qq[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0) * (1.0 + r * r) ** -0.25
    v = (q * vesc)[:, None] * _isotropic(rng, n)
    # Scale radius a = 3 pi / 16 gives E = -1/4 for G = M = 1.
    a = 3.0 * np.pi / 16.0
    x *= a
    v /= np.sqrt(a)
    x -= (m[:, None] * x).sum(0)
    v -= (m[:, None] * v).sum(0)
    return Cluster(m, x, v, {"model": "plummer", "seed": seed})


def plummer_segregated(n: int, seed: int = 1, m=None) -> Cluster:
    """Plummer sphere with full radial mass segregation (Baumgardt, De Marchi
    & Kroupa 2008, ApJ 685, 247): phase-space points of an equal-mass
    Plummer model, masses (Kroupa IMF unless ``m`` is given) assigned in
    order of specific binding energy, the heaviest to the most bound.  Then
    velocities are scaled to virial ratio 1/2 and lengths to E = -1/4, with
    the potential energy from the shell theorem about the centre of mass
    (exact for a spherical distribution; its error for the sampled cluster is
    of the order of the sampling noise)."""
    c = plummer(n, seed=seed)
    if m is None:
        m = kroupa(n, seed=seed + 1)
    m = np.sort(np.asarray(m, dtype=np.float64))[::-1]
    m = m / m.sum()
    a = 3.0 * np.pi / 16.0
    r = np.sqrt(np.sum(c.x * c.x, axis=1))
    eps = 0.5 * np.sum(c.v * c.v, axis=1) - 1.0 / np.sqrt(r * r + a * a)
    order = np.argsort(eps, kind="stable")
    x, v = c.x[order].copy(), c.v[order].copy()
    x -= (m[:, None] * x).sum(0)
    v -= (m[:, None] * v).sum(0)
    r = np.sqrt(np.sum(x * x, axis=1))
    srt = np.argsort(r, kind="stable")
    rs_, ms_ = r[srt], m[srt]
    inner = np.cumsum(ms_) - ms_                  # mass strictly inside
    outer = np.cumsum((ms_ / rs_)[::-1])[::-1] - ms_ / rs_
    pe = -0.5 * np.sum(ms_ * (inner / rs_ + outer))   # 1/2 sum m_i phi_i
    ke = 0.5 * np.sum(m * np.sum(v * v, axis=1))
    v *= np.sqrt(0.5 * abs(pe) / ke)
    s = -2.0 * pe                                 # E = pe/2 -> -1/4
    x *= s
    v /= np.sqrt(s)
    return Cluster(m, x, v, {"model": "plummer_segregated", "seed": seed})


def _king_profile(w0: float, nr: int = 4000):
    """Integrate the King model: returns r, W(r), rho(r), M(r) (King units,
    r0 = 1, rho(W0) = 1), up to the tidal radius where W = 0."""
    from scipy.integrate import solve_ivp
    from scipy.special import erf

    def rho(w):
        w = np.maximum(w, 0.0)
        return np.exp(w) * erf(np.sqrt(w)) - np.sqrt(4.0 * w / np.pi) * (
            1.0 + 2.0 * w / 3.0)

    rho0 = rho(w0)

    def rhs(r, y):
        w, dw = y
        if r == 0.0:
            return [dw, -3.0 * rho(w) / rho0]
        return [dw, -9.0 * rho(w) / rho0 - 2.0 * dw / r]

    def hit(r, y):
        return y[0]
    hit.terminal = True
    hit.direction = -1
    r0 = 1e-6
    y0 = [w0 - 1.5 * r0 * r0, -3.0 * r0]
    sol = solve_ivp(rhs, (r0, 1e4), y0, events=hit, rtol=1e-10, atol=1e-12,
                    dense_output=True, max_step=0.05)
    rt = sol.t_events[0][0]
    r = np.linspace(0.0, rt, nr)
    r[0] = r0
    w = np.maximum(sol.sol(r)[0], 0.0)
    dens = 9.0 * rho(w) / rho0 / (4.0 * np.pi)  # 4 pi G rho = 9 rho/rho0
    shell = 4.0 * np.pi * r * r * dens
    mass = np.concatenate([[0.0], np.cumsum(0.5 * (shell[1:] + shell[:-1])
                                            * np.diff(r))])
    return r, w, dens, mass


def king(n: int, w0: float = 6.0, seed: int = 1, m=None) -> Cluster:
    """King model of central potential w0 in Henon units (E = -1/4)."""
    rng = np.random.default_rng(seed)
    if m is None:
        m = np.full(n, 1.0 / n)
    m = np.asarray(m, dtype=np.float64)
    m = m / m.sum()
    r_t, w_t, dens, mass = _king_profile(w0)
    mt = mass[-1]
    u = rng.uniform(0.0, mt, n)
    r = np.interp(u, mass, r_t)
    w = np.interp(r, r_t, w_t)
    # Speeds in units of the velocity-dispersion parameter (sigma = 1):
    # f(E) ~ exp(W - v^2/2) - 1, v < sqrt(2W); density in v ~ v^2 f.
    v = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        ww = w[todo]
        vmax = np.sqrt(2.0 * ww)
        vv = rng.uniform(0.0, 1.0, todo.size) * vmax
        fmax = vmax ** 2 * (np.exp(ww) - 1.0) + 1e-300
        y = rng.uniform(0.0, 1.0, todo.size) * fmax
        ok = y < vv ** 2 * (np.exp(ww - 0.5 * vv ** 2) - 1.0)
        v[todo[ok]] = vv[ok]
        todo = todo[~ok]
    x = r[:, None] * _isotropic(rng, n)
    vel = v[:, None] * _isotropic(rng, n)
    # Potential energy of the continuum model (King units, G = 1, mass mt):
    # phi(r) = -mt/rt - int_r^rt M(s)/s^2 ds, PE = 1/2 int phi dM.
    g = mass / np.maximum(r_t, 1e-12) ** 2
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (g[1:] + g[:-1])
                                           * np.diff(r_t))])
    phi = -mt / r_t[-1] - (cum[-1] - cum)
    dm = 4.0 * np.pi * r_t * r_t * dens
    pe_model = 0.5 * np.trapezoid(dm * phi, r_t)
    # Mass 1 at the same lengths: PE scales with mass^2.  A length factor
    # alpha turns PE1 into PE1/alpha = -1/2; velocities are then rescaled to
    # the exact virial KE = 1/4, so E = -1/4 up to sampling noise in PE.
    pe1 = pe_model / mt ** 2
    alpha = -2.0 * pe1
    x = x * alpha
    ke = 0.5 * np.sum(m * np.sum(vel * vel, axis=1))
    vel = vel * np.sqrt(0.25 / ke)
    x -= (m[:, None] * x).sum(0)
    vel -= (m[:, None] * vel).sum(0)
    return Cluster(m, x, vel, {"model": "king", "w0": w0, "seed": seed,
                               "rt_king_units": float(r_t[-1])})


def kroupa(n: int, mlo: float = 0.08, mhi: float = 150.0, seed: int = 2,
           normalise: bool = True) -> np.ndarray:
    """Kroupa (2001) IMF: dN/dm ~ m^-1.3 (m < 0.5), m^-2.3 (m >= 0.5)."""
    rng = np.random.default_rng(seed)
    mb = 0.5
    a1, a2 = 1.3, 2.3

    def seg(lo, hi, a):
        return (hi ** (1 - a) - lo ** (1 - a)) / (1 - a)
    lo1, hi1 = mlo, min(mb, mhi)
    lo2, hi2 = max(mb, mlo), mhi
    w1 = seg(lo1, hi1, a1) if hi1 > lo1 else 0.0
    # continuity at mb: c2 = mb^(a2-a1)
    w2 = mb ** (a2 - a1) * seg(lo2, hi2, a2) if hi2 > lo2 else 0.0
    p1 = w1 / (w1 + w2)
    u = rng.uniform(0.0, 1.0, n)
    first = rng.uniform(0.0, 1.0, n) < p1

    def inv(u, lo, hi, a):
        return (lo ** (1 - a) + u * (hi ** (1 - a) - lo ** (1 - a))) ** (
            1.0 / (1 - a))
    m = np.where(first, inv(u, lo1, hi1, a1), inv(u, lo2, hi2, a2))
    if normalise:
        m = m / m.sum()
    return m


def with_binaries(c: Cluster, fraction: float = 0.1, amin: float = 1e-4,
                  amax: float = 1e-2, seed: int = 3) -> Cluster:
    """Pair ``fraction`` of the stars into binaries; each pair becomes one
    centre-of-mass particle (mass m1 + m2, the c.m. position and velocity).
    Singles come first, c.m. bodies last, as in the reference's j order
    (IFIRST..N singles, N+1..NTOT c.m.)."""
    rng = np.random.default_rng(seed)
    n = c.n
    nb = int(round(fraction * n)) // 2 * 2
    idx = rng.permutation(n)
    pairs = idx[:nb].reshape(-1, 2)
    singles = np.sort(idx[nb:])
    m1, m2 = c.m[pairs[:, 0]], c.m[pairs[:, 1]]
    mc = m1 + m2
    # The c.m. takes the first member's phase-space point.
    xc = c.x[pairs[:, 0]]
    vc = c.v[pairs[:, 0]]
    semi = np.exp(rng.uniform(np.log(amin), np.log(amax), pairs.shape[0]))
    m = np.concatenate([c.m[singles], mc])
    x = np.concatenate([c.x[singles], xc])
    v = np.concatenate([c.v[singles], vc])
    info = dict(c.info)
    info.update({"npairs": int(pairs.shape[0]), "semi_major": semi})
    return Cluster(m, x, v, info)


def neighbour_radii(c: Cluster, k: int = 48, smax: float = 0.125):
    """RS with exactly k distance-criterion neighbours (self excluded) and
    STEPR = min(SMAX/8 sqrt(1 + r^2), SMAX) (fpoly0.F:53-55).
    Returns (rs, stepr)."""
    from scipy.spatial import cKDTree
    tree = cKDTree(c.x)
    kk = min(k + 2, c.n)
    d, _ = tree.query(c.x, k=kk)
    d = np.reshape(d, (c.n, -1))
    if kk >= k + 2:
        rs = 0.5 * (d[:, k] + d[:, k + 1])
    else:
        rs = d[:, -1] * 1.01 + 1e-12
    r = np.sqrt(np.sum(c.x * c.x, axis=1))
    stepr = np.minimum(smax / 8.0 * np.sqrt(1.0 + r * r), smax)
    return rs, stepr


def energy(m, x, v, pot_positive):
    """Total energy from FP64 potentials phi_i = sum m_j / r_ij (positive)."""
    ke = 0.5 * np.sum(m * np.sum(v * v, axis=1))
    pe = -0.5 * np.sum(m * pot_positive)
    return ke + pe, ke, pe


def ks_system(nsingle=600, npairs=40, seed=3, rnb=0.3, rpert=0.4, lmax=None,
              complete=False, sep=2e-3):
    """A Plummer system of nsingle singles and npairs KS pairs in NBODY6++'s
    body order: components 1..2 NPAIRS (pair p: 2p-1, 2p, at separation
    ~sep about their c.m.), singles 2 NPAIRS+1..N, c.m. bodies N+1..NTOT.
    LIST rows as common6.h holds them: a single's neighbours within rnb
    (singles and c.m. bodies, ascending), pair p's perturbers within rpert of
    its c.m. in row 2p-1 (complete=True: every body of IFIRST..NTOT); other
    rows empty.  Returns dict(n, ntot, ifirst, npairs, lmax, body, x, xdot,
    list)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    nsys = nsingle + npairs
    c = plummer(nsys, seed=seed)
    N = 2 * npairs + nsingle
    ntot = N + npairs
    ifirst = 2 * npairs + 1
    body = np.zeros(ntot)
    x = np.zeros((ntot, 3))
    v = np.zeros((ntot, 3))
    # systems 0..npairs-1 are pairs, the rest singles
    for p in range(npairs):
        mc, xc, vc = c.m[p], c.x[p], c.v[p]
        q = rng.uniform(0.2, 0.8)
        m1, m2 = q * mc, (1 - q) * mc
        d = rng.normal(size=3)
        d *= sep * rng.uniform(0.5, 1.5) / np.linalg.norm(d)
        u = rng.normal(size=3) * 0.5
        i1, i2, icm = 2 * p, 2 * p + 1, N + p
        body[i1], body[i2], body[icm] = m1, m2, mc
        x[i1], x[i2], x[icm] = xc + (m2 / mc) * d, xc - (m1 / mc) * d, xc
        v[i1], v[i2], v[icm] = vc + (m2 / mc) * u, vc - (m1 / mc) * u, vc
    for s in range(nsingle):
        i = ifirst - 1 + s
        body[i], x[i], v[i] = c.m[npairs + s], c.x[npairs + s], c.v[npairs + s]
    # bodies that take part in the potential: IFIRST..NTOT (0-based ifirst-1..)
    active = np.arange(ifirst - 1, ntot)
    tree = cKDTree(x[active])
    rows = [[] for _ in range(ntot)]
    singles = np.arange(ifirst - 1, N)
    cms = np.arange(N, ntot)
    if complete:
        near_s = [list(range(len(active)))] * nsingle
        near_p = [list(range(len(active)))] * npairs
    else:
        near_s = tree.query_ball_point(x[singles], rnb) if nsingle else []
        near_p = tree.query_ball_point(x[cms], rpert) if npairs else []
    for s, i in enumerate(singles):
        nb = active[np.asarray(near_s[s], dtype=np.int64)]
        rows[i] = sorted(int(j) + 1 for j in nb if j != i)
    for p, icm in enumerate(cms):
        nb = active[np.asarray(near_p[p], dtype=np.int64)]
        rows[2 * p] = sorted(int(j) + 1 for j in nb if j != icm)
    if lmax is None:
        lmax = max(len(r) for r in rows) + 8
    lst = np.zeros((ntot, lmax), np.int32)
    for i, r in enumerate(rows):
        lst[i, 0] = len(r)
        lst[i, 1:1 + len(r)] = r
    return dict(n=N, ntot=ntot, ifirst=ifirst, npairs=npairs, lmax=lmax,
                body=body, x=x, xdot=v, list=lst)
