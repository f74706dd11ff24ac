"""Systems with KS pairs in NBODY6++'s body order for the energy-sum tests:
components 1..2 NPAIRS (pair p: 2p-1, 2p), singles 2 NPAIRS+1..N, c.m.
bodies N+1..N+NPAIRS = NTOT.  LIST rows: singles hold their neighbours
(singles and c.m. bodies, ascending, so c.m. bodies come last); pair p holds
its perturbers in row 2p-1; every other row is empty.
"""
from __future__ import annotations

import numpy as np

from nbody6ppgpu_amd import ics


def build(**kw):
    return ics.ks_system(**kw)


def gpupot_phidbl(s):
    """PHIDBL(IFIRST..NTOT) as GPUPOT returns it: sum over the other bodies
    of IFIRST..NTOT of m_j / r_ij (c.m. bodies as point masses), FP64."""
    ifirst, ntot = s["ifirst"], s["ntot"]
    x, m = s["x"], s["body"]
    phi = np.zeros(ntot)
    idx = np.arange(ifirst - 1, ntot)
    for i in idx:
        d = x[idx] - x[i]
        r = np.sqrt(np.sum(d * d, axis=1))
        ok = r > 0
        phi[i] = np.sum(m[idx][ok] / r[ok])
    return phi


def resolved_potential(s):
    """Potential energy of singles and KS components as point masses with
    each pair's own bond left out (it belongs to the binding energy)."""
    npairs, N = s["npairs"], s["n"]
    x, m = s["x"][:N], s["body"][:N]
    pot = 0.0
    for i in range(N):
        d = x[i + 1:] - x[i]
        r = np.sqrt(np.sum(d * d, axis=1))
        w = m[i + 1:] / r
        if i < 2 * npairs and i % 2 == 0:
            w[0] = 0.0                         # i and i+1 form a pair
        pot += m[i] * np.sum(w)
    return pot
