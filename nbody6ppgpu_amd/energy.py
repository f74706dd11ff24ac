"""Host-side front end of the device energy sum (gpunb_energy_dev).

It replaces the GPU branch of the reference's ENERGY (src/Main/energy.F:37-79:
GPUPOT over IFIRST..NTOT, the c.m. potential copied to the KS components,
PHICOR's differential corrections, src/Main/phicor.f:1-112, and the sum
POT = 1/2 sum BODY*PHIDBL) plus ZKIN (energy.F:141-146).  Bodies follow the
reference's order: KS components 1..2 NPAIRS, singles up to N, c.m. bodies
N+1..NTOT; LIST rows as common6.h holds them (singles: neighbour lists; pair
IPAIR: its perturber list in row 2 IPAIR - 1).

Every call goes through libgpunb.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import gpunb
from .gpunb import EnergyPar
from .hiprt import DeviceArray


def make_par(n: int, ntot: int, lmax: int, cmsep2_rmin2: float) -> EnergyPar:
    return EnergyPar(int(n), int(ntot), 2 * (int(ntot) - int(n)) + 1,
                     int(lmax), float(cmsep2_rmin2))


def energy_dev(par: EnergyPar, d_body, d_x, d_xdot, d_list, run_pot: bool,
               d_phidbl, d_out) -> None:
    """Enqueue gpunb_energy_dev on the library stream (device pointers)."""
    rc = gpunb.load().gpunb_energy_dev(ctypes.byref(par), d_body, d_x, d_xdot,
                                       d_list, 1 if run_pot else 0, d_phidbl,
                                       d_out)
    if rc != 0:
        gpunb._check()


def potential_energy(n, ntot, lmax, cmsep2_rmin2, body, x, xdot, lst,
                     phidbl=None):
    """From host arrays: returns (PHIDBL, POT, ZKIN) after synchronising.
    phidbl None: the potentials of IFIRST..NTOT come from the device's
    gpupot kernels; else the given values are corrected."""
    par = make_par(n, ntot, lmax, cmsep2_rmin2)
    d = {"body": DeviceArray.from_numpy(np.ascontiguousarray(body, np.float64)),
         "x": DeviceArray.from_numpy(np.ascontiguousarray(x, np.float64)),
         "xdot": DeviceArray.from_numpy(np.ascontiguousarray(xdot, np.float64)),
         "list": DeviceArray.from_numpy(np.ascontiguousarray(lst, np.int32)),
         "phi": DeviceArray.from_numpy(
             np.zeros(ntot) if phidbl is None else
             np.ascontiguousarray(phidbl, np.float64)),
         "out": DeviceArray((2,), np.float64)}
    energy_dev(par, d["body"].ptr, d["x"].ptr, d["xdot"].ptr, d["list"].ptr,
               phidbl is None, d["phi"].ptr, d["out"].ptr)
    if gpunb.load().gpunb_sync() != 0:
        gpunb._check()
    out = d["out"].numpy()
    return d["phi"].numpy(), float(out[0]), float(out[1])
