"""ctypes front end of orc_energy / orc_phicor (energy.c), the CPU
restatement of the potential-energy sum of the reference's energy check
(src/Main/energy.F:37-79 with src/Main/phicor.f:1-112, ZKIN energy.F:141-146).
TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py's cpu_baseline leg as
the checker of gpunb_energy_dev.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import oracle


class Par(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("ntot", ctypes.c_int),
                ("ifirst", ctypes.c_int), ("lmax", ctypes.c_int),
                ("cmsep2_rmin2", ctypes.c_double)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _lib():
    lib = oracle.load()
    if not hasattr(lib, "_energy_bound"):
        lib.orc_energy.restype = ctypes.c_int
        lib.orc_energy.argtypes = [ctypes.POINTER(Par), _dp, _dp, _dp, _ip,
                                   _dp, _dp]
        lib.orc_phicor.restype = ctypes.c_int
        lib.orc_phicor.argtypes = [ctypes.POINTER(Par), _dp, _dp, _ip,
                                   ctypes.c_int, _dp, _dp]
        lib._energy_bound = True
    return lib


def energy(n, ntot, lmax, cmsep2_rmin2, body, x, xdot, lst, phidbl):
    """Returns (phidbl corrected, POT, ZKIN); phidbl on entry holds the
    potential of bodies IFIRST..NTOT (IFIRST = 2 (ntot - n) + 1)."""
    par = Par(n, ntot, 2 * (ntot - n) + 1, lmax, cmsep2_rmin2)
    body = np.ascontiguousarray(body, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    xdot = np.ascontiguousarray(xdot, dtype=np.float64)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    phi = np.array(phidbl, dtype=np.float64, order="C")
    out = np.zeros(2)
    rc = _lib().orc_energy(ctypes.byref(par), body.ctypes.data_as(_dp),
                           x.ctypes.data_as(_dp), xdot.ctypes.data_as(_dp),
                           lst.ctypes.data_as(_ip), phi.ctypes.data_as(_dp),
                           out.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("orc_energy: invalid LIST row")
    return phi, float(out[0]), float(out[1])
