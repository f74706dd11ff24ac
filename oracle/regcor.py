"""ctypes front end of orc_regcor (regcor.c), the CPU restatement of the
reference's regular corrector REGCOR_GPU (src/Main/regcor_gpu.F:1-805) for a
block of bodies.  TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py's
cpu_baseline leg as the checker of the device corrector.

State is a dict of numpy arrays indexed by body p = I - 1 (vectors (ntot, 3),
scalars (ntot,), list (ntot, lmax) int32 rows: count then 1-based members),
plus x / xdot: predicted positions and velocities of every body at TIME.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import oracle

VEC = ("x0", "x0dot", "xn", "xndot", "fi", "fidot", "fr", "frdot", "d0",
       "d0r", "d1r", "d2r", "d3r")
SCAL = ("t0r", "stepr", "step", "t0", "timenw", "rs", "body")
COUNTERS = ("nbvoid", "nbsmin", "nbflux", "niconv")


class Par(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("rscale", ctypes.c_double),
                ("rdens", ctypes.c_double * 3), ("rc2", ctypes.c_double),
                ("rsfac", ctypes.c_double), ("znbmin", ctypes.c_double),
                ("smax", ctypes.c_double), ("dtmin", ctypes.c_double),
                ("etar", ctypes.c_double), ("smin", ctypes.c_double),
                ("bodym", ctypes.c_double), ("tidal4", ctypes.c_double),
                ("nnbopt", ctypes.c_int), ("nnbmax", ctypes.c_int),
                ("n", ctypes.c_int), ("ntot", ctypes.c_int),
                ("ifirst", ctypes.c_int), ("lmax", ctypes.c_int),
                ("nb_flag", ctypes.c_int), ("m_flag", ctypes.c_int),
                ("kz13", ctypes.c_int), ("kz14", ctypes.c_int),
                ("kz37", ctypes.c_int), ("kz38", ctypes.c_int)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _State(ctypes.Structure):
    _fields_ = ([(k, _dp) for k in VEC + SCAL + ("h",)] +
                [("list", _ip), ("x", _dp), ("xdot", _dp)])


def make_par(**kw) -> Par:
    p = Par()
    for k, v in kw.items():
        if k == "rdens":
            p.rdens[:] = list(v)
        else:
            setattr(p, k, v)
    return p


def regcor(par: Par, state: dict, idx, xi, vi, freg, fdr, nlist):
    """Correct bodies idx (1-based) in place in `state`; returns (freg, fdr,
    nlist, counters) with the block arrays as modified."""
    lib = oracle.load()
    if not hasattr(lib, "_regcor_bound"):
        lib.orc_regcor.restype = ctypes.c_int
        lib.orc_regcor.argtypes = [ctypes.POINTER(Par), ctypes.POINTER(_State),
                                   ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _ip,
                                   _ip]
        lib._regcor_bound = True
    keep = []

    def dp(a):
        keep.append(a)
        return a.ctypes.data_as(_dp)

    for k in VEC + SCAL:
        assert state[k].dtype == np.float64 and state[k].flags.c_contiguous, k
    assert state["list"].dtype == np.int32 and state["list"].flags.c_contiguous
    st = _State()
    for k in VEC + SCAL:
        setattr(st, k, state[k].ctypes.data_as(_dp))
    h = state.get("h")
    if h is None:
        h = np.zeros(1)
    st.h = dp(np.ascontiguousarray(h, dtype=np.float64))
    st.list = state["list"].ctypes.data_as(_ip)
    st.x = dp(np.ascontiguousarray(state["x"], dtype=np.float64))
    st.xdot = dp(np.ascontiguousarray(state["xdot"], dtype=np.float64))
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    vi = np.ascontiguousarray(vi, dtype=np.float64)
    freg = np.array(freg, dtype=np.float64, order="C")
    fdr = np.array(fdr, dtype=np.float64, order="C")
    nlist = np.array(nlist, dtype=np.int32, order="C")
    cnt = np.zeros(4, dtype=np.int32)
    rc = lib.orc_regcor(ctypes.byref(par), ctypes.byref(st), len(idx),
                        idx.ctypes.data_as(_ip), xi.ctypes.data_as(_dp),
                        vi.ctypes.data_as(_dp), freg.ctypes.data_as(_dp),
                        fdr.ctypes.data_as(_dp), nlist.ctypes.data_as(_ip),
                        cnt.ctypes.data_as(_ip))
    if rc != 0:
        raise ValueError("orc_regcor: invalid argument")
    return freg, fdr, nlist, dict(zip(COUNTERS, cnt.tolist()))
