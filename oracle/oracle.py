"""ctypes front end of liboracle.so (oracle.c), the FP64 restatement of the
reference's regular-force contract.  TEST INFRASTRUCTURE ONLY: used by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.

Function map (reference tree paths):
  regf     src/Main/gpunb.gpu.cu:149-241,752-794; src/Main/reg.avx.cpp:239-328
  pot      src/Main/gpupot.gpu.cu:32-57; src/Main/pot.avx.cpp:87-137
  predict  src/Main/xbpredall.f:18-26
  nearest  src/Main/nbint.F:29-47
  irr      src/Main/irr.avx.cpp:406-438 (with the xbpredall.f predictor)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with gcc (the oracle Makefile)."""
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", _HERE, "all"])
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.orc_regf.restype = None
        lib.orc_regf.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                                 _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                 ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int,
                                 _ip]
        lib.orc_pot.restype = None
        lib.orc_pot.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                _dp, _dp]
        lib.orc_predict.restype = None
        lib.orc_predict.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp,
                                    _dp, _dp, _dp, _dp, _dp]
        lib.orc_nearest.restype = ctypes.c_int
        lib.orc_nearest.argtypes = [_dp, ctypes.c_int, _ip, _dp]
        lib.orc_pair_is_nb.restype = ctypes.c_int
        lib.orc_pair_is_nb.argtypes = [_dp, _dp, ctypes.c_double,
                                       ctypes.c_double, _dp, _dp,
                                       ctypes.c_double, ctypes.c_int, _dp]
        lib.orc_irr.restype = None
        lib.orc_irr.argtypes = [ctypes.c_int, _ip, ctypes.c_double, _dp, _dp,
                                _dp, _dp, _dp, _dp, _ip, ctypes.c_int, _dp,
                                _dp, _ip]
        lib.orc_regf_ld.restype = None
        lib.orc_regf_ld.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                                    _dp, _dp, ctypes.c_void_p, _dp, _dp, _dp]
        _lib = lib
    return _lib


def _d(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def _p(a):
    return a.ctypes.data_as(_dp) if a.dtype == np.float64 else a.ctypes.data_as(_ip)


def regf(m, xj, vj, h2, dtr, xi, vi, lmax, nnbmax, m_flag=0):
    """Returns acc, jrk, pot, list (ni, lmax) int32, nnid (ni,) int32."""
    m = _d(m)
    nj = m.shape[0]
    xj = _d(xj, (nj, 3))
    vj = _d(vj, (nj, 3))
    h2 = _d(h2)
    ni = h2.shape[0]
    dtr = _d(dtr, (ni,))
    xi = _d(xi, (ni, 3))
    vi = _d(vi, (ni, 3))
    acc = np.zeros((ni, 3))
    jrk = np.zeros((ni, 3))
    pot = np.zeros(ni)
    lst = np.zeros((ni, lmax), dtype=np.int32)
    nnid = np.zeros(ni, dtype=np.int32)
    load().orc_regf(nj, _p(m), _p(xj), _p(vj), ni, _p(h2), _p(dtr), _p(xi),
                    _p(vi), _p(acc), _p(jrk), _p(pot), lmax, nnbmax, _p(lst),
                    int(m_flag), _p(nnid))
    return acc, jrk, pot, lst, nnid


def pot(istart, ni, m, x):
    m = _d(m)
    n = m.shape[0]
    x = _d(x, (n, 3))
    out = np.zeros(max(ni, 0))
    load().orc_pot(istart, ni, n, _p(m), _p(x), _p(out))
    return out


def predict(t, t0, x0, v0, f, fd):
    t0 = _d(t0)
    n = t0.shape[0]
    arrs = [_d(a, (n, 3)) for a in (x0, v0, f, fd)]
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    load().orc_predict(n, float(t), _p(t0), *[_p(a) for a in arrs], _p(x),
                       _p(v))
    return x, v


def irr(t, idx, t0, x0, v0, f, fd, m, lists):
    """Irregular force at time t of the particles idx (0-based) over their
    lists (lists[i] = sequence of 0-based neighbour indices of i).
    Returns (acc (ni, 3), jrk (ni, 3), nnbid (ni,), 1-based, 0 if empty)."""
    t0 = _d(t0)
    n = t0.shape[0]
    arrs = [_d(a, (n, 3)) for a in (x0, v0, f, fd)]
    stride = 1 + max([len(l) for l in lists] + [0])
    L = np.zeros((n, stride), dtype=np.int32)
    for i, l in enumerate(lists):
        L[i, 0] = len(l)
        L[i, 1:1 + len(l)] = l
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    ni = idx.shape[0]
    acc = np.zeros((ni, 3))
    jrk = np.zeros((ni, 3))
    nnbid = np.zeros(ni, dtype=np.int32)
    load().orc_irr(ni, _p(idx), float(t), _p(t0), *[_p(a) for a in arrs],
                   _p(_d(m)), _p(L), stride, _p(acc), _p(jrk), _p(nnbid))
    return acc, jrk, nnbid


def nearest(xi, idx, xj):
    xi = _d(xi, (3,))
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    xj = _d(xj)
    return load().orc_nearest(_p(xi), idx.shape[0], _p(idx), _p(xj))


def pair_is_nb(xi, vi, h2, dtr, xj, vj, mj, m_flag=0):
    r2 = ctypes.c_double(0.0)
    r = load().orc_pair_is_nb(_p(_d(xi, (3,))), _p(_d(vi, (3,))), float(h2),
                              float(dtr), _p(_d(xj, (3,))), _p(_d(vj, (3,))),
                              float(mj), int(m_flag), ctypes.byref(r2))
    return bool(r), r2.value


def regf_longdouble(m, xj, vj, xi, vi, mask):
    """Far-field sums in long double for a given exclusion mask (ni, nj)."""
    m = _d(m)
    nj = m.shape[0]
    xj = _d(xj, (nj, 3))
    vj = _d(vj, (nj, 3))
    xi = _d(xi)
    ni = xi.shape[0]
    xi = xi.reshape(ni, 3)
    vi = _d(vi, (ni, 3))
    mask = np.ascontiguousarray(mask, dtype=np.uint8).reshape(ni, nj)
    acc = np.zeros((ni, 3))
    jrk = np.zeros((ni, 3))
    pot = np.zeros(ni)
    load().orc_regf_ld(nj, _p(m), _p(xj), _p(vj), ni, _p(xi), _p(vi),
                       mask.ctypes.data_as(ctypes.c_void_p), _p(acc), _p(jrk),
                       _p(pot))
    return acc, jrk, pot


def list_rows(lst):
    """Split an (ni, lmax) reference-layout list into (counts, [arrays])."""
    counts = lst[:, 0].copy()
    rows = [lst[i, 1:1 + c] if c >= 0 else None for i, c in enumerate(counts)]
    return counts, rows
