"""ctypes front end of the reference's own CPU libraries, compiled from the
reference sources by oracle/Makefile into oracle/_ref/ (reg.avx.cpp,
reg.sse.cpp, pot.avx.cpp, pot.sse.cpp, irr.avx.cpp).  TEST INFRASTRUCTURE ONLY: used to
pin the oracle (tests/) and as the timed CPU baseline (bench.py).

Each library exports the same Fortran ABI as the engine (gpunb_open_ ...);
they are loaded RTLD_LOCAL so their symbols stay apart.  Caller-side padding
mirrors what the reference code assumes of its Fortran callers: reg.*.cpp
processes i in groups of 4 and writes whole groups (reg.avx.cpp:204-314), and
pot.avx.cpp processes i in groups of 8 at absolute offsets (pot.avx.cpp:87-137).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_libs = {}


def path(name: str) -> str:
    return os.path.join(REF_DIR, f"libref_{name}.so")


def available(name: str = "reg_avx") -> bool:
    return os.path.exists(path(name))


def build(ref_root: str = "/root/reference") -> bool:
    """Compile the reference CPU libraries when the reference tree exists."""
    if not os.path.isdir(ref_root):
        return False
    subprocess.check_call(["make", "-s", "-C", _HERE, "ref", f"REF={ref_root}"])
    return True


def load(name: str):
    if name not in _libs:
        p = path(name)
        if not os.path.exists(p):
            raise FileNotFoundError(p)
        lib = ctypes.CDLL(p, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        if name.startswith("irr"):
            lib.irr_simd_open_.argtypes = [_ip, _ip, _ip]
            lib.irr_simd_close_.argtypes = [_ip]
            lib.irr_simd_set_jp_.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp]
            lib.irr_simd_set_list_.argtypes = [_ip, _ip]
            lib.irr_simd_firr_vec_.argtypes = [_dp, _ip, _ip, _dp, _dp, _ip]
            for f in ("irr_simd_open_", "irr_simd_close_", "irr_simd_set_jp_",
                      "irr_simd_set_list_", "irr_simd_firr_vec_"):
                getattr(lib, f).restype = None
        elif name.startswith("reg"):
            lib.gpunb_open_.argtypes = [_ip, _ip]
            lib.gpunb_close_.argtypes = []
            lib.gpunb_send_.argtypes = [_ip, _dp, _dp, _dp]
            lib.gpunb_regf_.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                        _ip, _ip, _ip, _ip]
            for f in ("gpunb_open_", "gpunb_close_", "gpunb_send_",
                      "gpunb_regf_"):
                getattr(lib, f).restype = None
        else:
            lib.gpupot_.argtypes = [_ip, _ip, _ip, _ip, _dp, _dp, _dp]
            lib.gpupot_.restype = None
        _libs[name] = lib
    return _libs[name]


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _d(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


class RegLib:
    """One opened reference regular-force library (AVX or SSE)."""

    def __init__(self, name: str = "reg_avx", nbmax: int = 0):
        self.lib = load(name)
        self.nbmax = nbmax
        self.lib.gpunb_open_(_i(nbmax), _i(0))

    def close(self):
        self.lib.gpunb_close_()

    def send(self, m, x, v):
        m = _d(m)
        n = m.shape[0]
        x = _d(x, (n, 3))
        v = _d(v, (n, 3))
        self.lib.gpunb_send_(_i(n), m.ctypes.data_as(_dp),
                             x.ctypes.data_as(_dp), v.ctypes.data_as(_dp))

    def regf(self, h2, dtr, xi, vi, lmax, nnbmax, m_flag=0):
        h2 = _d(h2)
        ni = h2.shape[0]
        pad = (ni + 3) // 4 * 4 + 4
        H = np.zeros(pad)
        H[:ni] = h2
        D = np.zeros(pad)
        D[:ni] = _d(dtr, (ni,))
        X = np.zeros((pad, 3))
        X[:ni] = _d(xi, (ni, 3))
        V = np.zeros((pad, 3))
        V[:ni] = _d(vi, (ni, 3))
        acc = np.zeros((pad, 3))
        jrk = np.zeros((pad, 3))
        pot = np.zeros(pad)
        lst = np.zeros((pad, lmax), dtype=np.int32)
        self.lib.gpunb_regf_(_i(ni), H.ctypes.data_as(_dp),
                             D.ctypes.data_as(_dp), X.ctypes.data_as(_dp),
                             V.ctypes.data_as(_dp), acc.ctypes.data_as(_dp),
                             jrk.ctypes.data_as(_dp), pot.ctypes.data_as(_dp),
                             _i(lmax), _i(nnbmax), lst.ctypes.data_as(_ip),
                             _i(m_flag))
        return acc[:ni], jrk[:ni], pot[:ni], lst[:ni]


class IrrLib:
    """The reference's irregular-force library (irr.avx.cpp): predictor
    state and lists per particle (1-based addresses, as its Fortran callers
    pass them), FP32 pair arithmetic on double-single positions."""

    def __init__(self, name: str = "irr_avx", nmax: int = 0, lmax: int = 601):
        self.lib = load(name)
        self.lmax = lmax
        self.lib.irr_simd_open_(_i(nmax), _i(lmax), _i(0))

    def close(self):
        self.lib.irr_simd_close_(_i(0))

    def set_jp(self, t0, x0, v0, f, fd, m):
        """All particles 1..n: x0, v0, f = a/2, fd = jerk/6, mass, t0."""
        t0 = _d(t0)
        n = t0.shape[0]
        x0, v0, f, fd = (_d(a, (n, 3)) for a in (x0, v0, f, fd))
        m = _d(m)
        for i in range(n):
            self.lib.irr_simd_set_jp_(
                _i(i + 1), x0[i].ctypes.data_as(_dp), v0[i].ctypes.data_as(_dp),
                f[i].ctypes.data_as(_dp), fd[i].ctypes.data_as(_dp),
                m[i:i + 1].ctypes.data_as(_dp), t0[i:i + 1].ctypes.data_as(_dp))

    def set_list(self, i, members):
        """List of particle i (0-based) = members (0-based indices); the row
        is padded as a Fortran LIST(LMAX) row (the library reads whole
        groups of 8)."""
        row = np.zeros(max(self.lmax, len(members) + 9), dtype=np.int32)
        row[0] = len(members)
        row[1:1 + len(members)] = np.asarray(members) + 1
        self.lib.irr_simd_set_list_(_i(i + 1), row.ctypes.data_as(_ip))

    def firr(self, t, idx):
        """(acc, jrk, nnbid) of the particles idx (0-based) at time t;
        nnbid is 1-based."""
        idx = np.ascontiguousarray(np.asarray(idx) + 1, dtype=np.int32)
        ni = idx.shape[0]
        acc = np.zeros((ni, 3))
        jrk = np.zeros((ni, 3))
        nn = np.zeros(ni, dtype=np.int32)
        tt = ctypes.c_double(float(t))
        self.lib.irr_simd_firr_vec_(ctypes.byref(tt), _i(ni),
                                    idx.ctypes.data_as(_ip),
                                    acc.ctypes.data_as(_dp),
                                    jrk.ctypes.data_as(_dp),
                                    nn.ctypes.data_as(_ip))
        return acc, jrk, nn


def pot(name: str, m, x, istart: int = 1, ni: int = None):
    """Reference gpupot_ (AVX/SSE); istart must be 1 (the only form the
    reference's own caller uses, energy.F:38)."""
    lib = load(name)
    m = _d(m)
    n = m.shape[0]
    if ni is None:
        ni = n
    X = np.zeros((n + 16, 3))
    X[:n] = _d(x, (n, 3))
    M = np.zeros(n + 16)
    M[:n] = m
    out = np.zeros(n + 16)
    lib.gpupot_(_i(0), _i(istart), _i(ni), _i(n), M.ctypes.data_as(_dp),
                X.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out[istart - 1:istart - 1 + ni]
