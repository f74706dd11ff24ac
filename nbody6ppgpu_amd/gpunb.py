"""Python mirror of the regular-force library interface of NBODY6++GPU.

The functions keep the reference's names and argument meaning
(src/Main/gpunb.gpu.cu:821-848, src/Main/gpupot.gpu.cu:116-127): the Fortran
host calls ``gpunb_open(nbmax)``, ``gpunb_send(m, x, v)`` once per regular
block, ``gpunb_regf(...)`` per chunk of active particles and ``gpupot(...)``
per energy check.  Every call goes through the C-ABI of ``libgpunb.so`` (the
HIP engine); there is no CPU fallback, and loading fails loudly when the
library has not been built.

Arrays are numpy FP64 (positions/velocities shaped ``(n, 3)``, C order, which
is the memory layout of Fortran ``X(3,N)``); lists come back as an
``(ni, lmax)`` int32 array with the reference layout ``list[i, 0] = nnb`` (or
``-nnb`` on overflow) followed by ascending 0-based j indices.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgpunb.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)

_lib: Optional[ctypes.CDLL] = None


class GpunbError(RuntimeError):
    """An error reported by the library (precondition or device failure)."""


def library_path() -> str:
    return LIB_PATH


def load(path: Optional[str] = None) -> ctypes.CDLL:
    """Load libgpunb.so once; errors are returned, not aborted on."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"libgpunb.so not found at {p}: build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(p)
    _declare(lib)
    off = ctypes.c_int(0)
    lib.gpunb_set_abort_(ctypes.byref(off))
    _lib = lib
    return lib


class RegcorPar(ctypes.Structure):
    """gpunb_regcor_par: the common6.h scalars REGCOR_GPU reads."""
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


# gpunb_regcor_state fields, in order: per-body vectors, scalars, h, list.
REGCOR_VEC = ("x0", "x0dot", "xn", "xndot", "fi", "fidot", "fr", "frdot",
              "d0", "d0r", "d1r", "d2r", "d3r")
REGCOR_SCAL = ("t0r", "stepr", "step", "t0", "timenw", "rs", "body")
REGCOR_COUNTERS = ("nbvoid", "nbsmin", "nbflux", "niconv")


class RegcorState(ctypes.Structure):
    """gpunb_regcor_state: device pointers of the body state."""
    _fields_ = ([(k, ctypes.c_void_p) for k in REGCOR_VEC + REGCOR_SCAL] +
                [("h", ctypes.c_void_p), ("list", ctypes.c_void_p)])


class EnergyPar(ctypes.Structure):
    """gpunb_energy_par: body counts and the c.m. resolution radius."""
    _fields_ = [("n", ctypes.c_int), ("ntot", ctypes.c_int),
                ("ifirst", ctypes.c_int), ("lmax", ctypes.c_int),
                ("cmsep2_rmin2", ctypes.c_double)]


def _declare(lib: ctypes.CDLL) -> None:
    v = None
    sig = {
        "gpunb_devinit_": (v, [_ip]),
        "gpunb_open_": (v, [_ip, _ip]),
        "gpunb_close_": (v, []),
        "gpunb_send_": (v, [_ip, _dp, _dp, _dp]),
        "gpunb_regf_": (v, [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip,
                            _ip, _ip]),
        "gpunb_regfnn_": (v, [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                              _ip, _ip, _ip, _ip]),
        "gpunb_profile_": (v, [_ip]),
        "gpupot_": (v, [_ip, _ip, _ip, _ip, _dp, _dp, _dp]),
        "gpunb_jpset_": (v, [_ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp]),
        "gpunb_predict_": (v, [_dp]),
        "gpunb_jpget_": (v, [_ip, _dp, _dp]),
        "gpunb_set_abort_": (v, [_ip]),
        "gpunb_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
        "gpunb_regf_dev": (ctypes.c_int,
                           [ctypes.c_int] + [ctypes.c_void_p] * 7 +
                           [ctypes.c_int, ctypes.c_int] +
                           [ctypes.c_void_p] * 5 + [ctypes.c_longlong]),
        "gpunb_regf_dev_noalloc": (ctypes.c_int,
                                   [ctypes.c_int] + [ctypes.c_void_p] * 7 +
                                   [ctypes.c_int, ctypes.c_int] +
                                   [ctypes.c_void_p] * 5 +
                                   [ctypes.c_longlong]),
        "gpunb_regcor_dev": (ctypes.c_int,
                             [ctypes.POINTER(RegcorPar),
                              ctypes.POINTER(RegcorState), ctypes.c_int] +
                             [ctypes.c_void_p] * 8 +
                             [ctypes.c_longlong, ctypes.c_void_p]),
        "gpunb_energy_dev": (ctypes.c_int,
                             [ctypes.POINTER(EnergyPar)] +
                             [ctypes.c_void_p] * 4 + [ctypes.c_int] +
                             [ctypes.c_void_p] * 2),
        "gpunb_sync": (ctypes.c_int, []),
        "gpunb_stream": (ctypes.c_void_p, []),
        "gpunb_version": (ctypes.c_char_p, []),
        "gpunb_set_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_double]),
        "gpunb_set_precision": (ctypes.c_int, [ctypes.c_int]),
        "gpunb_ens_open": (ctypes.c_int, [ctypes.c_int, _ip]),
        "gpunb_ens_close": (ctypes.c_int, []),
        "gpunb_ens_send": (ctypes.c_int, [ctypes.c_int, _ip, _dp, _dp, _dp]),
        "gpunb_ens_regf_dev": (ctypes.c_int,
                               [ctypes.c_int, _ip] + [ctypes.c_void_p] * 7 +
                               [ctypes.c_int, ctypes.c_int] +
                               [ctypes.c_void_p] * 5 + [ctypes.c_longlong]),
        "gpunb_ens_graph": (ctypes.c_int, [ctypes.c_int]),
        "gpunb_ens_persistent": (ctypes.c_int, [ctypes.c_int]),
        "gpunb_ens_set_bucket_tiles": (ctypes.c_int, [ctypes.c_int, _ip]),
        "gpunb_ens_pairs": (ctypes.c_longlong, []),
        "gpunb_mlp_nparams": (ctypes.c_int, []),
        "gpunb_mlp_init": (ctypes.c_int, [ctypes.c_uint]),
        "gpunb_mlp_set_norm": (ctypes.c_int, [_fp, _fp, ctypes.c_float,
                                              ctypes.c_float]),
        "gpunb_mlp_get_params": (ctypes.c_int, [_fp]),
        "gpunb_mlp_set_params": (ctypes.c_int, [_fp]),
        "gpunb_mlp_predict": (ctypes.c_int, [ctypes.c_int, _fp, _fp]),
        "gpunb_mlp_grad": (ctypes.c_int,
                           [ctypes.c_int, _fp, _fp, _fp,
                            ctypes.POINTER(ctypes.c_double)]),
        "gpunb_mlp_train": (ctypes.c_int,
                            [ctypes.c_int, _fp, _fp, ctypes.c_int,
                             ctypes.c_int, ctypes.c_float, ctypes.c_uint,
                             ctypes.POINTER(ctypes.c_double)]),
        "irr_simd_open_": (v, [_ip, _ip, _ip]),
        "irr_simd_close_": (v, [_ip]),
        "irr_simd_profile_": (v, [_ip]),
        "irr_simd_set_jp_": (v, [_ip, _dp, _dp, _dp, _dp, _dp, _dp]),
        "irr_simd_set_list_": (v, [_ip, _ip]),
        "irr_simd_firr_vec_": (v, [_dp, _ip, _ip, _dp, _dp, _ip]),
        "gpunb_irr_set_jp": (ctypes.c_int, [ctypes.c_int, _ip, _dp, _dp, _dp,
                                            _dp, _dp, _dp]),
        "gpunb_irr_set_lists": (ctypes.c_int, [ctypes.c_int, _ip, _ip, _ip]),
        "gpunb_irr_kernel_ms": (ctypes.c_double, []),
        "gpunb_timing": (ctypes.c_int, [ctypes.c_int]),
        "gpunb_timing_read": (ctypes.c_int,
                              [ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_longlong),
                               ctypes.POINTER(ctypes.c_longlong)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def _check() -> None:
    lib = load()
    buf = ctypes.create_string_buffer(4096)
    n = lib.gpunb_last_error(buf, 4096)
    if n:
        raise GpunbError(buf.value.decode(errors="replace"))


def _f64(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp) if a.dtype == np.float64 else a.ctypes.data_as(_ip)


def _i(v: int):
    return ctypes.byref(ctypes.c_int(int(v)))


def gpunb_devinit(irank: int = 0) -> None:
    load().gpunb_devinit_(_i(irank))
    _check()


def gpunb_open(nbmax: int, irank: int = 0) -> None:
    load().gpunb_open_(_i(nbmax), _i(irank))
    _check()


def gpunb_close() -> None:
    load().gpunb_close_()
    _check()


def gpunb_send(m, x, v) -> None:
    m = _f64(m)
    n = m.shape[0]
    x = _f64(x, (n, 3))
    v = _f64(v, (n, 3))
    load().gpunb_send_(_i(n), _p(m), _p(x), _p(v))
    _check()


def gpunb_regf(h2, dtr, xi, vi, lmax: int, nnbmax: int, m_flag: int = 0,
               nearest: bool = False):
    """Regular force of the i-particles against the sent j set.

    Returns ``(acc, jrk, pot, list)`` and, with ``nearest=True``, also the
    nearest-neighbour index array (``gpunb_regfnn_``).
    """
    h2 = _f64(h2)
    ni = h2.shape[0]
    dtr = _f64(dtr, (ni,))
    xi = _f64(xi, (ni, 3))
    vi = _f64(vi, (ni, 3))
    acc = np.zeros((ni, 3))
    jrk = np.zeros((ni, 3))
    pot = np.zeros(ni)
    lst = np.zeros((ni, lmax), dtype=np.int32)
    lib = load()
    if nearest:
        nnid = np.full(ni, -2, dtype=np.int32)
        lib.gpunb_regfnn_(_i(ni), _p(h2), _p(dtr), _p(xi), _p(vi), _p(acc),
                          _p(jrk), _p(pot), _i(lmax), _i(nnbmax), _p(lst),
                          _i(m_flag), _p(nnid))
        _check()
        return acc, jrk, pot, lst, nnid
    lib.gpunb_regf_(_i(ni), _p(h2), _p(dtr), _p(xi), _p(vi), _p(acc),
                    _p(jrk), _p(pot), _i(lmax), _i(nnbmax), _p(lst),
                    _i(m_flag))
    _check()
    return acc, jrk, pot, lst


def gpunb_profile(irank: int = 0) -> None:
    load().gpunb_profile_(_i(irank))
    _check()


def gpupot(istart: int, ni: int, m, x, irank: int = 0) -> np.ndarray:
    """phi of particles istart-1 .. istart+ni-2 (1-based istart) over all n."""
    m = _f64(m)
    n = m.shape[0]
    x = _f64(x, (n, 3))
    pot = np.zeros(max(ni, 0))
    load().gpupot_(_i(irank), _i(istart), _i(ni), _i(n), _p(m), _p(x),
                   _p(pot))
    _check()
    return pot


def gpunb_jpset(addr, x0, v0, f, fdot, mass, t0) -> None:
    addr = np.ascontiguousarray(addr, dtype=np.int32)
    n = addr.shape[0]
    arrs = [_f64(a, (n, 3)) for a in (x0, v0, f, fdot)]
    mass = _f64(mass, (n,))
    t0 = _f64(t0, (n,))
    load().gpunb_jpset_(_i(n), _p(addr), *[_p(a) for a in arrs], _p(mass),
                        _p(t0))
    _check()


def gpunb_predict(time: float) -> None:
    load().gpunb_predict_(ctypes.byref(ctypes.c_double(float(time))))
    _check()


def gpunb_jpget(n: int):
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    load().gpunb_jpget_(_i(n), _p(x), _p(v))
    _check()
    return x, v


def gpunb_set_tuning(group: int = 0, rounds: float = 0.0) -> None:
    if load().gpunb_set_tuning(int(group), float(rounds)) != 0:
        _check()


def gpunb_set_precision(prec: int = 0) -> None:
    if load().gpunb_set_precision(int(prec)) != 0:
        _check()


def gpunb_timing(enable: bool) -> None:
    if load().gpunb_timing(1 if enable else 0) != 0:
        _check()


def gpunb_timing_read():
    """(kernel ms summed, launches, pair interactions) since the last read."""
    ms = ctypes.c_double()
    n = ctypes.c_longlong()
    w = ctypes.c_longlong()
    if load().gpunb_timing_read(ctypes.byref(ms), ctypes.byref(n),
                                ctypes.byref(w)) != 0:
        _check()
    return ms.value, n.value, w.value


def gpunb_ens_open(nbmax) -> None:
    nb = np.ascontiguousarray(nbmax, dtype=np.int32)
    if load().gpunb_ens_open(nb.shape[0], _p(nb)) != 0:
        _check()


def gpunb_ens_close() -> None:
    if load().gpunb_ens_close() != 0:
        _check()


def gpunb_ens_send(clusters) -> None:
    """clusters: sequence of (m, x, v) per cluster."""
    nj = np.array([len(c[0]) for c in clusters], dtype=np.int32)
    m = _f64(np.concatenate([np.asarray(c[0], float) for c in clusters]))
    x = _f64(np.concatenate([np.asarray(c[1], float).reshape(-1, 3) for c in clusters]))
    v = _f64(np.concatenate([np.asarray(c[2], float).reshape(-1, 3) for c in clusters]))
    if load().gpunb_ens_send(nj.shape[0], _p(nj), _p(m), _p(x), _p(v)) != 0:
        _check()


def gpunb_ens_graph(enable: bool) -> None:
    if load().gpunb_ens_graph(1 if enable else 0) != 0:
        _check()


def gpunb_ens_persistent(enable: bool) -> None:
    if load().gpunb_ens_persistent(1 if enable else 0) != 0:
        _check()


def gpunb_ens_set_bucket_tiles(tiles) -> None:
    t = np.ascontiguousarray(tiles, dtype=np.int32)
    if load().gpunb_ens_set_bucket_tiles(t.shape[0], _p(t)) != 0:
        _check()


def gpunb_sync() -> None:
    if load().gpunb_sync() != 0:
        _check()


def gpunb_stream() -> int:
    return int(load().gpunb_stream() or 0)


def version() -> str:
    return load().gpunb_version().decode()


# --- irregular force (irr_simd_* drop-in; 1-based addresses) ---------------

def irr_open(nmax: int, lmax: int, rank: int = 0) -> None:
    load().irr_simd_open_(_i(nmax), _i(lmax), _i(rank))
    _check()


def irr_close(rank: int = 0) -> None:
    load().irr_simd_close_(_i(rank))
    _check()


def irr_profile(rank: int = 0) -> None:
    load().irr_simd_profile_(_i(rank))
    _check()


def irr_set_jp(addr, x0, v0, f, fd, m, t0) -> None:
    """Predictor state of particles addr (1-based): x0, v0, f = a/2,
    fd = jerk/6 ([n][3]), mass and t0 ([n])."""
    a = np.ascontiguousarray(addr, dtype=np.int32)
    n = a.shape[0]
    arrs = [_f64(q, (n, 3)) for q in (x0, v0, f, fd)]
    if load().gpunb_irr_set_jp(n, _p(a), *[_p(q) for q in arrs],
                               _p(_f64(m, (n,))), _p(_f64(t0, (n,)))) != 0:
        _check()


def irr_set_lists(addr, lists) -> None:
    """lists[q]: 1-based members of particle addr[q]."""
    a = np.ascontiguousarray(addr, dtype=np.int32)
    nnb = np.array([len(l) for l in lists], dtype=np.int32)
    flat = np.ascontiguousarray(
        np.concatenate([np.asarray(l, dtype=np.int32) for l in lists])
        if len(lists) else np.zeros(0, np.int32), dtype=np.int32)
    if flat.size == 0:
        flat = np.zeros(1, np.int32)
    if load().gpunb_irr_set_lists(a.shape[0], _p(a), _p(nnb), _p(flat)) != 0:
        _check()


def irr_kernel_ms() -> float:
    """Duration (ms) of the force kernel of the last irr_firr with ni > 0."""
    return float(load().gpunb_irr_kernel_ms())


def irr_firr(t: float, addr):
    """(acc [ni][3], jrk [ni][3], nnbid [ni] 1-based) at time t."""
    a = np.ascontiguousarray(addr, dtype=np.int32)
    ni = a.shape[0]
    acc = np.zeros((max(ni, 1), 3))
    jrk = np.zeros((max(ni, 1), 3))
    nn = np.zeros(max(ni, 1), dtype=np.int32)
    tt = ctypes.c_double(float(t))
    load().irr_simd_firr_vec_(ctypes.byref(tt), _i(ni), _p(a if ni else nn),
                              _p(acc), _p(jrk), _p(nn))
    _check()
    return acc[:ni], jrk[:ni], nn[:ni]
