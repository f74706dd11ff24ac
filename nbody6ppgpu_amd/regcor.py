"""Host-side front end of the device regular corrector (gpunb_regcor_dev).

It replaces the host loop of the reference's util_gpu.F:152-180,

    DO II = 1, NI
       CALL REGCOR_GPU(I, x_i(1,II), v_i(1,II), GPU_A(1,II), GPU_JERK(1,II),
                       LISTGP(1,II))
    END DO

(REGCOR_GPU: src/Main/regcor_gpu.F:1-805) with one call for the whole block
on the device.  ``DeviceState`` keeps the body arrays of common6.h that the
corrector reads and writes (X0, X0DOT, XN, XNDOT, FI, FIDOT, FR, FRDOT, D0,
D0R..D3R, T0R, STEPR, STEP, T0, TIMENW, RS, BODY, H, LIST) resident in HBM;
``correct`` runs one block from device arrays, ``correct_host`` uploads a
block from numpy, runs it and synchronises (the form the tests use).

Every call goes through libgpunb.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import gpunb
from .gpunb import (REGCOR_COUNTERS, REGCOR_SCAL, REGCOR_VEC, RegcorPar,
                    RegcorState)
from .hiprt import DeviceArray


def make_par(**kw) -> RegcorPar:
    """A gpunb_regcor_par from keyword values (rdens as a 3-sequence)."""
    p = RegcorPar()
    for k, v in kw.items():
        if k == "rdens":
            p.rdens[:] = [float(c) for c in v]
        else:
            setattr(p, k, v)
    return p


class DeviceState:
    """The corrector's body state in device memory.

    state: dict of numpy arrays indexed by body I - 1: REGCOR_VEC entries
    (ntot, 3) float64, REGCOR_SCAL entries (ntot,) float64, ``list``
    (ntot, lmax) int32 and optionally ``h`` (npairs,) float64."""

    def __init__(self, state: dict):
        self.arr = {}
        for k in REGCOR_VEC:
            self.arr[k] = DeviceArray.from_numpy(
                np.ascontiguousarray(state[k], dtype=np.float64))
        for k in REGCOR_SCAL:
            self.arr[k] = DeviceArray.from_numpy(
                np.ascontiguousarray(state[k], dtype=np.float64))
        self.arr["list"] = DeviceArray.from_numpy(
            np.ascontiguousarray(state["list"], dtype=np.int32))
        h = state.get("h")
        self.arr["h"] = (DeviceArray.from_numpy(
            np.ascontiguousarray(h, dtype=np.float64)) if h is not None
            else None)
        self.c = RegcorState()
        for k in REGCOR_VEC + REGCOR_SCAL + ("list",):
            setattr(self.c, k, self.arr[k].ptr)
        self.c.h = self.arr["h"].ptr if self.arr["h"] is not None else None

    def numpy(self) -> dict:
        out = {k: a.numpy() for k, a in self.arr.items() if a is not None}
        return out

    def copy_from(self, other: "DeviceState") -> None:
        """Device-to-device copy of another state of the same shapes."""
        for k, a in self.arr.items():
            if a is not None:
                a.copy_from(other.arr[k])


def compact_to_rows(nnb, off, lst, lmax: int, first: int = 1) -> np.ndarray:
    """gpunb_regf_dev's compact lists (0-based j, body included) as LIST /
    NLIST rows (util_gpu.F:103-110): count, then the 1-based members
    j + first without the body itself.  Shape (ni, lmax) int32."""
    nnb = np.asarray(nnb, dtype=np.int64)
    off = np.asarray(off, dtype=np.int64)
    ni = len(nnb)
    if np.any(nnb < 1) or np.any(off < 0):
        raise ValueError("every list must hold the body itself, no overflow")
    row = np.repeat(np.arange(ni), nnb)
    k = np.arange(int(nnb.sum())) - np.repeat(np.cumsum(nnb) - nnb, nnb)
    mem = np.asarray(lst)[np.repeat(off, nnb) + k] + first
    keep = mem != row + first
    if int((~keep).sum()) != ni:
        raise ValueError("each list must contain its body exactly once")
    kself = np.empty(ni, dtype=np.int64)
    kself[row[~keep]] = k[~keep]
    col = 1 + k - (k > kself[row])
    if int(nnb.max(initial=1)) + 1 > lmax:
        raise ValueError("lmax too small for these lists")
    rows = np.zeros((ni, lmax), dtype=np.int32)
    rows[:, 0] = nnb - 1
    rows[row[keep], col[keep]] = mem[keep]
    return rows


def correct(par: RegcorPar, dev: DeviceState, ni: int, d_idx, d_xi, d_vi,
            d_freg, d_fdr, d_nnb, d_off, d_list, list_len: int,
            d_cnt) -> None:
    """Enqueue gpunb_regcor_dev on the library stream (device pointers)."""
    lib = gpunb.load()
    rc = lib.gpunb_regcor_dev(ctypes.byref(par), ctypes.byref(dev.c), int(ni),
                              d_idx, d_xi, d_vi, d_freg, d_fdr, d_nnb, d_off,
                              d_list, int(list_len), d_cnt)
    if rc != 0:
        gpunb._check()


def correct_host(par: RegcorPar, dev: DeviceState, idx, xi, vi, freg, fdr,
                 nnb, off, lst):
    """One block from host arrays: idx (1-based I), xi / vi / freg / fdr
    (ni, 3), the compact lists of gpunb_regf_dev (nnb, off, lst; 0-based j,
    body included).  Returns (freg, fdr, counters) after synchronising."""
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    ni = len(idx)
    d = {"idx": DeviceArray.from_numpy(idx),
         "xi": DeviceArray.from_numpy(np.ascontiguousarray(xi, np.float64)),
         "vi": DeviceArray.from_numpy(np.ascontiguousarray(vi, np.float64)),
         "freg": DeviceArray.from_numpy(np.ascontiguousarray(freg, np.float64)),
         "fdr": DeviceArray.from_numpy(np.ascontiguousarray(fdr, np.float64)),
         "nnb": DeviceArray.from_numpy(np.ascontiguousarray(nnb, np.int32)),
         "off": DeviceArray.from_numpy(np.ascontiguousarray(off, np.int32)),
         "lst": DeviceArray.from_numpy(
             np.ascontiguousarray(lst, np.int32) if len(lst) else
             np.zeros(1, np.int32)),
         "cnt": DeviceArray.from_numpy(np.zeros(4, np.int32))}
    correct(par, dev, ni, d["idx"].ptr, d["xi"].ptr, d["vi"].ptr,
            d["freg"].ptr, d["fdr"].ptr, d["nnb"].ptr, d["off"].ptr,
            d["lst"].ptr, len(lst), d["cnt"].ptr)
    if gpunb.load().gpunb_sync() != 0:
        gpunb._check()
    cnt = d["cnt"].numpy()
    return (d["freg"].numpy(), d["fdr"].numpy(),
            dict(zip(REGCOR_COUNTERS, cnt.tolist())))
