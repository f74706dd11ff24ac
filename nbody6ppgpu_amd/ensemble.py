"""Ensemble of independent star clusters on one GPU: workload, timing of one
regular block step of every cluster, and the autotuner of the ensemble
launch.

Workload: ``ncl`` Plummer clusters with N drawn log-uniform in [nmin, nmax],
equal-mass and Kroupa-IMF clusters alternating; every body of every cluster
is active (one full regular block step per cluster).

Three ways to advance them are timed:

* sequential -- one launch pair per cluster (what a per-cluster caller does;
  ni of the other clusters set to 0);
* batched    -- one launch pair for all clusters (gpunb_ens_regf_dev);
* graph      -- the batched launch captured once and replayed as a hipGraph.

The autotuner picks the pass-1 group size and wave budget for the whole
launch, then the j tiles per slice for every size bucket [2^b, 2^(b+1)),
keeping a change only when it lowers the measured step time.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import gpunb, ics
from .hiprt import DeviceArray, device_synchronize


def make_ensemble(ncl=64, nmin=1024, nmax=65536, seed=11, k=48):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), ncl)).astype(int)
    out = []
    for c, n in enumerate(sizes):
        m = ics.kroupa(int(n), seed=seed + 3 * c) if c % 2 else None
        cl = ics.plummer(int(n), seed=seed + 3 * c + 1, m=m)
        rs, dtr = ics.neighbour_radii(cl, k=k)
        out.append((cl, rs * rs, dtr))
    return out


@dataclass
class EnsembleBuffers:
    x: DeviceArray
    v: DeviceArray
    h2: DeviceArray
    dt: DeviceArray
    acc: DeviceArray
    jrk: DeviceArray
    pot: DeviceArray
    nnb: DeviceArray
    nnid: DeviceArray
    off: DeviceArray
    lst: DeviceArray
    total: DeviceArray
    cap: int
    ni: np.ndarray


def upload(clusters, nnbmax):
    cat = lambda f: np.concatenate([f(c) for c in clusters])
    tot = sum(c[0].n for c in clusters)
    cap = tot * nnbmax
    return EnsembleBuffers(
        DeviceArray.from_numpy(cat(lambda c: c[0].x)),
        DeviceArray.from_numpy(cat(lambda c: c[0].v)),
        DeviceArray.from_numpy(cat(lambda c: c[1])),
        DeviceArray.from_numpy(cat(lambda c: c[2])),
        DeviceArray((tot, 3), np.float64), DeviceArray((tot, 3), np.float64),
        DeviceArray((tot,), np.float64), DeviceArray((tot,), np.int32),
        DeviceArray((tot,), np.int32), DeviceArray((tot,), np.int32),
        DeviceArray((cap,), np.int32), DeviceArray((1,), np.int32), cap,
        np.array([c[0].n for c in clusters], dtype=np.int32))


def call(b: EnsembleBuffers, ni, nnbmax):
    niv = np.ascontiguousarray(ni, dtype=np.int32)
    rc = gpunb.load().gpunb_ens_regf_dev(
        len(niv), niv.ctypes.data_as(gpunb._ip), b.h2.ptr, b.dt.ptr, b.x.ptr,
        b.v.ptr, b.acc.ptr, b.jrk.ptr, b.pot.ptr, nnbmax, 0, b.nnb.ptr,
        b.nnid.ptr, b.lst.ptr, b.off.ptr, b.total.ptr, b.cap)
    if rc != 0:
        gpunb._check()


def time_step(b, nnbmax, mode, reps=5):
    """Seconds per block step of all clusters in the given mode."""
    ncl = len(b.ni)
    zeros = np.zeros(ncl, dtype=np.int32)

    def step():
        if mode == "sequential":
            for c in range(ncl):
                ni = zeros.copy()
                ni[c] = b.ni[c]
                call(b, ni, nnbmax)
        else:
            call(b, b.ni, nnbmax)

    gpunb.gpunb_ens_graph(mode == "graph")
    step()
    gpunb.gpunb_sync()
    device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    gpunb.gpunb_sync()
    device_synchronize()
    dt = (time.perf_counter() - t0) / reps
    gpunb.gpunb_ens_graph(False)
    return dt


def autotune(b, nnbmax, reps=3, log=None):
    """Returns (group, rounds, bucket_tiles, best_seconds)."""
    best = (None, None, [], 1e30)
    gpunb.gpunb_ens_set_bucket_tiles([])
    # Small clusters need many more slices per resident wave than one large
    # cluster to even out the launch tail (measured on 64 clusters of
    # 1024..65536: 6.8e11 pairs/s at 4 rounds, 1.05e12 at 64).
    for group in (2, 4):
        for rounds in (4.0, 16.0, 64.0, 128.0):
            gpunb.gpunb_set_tuning(group, rounds)
            t = time_step(b, nnbmax, "graph", reps)
            if log:
                log(f"autotune group {group} rounds {rounds}: {t * 1e3:.3f} ms")
            if t < best[3]:
                best = (group, rounds, [], t)
    group, rounds = best[0], best[1]
    gpunb.gpunb_set_tuning(group, rounds)
    buckets = sorted({int(np.log2(n)) for n in b.ni})
    tiles = [0] * 24
    t_best = best[3]
    for bk in buckets:
        for cand in (2, 8, 32, 128):
            trial = list(tiles)
            trial[bk] = cand
            gpunb.gpunb_ens_set_bucket_tiles(trial)
            t = time_step(b, nnbmax, "graph", reps)
            if log:
                log(f"autotune bucket 2^{bk} tiles {cand}: {t * 1e3:.3f} ms")
            if t < t_best:
                t_best, tiles = t, trial
    gpunb.gpunb_ens_set_bucket_tiles(tiles)
    return group, rounds, tiles, t_best
