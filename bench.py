"""Regular-force benchmark of the MI355X engine (one JSON line on stdout).

Workload: King W0=6 cluster, N = 131072 bodies, Kroupa IMF (0.08-150 Msun,
total mass 1), Henon units; neighbour radii with 48 distance neighbours
(~55 with the predicted-distance criterion), STEPR as FPOLY0 sets it.
One step = one regular block over ALL bodies: the device predictor
(gpunb_predict_) followed by the regular force of all N i-particles against
all N j-particles (gpunb_regf_dev: inputs and outputs resident in HBM),
i.e. N^2 = 1.72e10 pair interactions per step.

value = pair interactions / s over the timed steps (wall clock between two
device synchronisations).  roofline: the pair kernel (k_regf_pass1), timed
with HIP events on the library stream around every launch in the timed
region; achieved = 60 flops per pair (the reference's convention,
gpunb.gpu.cu:811) x pairs per launch / mean launch time, against the MI355X
FP32 peak.  cpu_baseline: the reference's own AVX regular force
(src/Main/reg.avx.cpp, compiled from source under oracle/_ref) on the same
j set, i in chunks of 1024 as the Fortran host calls it (util_gpu.F:33-37),
timed on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 (vector and matrix) peak, AMD spec
FLOPS_PER_PAIR = 60            # reference convention (gpunb.gpu.cu:811)
N_BODIES = 131072


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(n=N_BODIES, seed=7):
    from nbody6ppgpu_amd import ics
    c = ics.king(n, 6.0, seed=seed, m=ics.kroupa(n, seed=seed + 1))
    rs, stepr = ics.neighbour_radii(c, k=48)
    return c, rs * rs, stepr


def load_traffic():
    """HBM bytes per pair-kernel launch from the committed rocprofv3 PMC
    summary (profiles/regf_pass1_king131072_traffic.json), if present."""
    p = os.path.join(ROOT, "profiles", "regf_pass1_king131072_traffic.json")
    if os.path.exists(p):
        try:
            with open(p) as f:
                return json.load(f).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            return None
    return None


def regcor_bench(c, h2, stepr, reps=5, cpu=True):
    """Regular corrector on the device (gpunb_regcor_dev) for every body of
    the King block, fed by gpunb_regf_dev lists at 0.9 RS (the old LIST) and
    RS (the new lists), DTR = STEPR, TIME = 1: one call per rep on the same
    input (the state is restored by a device copy outside the timed region),
    HIP events on the library stream around the call.  The CPU restatement
    of REGCOR_GPU (oracle/regcor.c, serial as util_gpu.F:152-180 calls it)
    runs on the same block as the baseline.  Expects the engine open with
    the bodies sent."""
    from nbody6ppgpu_amd import gpunb, regcor
    from nbody6ppgpu_amd.hiprt import DeviceArray, Event
    n = c.n
    lmax, nnbmax = 600, 550
    lib = gpunb.load()
    stream = gpunb.gpunb_stream()
    d = {k: DeviceArray.from_numpy(a) for k, a in
         (("x", c.x), ("v", c.v), ("dtr", stepr))}
    outs = {}
    for tag, fac in (("old", 0.9), ("new", 1.0)):
        dh2 = DeviceArray.from_numpy(h2 * fac * fac)
        o = {k: DeviceArray(sh, dt) for k, sh, dt in
             (("acc", (n, 3), np.float64), ("jrk", (n, 3), np.float64),
              ("pot", (n,), np.float64), ("nnb", (n,), np.int32),
              ("off", (n,), np.int32), ("lst", (n * nnbmax,), np.int32),
              ("tot", (1,), np.int32))}
        rc = lib.gpunb_regf_dev(n, dh2.ptr, d["dtr"].ptr, d["x"].ptr,
                                d["v"].ptr, o["acc"].ptr, o["jrk"].ptr,
                                o["pot"].ptr, nnbmax, 0, o["nnb"].ptr, 0,
                                o["lst"].ptr, o["off"].ptr, o["tot"].ptr,
                                n * nnbmax)
        if rc != 0 or lib.gpunb_sync() != 0:
            gpunb._check()
        outs[tag] = o
    old = outs["old"]
    LIST = regcor.compact_to_rows(old["nnb"].numpy(), old["off"].numpy(),
                                  old["lst"].numpy(), lmax)
    rng = np.random.default_rng(11)
    state = {k: rng.normal(0.0, 1.0, (n, 3)) for k in
             ("fi", "fidot", "fr", "frdot")}
    for k in ("xn", "xndot", "d0", "d0r", "d1r", "d2r", "d3r"):
        state[k] = np.zeros((n, 3))
    state.update(x0=c.x, x0dot=c.v, t0r=1.0 - stepr, stepr=stepr.copy(),
                 step=stepr / 2.0 ** rng.integers(0, 5, n), t0=np.ones(n),
                 rs=np.sqrt(h2) * 0.9, body=c.m, list=LIST)
    state["timenw"] = state["t0"] + state["step"]
    par_kw = dict(time=1.0, rscale=1.0, rdens=(0.0, 0.0, 0.0), rc2=0.05 ** 2,
                  rsfac=2.0, znbmin=10.0, smax=1.0, dtmin=1e-9, etar=0.02,
                  smin=0.0, bodym=float(c.m.mean()), tidal4=0.0, nnbopt=56,
                  nnbmax=nnbmax, n=n, ntot=n, ifirst=1, lmax=lmax, nb_flag=1,
                  m_flag=0)
    par = regcor.make_par(**par_kw)
    new = outs["new"]
    pristine = regcor.DeviceState(state)
    work = regcor.DeviceState(state)
    freg0 = DeviceArray((n, 3), np.float64).copy_from(new["acc"])
    fdr0 = DeviceArray((n, 3), np.float64).copy_from(new["jrk"])
    idx = DeviceArray.from_numpy(np.arange(1, n + 1, dtype=np.int32))
    cnt = DeviceArray((4,), np.int32)
    times = []
    for r in range(reps + 1):
        work.copy_from(pristine)
        new["acc"].copy_from(freg0)
        new["jrk"].copy_from(fdr0)
        cnt.zero()
        e0, e1 = Event(), Event()
        e0.record(stream)
        regcor.correct(par, work, n, idx.ptr, d["x"].ptr, d["v"].ptr,
                       new["acc"].ptr, new["jrk"].ptr, new["nnb"].ptr,
                       new["off"].ptr, new["lst"].ptr, n * nnbmax, cnt.ptr)
        e1.record(stream)
        if lib.gpunb_sync() != 0:
            gpunb._check()
        got = cnt.numpy().tolist()
        if r == 0:
            first = got
        elif got != first:
            raise RuntimeError(f"regcor counters differ between calls on the "
                               f"same input: {got} vs {first}")
        if r:                      # the first call warms up
            times.append(e0.elapsed_ms(e1))
    counts = dict(zip(gpunb.REGCOR_COUNTERS, first))
    nnb_new = new["nnb"].numpy().astype(np.int64)
    nnb0 = LIST[:, 0].astype(np.int64)
    flux = counts["nbflux"]
    # Algorithmic bytes of one call, what any implementation moves: per body
    # the block inputs (idx, x, v, FREG, FDR, count, offset: 108 B), the
    # state read (X0, X0DOT, FI, FIDOT, FR, FRDOT, T0R, STEPR, STEP, TIMENW,
    # RS: 184 B) and written (XN, XNDOT, FI, FIDOT, FR, FRDOT, D0, D0R..D3R,
    # T0R, RS, STEPR, FREG, FDR, STEP, TIMENW: 368 B); per list entry 4 B
    # (new list read, old LIST read, new LIST written); per gained or lost
    # neighbour its 56-B record and its STEP.  Working storage of this
    # implementation (the STEP / TIMENW staging for the commit launch, the
    # JJLIST scratch) is traffic, not algorithmic bytes.
    nbytes = (n * (108 + 184 + 368) +
              4 * int(nnb_new.sum() + (nnb0 + 1).sum() + nnb_new.sum()) +
              64 * flux)
    ms = float(np.median(times))
    out = {"workload": f"King W0=6 N={n}: every body corrected after a "
                       "regular step whose lists grew from 0.9 RS to RS",
           "bodies": n, "counters": counts, "kernel_ms_median": ms,
           "kernel_ms": times, "bodies_per_s": n / (ms * 1e-3),
           "algorithmic_bytes": nbytes,
           "achieved_GBps": nbytes / (ms * 1e-3) / 1e9,
           "peak_GBps": 8000.0}
    if cpu:
        from oracle import regcor as orc
        freg_h, fdr_h = freg0.numpy(), fdr0.numpy()
        rows = regcor.compact_to_rows(nnb_new, new["off"].numpy(),
                                      new["lst"].numpy(), lmax)
        st = {k: (np.array(v, dtype=np.int32) if k == "list" else
                  np.array(v, dtype=np.float64)) for k, v in state.items()}
        st["x"], st["xdot"] = c.x, c.v
        t0 = time.perf_counter()
        orc.regcor(orc.make_par(**par_kw), st, np.arange(1, n + 1), c.x, c.v,
                   freg_h, fdr_h, rows)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / dt, "unit": "bodies/s",
                               "cores": 1, "kind": "port",
                               "sample": f"oracle/regcor.c, all {n} bodies "
                                         f"once, {dt:.3f} s"}
    return out


def irr_bench(c, h2, stepr, reps=5, cpu=True, min_s=5.0):
    """Irregular force through the irr_simd_* drop-in for every body of the
    King block: lists from gpunb_regf_dev at RS, predictor records x0, v0,
    a/2, jerk/6 of that regular force, all bodies active at t = 1/64 past
    t0.  The first call uploads the records and lists; each timed call is
    one irr_simd_firr_vec_ of all bodies, timed around the C entry point
    (index upload, kernel, result copy).  The reference's irr.avx.cpp
    (oracle/_ref, OpenMP over the host cores) on the same state is the
    baseline, and its forces and nearest neighbours are compared.  Expects
    the engine open with the bodies sent."""
    from nbody6ppgpu_amd import gpunb, regcor
    from nbody6ppgpu_amd.hiprt import DeviceArray
    n = c.n
    lmax, nnbmax = 600, 550
    lib = gpunb.load()
    d = {k: DeviceArray.from_numpy(a) for k, a in
         (("x", c.x), ("v", c.v), ("dtr", stepr), ("h2", h2))}
    o = {k: DeviceArray(sh, dt) for k, sh, dt in
         (("acc", (n, 3), np.float64), ("jrk", (n, 3), np.float64),
          ("pot", (n,), np.float64), ("nnb", (n,), np.int32),
          ("off", (n,), np.int32), ("lst", (n * nnbmax,), np.int32),
          ("tot", (1,), np.int32))}
    rc = lib.gpunb_regf_dev(n, d["h2"].ptr, d["dtr"].ptr, d["x"].ptr,
                            d["v"].ptr, o["acc"].ptr, o["jrk"].ptr,
                            o["pot"].ptr, nnbmax, 0, o["nnb"].ptr, 0,
                            o["lst"].ptr, o["off"].ptr, o["tot"].ptr,
                            n * nnbmax)
    if rc != 0 or lib.gpunb_sync() != 0:
        gpunb._check()
    rows = regcor.compact_to_rows(o["nnb"].numpy(), o["off"].numpy(),
                                  o["lst"].numpy(), lmax)
    acc0, jrk0 = o["acc"].numpy(), o["jrk"].numpy()
    lists = [rows[i, 1:1 + rows[i, 0]] for i in range(n)]
    pairs = int(rows[:, 0].sum())
    addr = np.arange(1, n + 1, dtype=np.int32)
    t0 = np.zeros(n)
    t = 1.0 / 64
    f, fd = 0.5 * acc0, jrk0 / 6.0
    gpunb.irr_open(n + 10, lmax)
    gpunb.irr_set_jp(addr, c.x, c.v, f, fd, c.m, t0)
    gpunb.irr_set_lists(addr, lists)
    acc, jrk, nn = gpunb.irr_firr(t, addr)
    # Timed calls: the C entry point with output arrays allocated once, as
    # the Fortran host holds them.
    import ctypes
    oa, oj = np.zeros((n, 3)), np.zeros((n, 3))
    on = np.zeros(n, dtype=np.int32)
    tt, nn_c = ctypes.c_double(t), ctypes.c_int(n)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    args = (ctypes.byref(tt), ctypes.byref(nn_c), addr.ctypes.data_as(ip),
            oa.ctypes.data_as(dp), oj.ctypes.data_as(dp), on.ctypes.data_as(ip))
    times, kms = [], []
    for _ in range(reps):
        t1 = time.perf_counter()
        lib.irr_simd_firr_vec_(*args)
        times.append(time.perf_counter() - t1)
        gpunb._check()
        kms.append(gpunb.irr_kernel_ms())
    if not (np.array_equal(oa, acc) and np.array_equal(oj, jrk) and
            np.array_equal(on, nn)):
        raise RuntimeError("irregular force: repeated calls differ")
    gpunb.irr_close()
    ms = float(np.median(times)) * 1e3
    kms_med = float(np.median(kms))
    # Algorithmic bytes of the call's kernels.  With more than three pairs
    # per slot (slots = n + 10 here) the library predicts every slot once
    # (128 B read, 64 B written per slot) and the force kernel reads per
    # body its index, predicted record (64 B), list count and outputs
    # (6 doubles + the nearest index), per pair the member's index and
    # predicted record (4 + 64 B); otherwise 128-byte records per member.
    nslot = n + 10
    if pairs > 3 * nslot:
        abytes = nslot * (128 + 64) + n * (4 + 64 + 4 + 52) + pairs * (4 + 64)
    else:
        abytes = n * (4 + 128 + 4 + 52) + pairs * (4 + 128)
    leg = {"workload": "King W0=6 N=131072: every body active, its "
                       "gpunb_regf_dev list at RS, records predicted to "
                       "t0 + 1/64",
           "bodies": n, "pairs_per_call": pairs,
           "ms_per_call": ms, "ms_per_call_all": [x * 1e3 for x in times],
           "pairs_per_s": pairs / (ms * 1e-3),
           "bodies_per_s": n / (ms * 1e-3),
           "note": "ms_per_call is timed around the C entry point "
                   "irr_simd_firr_vec_ with output arrays allocated once: "
                   "index upload, kernels, result copy and the copy into "
                   "the caller's arrays",
           "kernel_ms": kms_med, "kernel_ms_all": kms,
           "kernel_pairs_per_s": pairs / (kms_med * 1e-3),
           "algorithmic_bytes": abytes,
           "achieved_GBps": abytes / (kms_med * 1e-3) / 1e9,
           "peak_GBps": 8000.0}
    if cpu:
        from oracle import ref
        if ref.available("irr_avx"):
            threads = int(os.environ.get("OMP_NUM_THREADS",
                                         os.cpu_count() or 1))
            R = ref.IrrLib("irr_avx", n + 10, lmax + 1)
            R.set_jp(t0, c.x, c.v, f, fd, c.m)
            for i in range(n):
                R.set_list(i, lists[i] - 1)
            ra, rj, rn = R.firr(t, np.arange(n))
            calls = 0
            t1 = time.perf_counter()
            while True:
                R.firr(t, np.arange(n))
                calls += 1
                if time.perf_counter() - t1 >= min_s:
                    break
            dt = (time.perf_counter() - t1) / calls
            R.close()
            rel = np.linalg.norm(ra - acc, axis=1) / np.linalg.norm(acc, axis=1)
            relj = np.linalg.norm(rj - jrk, axis=1) / np.linalg.norm(jrk, axis=1)
            leg["vs_reference"] = {
                "acc_rel_median": float(np.median(rel)),
                "jrk_rel_median": float(np.median(relj)),
                "nearest_equal_frac": float(np.mean(rn == nn))}
            if not np.median(rel) < 1e-5:
                raise RuntimeError(f"irregular force: median relative "
                                   f"difference {np.median(rel)!r} from the "
                                   f"reference library")
            leg["cpu_baseline"] = {
                "value": pairs / dt, "unit": "pair interactions/s",
                "cores": threads, "kind": "reference",
                "sample": f"reference irr.avx.cpp irr_simd_firr_vec_ of all "
                          f"{n} bodies, {calls} calls in {dt * calls:.1f} s, "
                          f"OMP_NUM_THREADS={threads}"}
    return leg


def predictor_pot_bench(c, cpu=True, reps=50):
    """The device predictor (gpunb_predict_, k_predict) and the O(N^2)
    potential (gpupot_).

    Predictor: synthetic predictor states of N = 131072 and 1048576 bodies
    sent by gpunb_jpset_, HIP events on the library stream around `reps`
    back-to-back gpunb_predict_ calls; 176 algorithmic bytes per body (13
    doubles and the mass read, a 64-byte j record written).  The predicted
    records of the 131072-body set must equal the oracle's xbpredall.f
    restatement bit for bit; the oracle (serial C) is the CPU baseline.

    Potential: gpupot_ of all bodies of the King block (host arrays in and
    out, as energy.F:38 calls it), wall time per call; the reference
    pot.avx.cpp (oracle/_ref, OpenMP over the host cores) on the same
    bodies as the baseline and as a check."""
    from nbody6ppgpu_amd import gpunb
    from nbody6ppgpu_amd.hiprt import Event
    leg = {"predict": []}
    rng = np.random.default_rng(31)
    for n in (131072, 1048576):
        x0, v0 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
        f, fd = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
        m, t0 = np.full(n, 1.0 / n), rng.uniform(-0.01, 0.0, n)
        t = 0.001
        gpunb.gpunb_open(n + 10)
        try:
            gpunb.gpunb_jpset(np.arange(1, n + 1), x0, v0, f, fd, m, t0)
            st = gpunb.gpunb_stream()
            for _ in range(3):
                gpunb.gpunb_predict(t)
            gpunb.gpunb_sync()
            e0, e1 = Event(), Event()
            e0.record(st)
            for _ in range(reps):
                gpunb.gpunb_predict(t)
            e1.record(st)
            gpunb.gpunb_sync()
            us = e0.elapsed_ms(e1) / reps * 1e3
            xd, vd = gpunb.gpunb_jpget(n) if n == 131072 else (None, None)
        finally:
            gpunb.gpunb_close()
        row = {"bodies": n, "us_per_predict": us,
               "achieved_GBps": 176 * n / (us * 1e-6) / 1e9, "peak_GBps": 8000.0}
        if xd is not None:
            from oracle import oracle
            xo, vo = oracle.predict(t, t0, x0, v0, f, fd)
            row["bit_exact_vs_oracle"] = bool(np.array_equal(xd, xo) and
                                              np.array_equal(vd, vo))
            if not row["bit_exact_vs_oracle"]:
                raise RuntimeError("device predictor differs from the oracle")
            if cpu:
                calls, t1 = 0, time.perf_counter()
                while time.perf_counter() - t1 < 2.0:
                    oracle.predict(t, t0, x0, v0, f, fd)
                    calls += 1
                dt = (time.perf_counter() - t1) / calls
                row["cpu_baseline"] = {
                    "value": n / dt, "unit": "bodies/s", "cores": 1,
                    "kind": "port",
                    "sample": f"oracle/oracle.c orc_predict of {n} bodies, "
                              f"{calls} calls, {dt * 1e3:.2f} ms each"}
        leg["predict"].append(row)
    n = c.n
    gpunb.gpunb_open(n + 10)
    try:
        phi = gpunb.gpupot(1, n, c.m, c.x)
        times = []
        for _ in range(3):
            t1 = time.perf_counter()
            gpunb.gpupot(1, n, c.m, c.x)
            times.append(time.perf_counter() - t1)
    finally:
        gpunb.gpunb_close()
    ms = float(np.median(times)) * 1e3
    pot = {"workload": f"gpupot_ of all {n} bodies of the King W0=6 block",
           "ms_per_call": ms, "pairs_per_s": float(n) * n / (ms * 1e-3)}
    if cpu:
        from oracle import ref
        if ref.available("pot_avx"):
            threads = int(os.environ.get("OMP_NUM_THREADS",
                                         os.cpu_count() or 1))
            ni = n
            t1 = time.perf_counter()
            pr = ref.pot("pot_avx", c.m, c.x, 1, ni)
            dt = time.perf_counter() - t1
            rel = np.abs(pr - phi[:ni]) / np.abs(phi[:ni])
            pot["vs_reference_rel_max"] = float(rel.max())
            if not rel.max() < 1e-5:
                raise RuntimeError(f"gpupot differs from pot.avx.cpp by "
                                   f"{rel.max()!r}")
            pot["cpu_baseline"] = {
                "value": float(ni) * n / dt, "unit": "pair interactions/s",
                "cores": threads, "kind": "reference",
                "sample": f"reference pot.avx.cpp gpupot_ of all {n} "
                          f"bodies, {dt:.2f} s, OMP_NUM_THREADS={threads}"}
    leg["gpupot"] = pot
    return leg


def energy_bench(reps=5, cpu=True):
    """Energy sum on the device (gpunb_energy_dev): a Plummer system of
    N = 131072 (126976 singles and 2048 KS pairs), neighbour lists within
    0.08 and pair perturber lists within 0.1, c.m. resolution radius 0.03.
    HIP events around each call, once with the device potential (the O(N^2)
    gpupot kernels) and once with the corrections and sums alone; the CPU
    restatement of ENERGY + PHICOR (serial, as the reference's loop runs)
    on the same system as the baseline of the corrections."""
    from nbody6ppgpu_amd import energy, gpunb, ics
    from nbody6ppgpu_amd.hiprt import DeviceArray, Event
    s = ics.ks_system(nsingle=126976, npairs=2048, seed=21, rnb=0.08,
                      rpert=0.1)
    cut = 0.03 ** 2
    par = energy.make_par(s["n"], s["ntot"], s["lmax"], cut)
    d = {k: DeviceArray.from_numpy(np.ascontiguousarray(s[k])) for k in
         ("body", "x", "xdot", "list")}
    phi = DeviceArray((s["ntot"],), np.float64).zero()
    out = DeviceArray((2,), np.float64)
    stream = gpunb.gpunb_stream()
    lib = gpunb.load()
    res = {}
    for mode, run_pot in (("with_potential", True), ("corrections_only", False)):
        times = []
        for r in range(reps + 1):
            e0, e1 = Event(), Event()
            e0.record(stream)
            energy.energy_dev(par, d["body"].ptr, d["x"].ptr, d["xdot"].ptr,
                              d["list"].ptr, run_pot, phi.ptr, out.ptr)
            e1.record(stream)
            if lib.gpunb_sync() != 0:
                gpunb._check()
            if r:
                times.append(e0.elapsed_ms(e1))
        res[mode] = float(np.median(times))
        if run_pot:
            pot_zkin = out.numpy().tolist()
    nb = s["ntot"] - s["ifirst"] + 1
    leg = {"workload": "Plummer N=131072: 126976 singles + 2048 KS pairs, "
                       "lists within 0.08 (pairs: 0.1), resolution 0.03",
           "bodies": nb, "pot_zkin": pot_zkin,
           "ms_with_potential": res["with_potential"],
           "ms_corrections_only": res["corrections_only"],
           "bodies_per_s_corrections": nb / (res["corrections_only"] * 1e-3)}
    if cpu:
        from oracle import energy as orc
        # The same potentials through gpupot_ (same kernels), corrected and
        # summed by the serial restatement: timed, and its POT must agree.
        ifirst = s["ifirst"]
        phi_in = np.zeros(s["ntot"])
        phi_in[ifirst - 1:] = gpunb.gpupot(1, nb, s["body"][ifirst - 1:],
                                           s["x"][ifirst - 1:])
        t0 = time.perf_counter()
        _, pot, zkin = orc.energy(s["n"], s["ntot"], s["lmax"], cut, s["body"],
                                  s["x"], s["xdot"], s["list"], phi_in)
        dt = time.perf_counter() - t0
        if abs(pot - pot_zkin[0]) > 1e-12 * abs(pot):
            raise RuntimeError(f"energy sum: device POT {pot_zkin[0]!r} "
                               f"against {pot!r}")
        leg["cpu_baseline"] = {"value": nb / dt, "unit": "bodies/s",
                               "cores": 1, "kind": "port",
                               "sample": f"oracle/energy.c corrections and "
                                         f"sums of all {nb} bodies, "
                                         f"{dt * 1e3:.1f} ms"}
    return leg


def cpu_baseline(c, h2, stepr, min_s=10.0):
    """Reference AVX regular force (oracle/_ref) on the same j set: repeated
    sweeps over all i in 1024-i calls until at least min_s seconds."""
    from oracle import ref
    n = c.n
    lmax, nnbmax = 601, 550
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = min(threads, 32)      # reg.avx.cpp:7,103 asserts <= 32 threads
    if ref.available("reg_avx"):
        R = ref.RegLib("reg_avx", n + 10)
        R.send(c.m, c.x, c.v)
        done, k = 0, 0
        t0 = time.perf_counter()
        while True:
            sl = slice(k, min(k + 1024, n))
            R.regf(h2[sl], stepr[sl], c.x[sl], c.v[sl], lmax, nnbmax, 0)
            done += sl.stop - sl.start
            k = sl.stop % n
            if time.perf_counter() - t0 >= min_s:
                break
        dt = time.perf_counter() - t0
        R.close()
        return {"value": done * n / dt, "unit": "pair interactions/s",
                "cores": threads, "kind": "reference",
                "sample": f"reference reg.avx.cpp gpunb_regf_, {done / n:.1f} "
                          f"sweeps of all {n} i in 1024-i calls against all "
                          f"{n} j, OMP_NUM_THREADS={threads}, {dt:.1f} s"}
    from oracle import oracle
    ni = 256
    t0 = time.perf_counter()
    oracle.regf(c.m, c.x, c.v, h2[:ni], stepr[:ni], c.x[:ni], c.v[:ni], lmax,
                nnbmax, 0)
    dt = time.perf_counter() - t0
    return {"value": ni * n / dt, "unit": "pair interactions/s", "cores": 1,
            "kind": "port", "sample": f"FP64 oracle, {ni} i against {n} j"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline leg")
    ap.add_argument("--prec", type=int, default=0,
                    help="pair arithmetic: 32, 64 or 65 (0 = library default)")
    ap.add_argument("--no-regcor", action="store_true",
                    help="skip the device regular-corrector leg")
    ap.add_argument("--no-energy", action="store_true",
                    help="skip the device energy-sum leg")
    ap.add_argument("--no-irr", action="store_true",
                    help="skip the irregular-force leg")
    ap.add_argument("--no-predict-pot", action="store_true",
                    help="skip the predictor and gpupot legs")
    ap.add_argument("--no-hermite", action="store_true",
                    help="skip the Hermite/Ahmad-Cohen integration leg")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the ensemble (64 clusters) leg")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned neighbour-radius leg")
    ap.add_argument("--learned-n", type=int, default=262144,
                    help="bodies of the learned-radius leg (the full "
                         "workload is 1048576: about 6 minutes)")
    ap.add_argument("--learned-t", type=float, default=0.25)
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py runs on one GPU only: this build measures the "
                 "regular-force engine on a single MI355X (--gpus 1).")

    from nbody6ppgpu_amd import gpunb
    from nbody6ppgpu_amd.hiprt import DeviceArray, Event, device_synchronize

    log("bench: building workload (King W0=6, N=131072, Kroupa IMF)")
    c, h2, stepr = workload()
    n = c.n
    nnbmax = 550
    gpunb.gpunb_open(n + 10)
    gpunb.gpunb_jpset(np.arange(1, n + 1), c.x, c.v, np.zeros((n, 3)),
                      np.zeros((n, 3)), c.m, np.zeros(n))
    d = {k: DeviceArray.from_numpy(a) for k, a in
         (("x", c.x), ("v", c.v), ("h2", h2), ("dtr", stepr))}
    acc = DeviceArray((n, 3), np.float64)
    jrk = DeviceArray((n, 3), np.float64)
    pot = DeviceArray((n,), np.float64)
    nnb = DeviceArray((n,), np.int32)
    nnid = DeviceArray((n,), np.int32)
    off = DeviceArray((n,), np.int32)
    cap = n * nnbmax
    lst = DeviceArray((cap,), np.int32)
    tot = DeviceArray((1,), np.int32)
    lib = gpunb.load()
    stream = gpunb.gpunb_stream()
    if args.prec:
        gpunb.gpunb_set_precision(args.prec)

    def step():
        gpunb.gpunb_predict(0.0)
        rc = lib.gpunb_regf_dev(n, d["h2"].ptr, d["dtr"].ptr, d["x"].ptr,
                                d["v"].ptr, acc.ptr, jrk.ptr, pot.ptr, nnbmax,
                                0, nnb.ptr, nnid.ptr, lst.ptr, off.ptr,
                                tot.ptr, cap)
        if rc != 0:
            gpunb._check()

    log(f"bench: warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    gpunb.gpunb_sync()
    counts = nnb.numpy()
    assert np.all(counts >= 1), "every body must list itself"
    nnb_mean = float(counts.mean())

    log(f"bench: timing {args.steps} steps")
    gpunb.gpunb_timing(True)
    e0, e1 = Event(), Event()
    device_synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    gpunb.gpunb_sync()
    device_synchronize()
    t1 = time.perf_counter()
    k_ms, k_launches, k_pairs = gpunb.gpunb_timing_read()
    gpunb.gpunb_timing(False)
    ev_ms = e0.elapsed_ms(e1)
    wall_ms = (t1 - t0) * 1e3
    ms_per_step = wall_ms / args.steps
    pairs_per_step = float(n) * n
    value = pairs_per_step / (ms_per_step * 1e-3)

    avg_launch_s = k_ms * 1e-3 / max(k_launches, 1)
    pairs_per_launch = k_pairs / max(k_launches, 1)
    achieved = FLOPS_PER_PAIR * pairs_per_launch / avg_launch_s / 1e12
    log(f"bench: {ms_per_step:.3f} ms/step (events {ev_ms / args.steps:.3f}), "
        f"pair kernel {k_launches} launches, mean {avg_launch_s * 1e3:.3f} ms")

    # Drop-in ABI path with host arrays (PCIe-inclusive), i in chunks of
    # 1024 as the Fortran host calls it (util_gpu.F:33-37), into output
    # arrays allocated once as the host's LISTGP/FREG arrays are; reported
    # next to value, never as it.
    gpunb.gpunb_send(c.m, c.x, c.v)
    nchunk, ci = 32, 1024
    lmax_abi = nnbmax + 1
    X = np.ascontiguousarray(c.x)
    V = np.ascontiguousarray(c.v)
    H2 = np.ascontiguousarray(h2)
    DT = np.ascontiguousarray(stepr)
    acc_o = np.zeros((ci, 3))
    jrk_o = np.zeros((ci, 3))
    pot_o = np.zeros(ci)
    lst_o = np.zeros((ci, lmax_abi), dtype=np.int32)
    lib = gpunb.load()
    dptr = ctypes.POINTER(ctypes.c_double)

    def rows(a, first):
        """Pointer to row `first` of a C-contiguous FP64 array."""
        return ctypes.cast(a.ctypes.data + first * a.strides[0], dptr)

    ni_c, lmax_c, nnb_c, mf_c = (ctypes.c_int(v) for v in (ci, lmax_abi, nnbmax, 0))
    outs = [rows(acc_o, 0), rows(jrk_o, 0), rows(pot_o, 0)]
    lst_p = lst_o.ctypes.data_as(ctypes.POINTER(ctypes.c_int))

    def call(q):
        o = q * ci
        lib.gpunb_regf_(ctypes.byref(ni_c), rows(H2, o), rows(DT, o),
                        rows(X, o), rows(V, o), *outs, ctypes.byref(lmax_c),
                        ctypes.byref(nnb_c), lst_p, ctypes.byref(mf_c))

    call(0)
    gpunb._check()
    # Host-latency bound (one synchronising call per 1024 i): best of three
    # passes of 32 calls, so a busy host core does not set the figure.
    abi_rate = 0.0
    for _ in range(3):
        t2 = time.perf_counter()
        for q in range(nchunk):
            call(q)
        abi_rate = max(abi_rate, nchunk * ci * n / (time.perf_counter() - t2))
    gpunb._check()

    # Regular corrector on the device for every body of the block (while the
    # engine still holds the sent bodies as its j records).
    regcor_leg = None
    if not args.no_regcor:
        log("bench: regular corrector (gpunb_regcor_dev), all bodies")
        regcor_leg = regcor_bench(c, h2, stepr, cpu=not args.no_cpu)
    # Irregular force of every body through the irr_simd_* drop-in.
    irr_leg = None
    if not args.no_irr:
        log("bench: irregular force (irr_simd_firr_vec_), all bodies")
        irr_leg = irr_bench(c, h2, stepr, cpu=not args.no_cpu)
    gpunb.gpunb_close()

    # Device predictor and the O(N^2) potential.
    pp_leg = None
    if not args.no_predict_pot:
        log("bench: device predictor and gpupot_")
        pp_leg = predictor_pot_bench(c, cpu=not args.no_cpu)

    # Potential-energy sum of the energy check with KS pairs on the device.
    energy_leg = None
    if not args.no_energy:
        log("bench: energy sum (gpunb_energy_dev), N=131072 with KS pairs")
        energy_leg = energy_bench(cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (reference AVX path)")
        cpu = cpu_baseline(c, h2, stepr)

    # |dE/E| over one N-body time unit: the host Hermite / Ahmad-Cohen driver
    # (host/hermite_ac.cpp) integrates the same cluster through the engine's
    # ABI (device predictor + gpunb_regf_ + gpupot_) with ETAI 0.05 and
    # ETAR 0.1, once with the library's default kernel choice (the mixed
    # kernel for blocks of at least 8192 i, run in Hilbert order; the FP64
    # kernel for smaller, scattered blocks) and once with the FP64 kernel
    # for every block, side by side.
    hermite = []
    if not args.no_hermite:
        from nbody6ppgpu_amd import integrate
        for prec, label in ((0, "default: mixed FP32/FP64 for Hilbert-sorted "
                                "blocks >= 8192 i, FP64 below"),
                            (65, "FP64 (v_rsq_f64 + one Newton step)")):
            log(f"bench: Hermite/AC integration, ETAI 0.05 ETAR 0.1, {label}")
            rec = integrate.run(c.m, c.x, c.v, t_end=1.0, etai=0.05, etar=0.1,
                                prec=prec, timeout=900)
            last = rec[-1]
            hermite.append({
                "etai": 0.05, "etar": 0.1, "kernel": label, "prec": prec,
                "t_end": last["t"], "dE_E": last["dE_E"],
                "max_abs_dE_E": max(abs(r["dE_E"]) for r in rec),
                "dE_E_checks": [r["dE_E"] for r in rec[1:]],
                "regular_pairs": last["reg_pairs"],
                "regular_pairs_per_s": last["reg_pairs"] / max(last["t_reg"], 1e-9),
                "irregular_pairs": last["irr_pairs"],
                "interactions_per_s": (last["reg_pairs"] + last["irr_pairs"])
                                      / max(last["wall"], 1e-9),
                "irregular_steps": last["irr_steps"],
                "regular_steps": last["reg_steps"],
                "blocks": last["blocks"], "nnb_mean": last["nnb_mean"],
                "rs_rule": last["rs_rule"], "wall_s": last["wall"]})

    # Ensemble: 64 independent Plummer clusters, N log-uniform in
    # [1024, 65536], one regular block step of every body of every cluster,
    # sequential per-cluster launches vs one batched launch vs its hipGraph,
    # after the autotuner has chosen the launch shape.
    ensemble = None
    if not args.no_ensemble:
        from nbody6ppgpu_amd import ensemble as ens
        log("bench: ensemble workload (64 clusters)")
        cl = ens.make_ensemble()
        nnb_e = 300
        gpunb.gpunb_ens_open([c[0].n for c in cl])
        gpunb.gpunb_ens_send([(c[0].m, c[0].x, c[0].v) for c in cl])
        eb = ens.upload(cl, nnb_e)
        pairs = float(sum(c[0].n ** 2 for c in cl))
        # The library default for the ensemble: the mixed kernel on every
        # cluster's i in Hilbert order, pass 1 as a persistent launch.
        t_seq = ens.time_step(eb, nnb_e, "sequential", reps=3)
        t_bat = ens.time_step(eb, nnb_e, "batched", reps=5)
        t_gra = ens.time_step(eb, nnb_e, "graph", reps=5)
        gpunb.gpunb_ens_persistent(False)
        t_bat_plain = ens.time_step(eb, nnb_e, "batched", reps=5)
        gpunb.gpunb_ens_persistent(True)
        group, rounds, tiles, t_tuned = ens.autotune(eb, nnb_e, reps=3, log=log)
        t_final = ens.time_step(eb, nnb_e, "graph", reps=10)
        # The FP64 pair kernel (one Newton step), one launch per cluster at
        # the default decomposition: the second baseline.
        gpunb.gpunb_ens_set_bucket_tiles([])
        gpunb.gpunb_set_tuning(0, 0.0)
        gpunb.gpunb_set_precision(65)
        t_seq64 = ens.time_step(eb, nnb_e, "sequential", reps=3)
        gpunb.gpunb_set_precision(0)
        gpunb.gpunb_ens_close()
        gpunb.gpunb_set_tuning(0, 0.0)
        ensemble = {
            "workload": "64 Plummer clusters, N log-uniform in [1024, 65536], "
                        "equal-mass and Kroupa alternating, all bodies active",
            "clusters": len(cl), "bodies": int(sum(c[0].n for c in cl)),
            "pairs_per_step": pairs,
            "kernel": "mixed (FP32 far tiles, FP64 near tiles), every "
                      "cluster's i in Hilbert order",
            "sequential_pairs_per_s": pairs / t_seq,
            "batched_pairs_per_s": pairs / t_bat,
            "batched_one_workgroup_per_task_pairs_per_s": pairs / t_bat_plain,
            "graph_pairs_per_s": pairs / t_gra,
            "tuned_graph_pairs_per_s": pairs / t_final,
            "fp64_sequential_pairs_per_s": pairs / t_seq64,
            # Headline: the same mixed-precision kernel, tuned persistent
            # batched graph launch over one launch pair per cluster.
            "batched_over_sequential": t_seq / t_final,
            "mixed_batched_over_fp64_sequential": t_seq64 / t_final,
            "tuned": {"group": group, "rounds": rounds,
                      "bucket_tiles": {f"2^{b}": t for b, t in enumerate(tiles) if t}},
        }

    # Learned neighbour radius: mass-segregated Plummer with a Kroupa IMF;
    # 2M (at most) training pairs logged from a plain run, the network
    # trained on the GPU, then an independent realisation run with the
    # reference's Ahmad-Cohen rule, the plain volume rule and the learned
    # radius (Ahmad-Cohen fallback), each for learned_t N-body time units.
    learned = None
    if not args.no_learned:
        from nbody6ppgpu_amd import rsnet
        log(f"bench: learned neighbour radius, N={args.learned_n}")
        ex = rsnet.experiment(n=args.learned_n, t_log=args.learned_t,
                              t_run=args.learned_t, epochs=20, log=log)
        learned = {
            "workload": "mass-segregated Plummer, Kroupa IMF 0.08-150 Msun, "
                        f"N={args.learned_n}, {args.learned_t} N-body time "
                        "units per run; MLP 8-64-64-1 BF16/FP32 on MFMA",
            "fit": ex["fit"], "runs": ex["runs"]}

    traffic = load_traffic()
    out = {
        "metric": "regular-force pair interactions/s at N=128k, % FP32 peak; "
                  "dE/E per t_NB",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.prec >= 64 else "f32",
        "data": "synthetic",
        "config": {
            "workload": "King W0=6, N=131072, Kroupa IMF 0.08-150 Msun; one "
                        "regular block over all N bodies (device predictor + "
                        "regular force of N i vs N j), inputs resident in HBM",
            "n": n, "ni_per_step": n, "pairs_per_step": pairs_per_step,
            "nnb_mean": nnb_mean, "nnbmax": nnbmax,
            "precision": ("FP64 pair arithmetic (hardware v_rsq_f64, %d "
                          "Newton steps)" % (args.prec - 64))
                         if args.prec >= 64 else
                         ("FP32 pair arithmetic on tile-relative "
                          "differences in far tiles (j Hilbert-sorted in 64-j "
                          "tiles with FP64 centres, i Hilbert-sorted); FP64 "
                          "pair arithmetic for every tile within a "
                          "wavefront's near band (12 h2 w_j), the neighbour "
                          "predicate and the accumulation"),
        },
        "roofline": {
            "bound": "valu",
            "compute_unit": "FP32 VALU (pairwise 1/r^3 sums are not a dense "
                            "contraction); peak = MI355X FP32 vector peak",
            "kernel": "k_regf_pass1",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "flops_per_pair": FLOPS_PER_PAIR,
            "pairs_per_launch": pairs_per_launch,
            "mean_launch_ms": avg_launch_s * 1e3,
            "launches": k_launches,
        },
        "cpu_baseline": cpu,
        "abi_ni1024_pairs_per_s": abi_rate,
        "regcor": regcor_leg,
        "irr": irr_leg,
        "predictor_pot": pp_leg,
        "energy": energy_leg,
        "hermite_1tnb": hermite,
        "ensemble": ensemble,
        "learned_rs": learned,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
