"""Body states for the regular-corrector tests (CPU pins and GPU parity).

A state is what REGCOR_GPU (src/Main/regcor_gpu.F) reads: predicted X / XDOT
of every body at TIME, the old lists LIST and irregular force FI / FIDOT
summed over them, the regular force FR / FRDOT at T0R = TIME - DTR, and the
new regular force FREG / FDR with the new lists (in gpunb_regf_dev's compact
form: 0-based j, the body itself included).  The regular force is a cubic in
time per body, F(s) = F0 + F1 s + F2 s^2/2 + F3 s^3/6 with s = t - T0R, so
the corrector's derivatives have closed forms.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

from nbody6ppgpu_amd import ics


def pair_sums(x, v, m, i, members):
    """Newtonian force and its time derivative on body i (0-based) from the
    0-based members, FP64 numpy."""
    members = np.asarray(members, dtype=int)
    if members.size == 0:
        return np.zeros(3), np.zeros(3)
    d = x[members] - x[i]
    u = v[members] - v[i]
    r2 = np.sum(d * d, axis=1)
    r3 = r2 * np.sqrt(r2)
    rv = np.sum(d * u, axis=1)
    f = np.sum(m[members, None] * d / r3[:, None], axis=0)
    fd = np.sum(m[members, None] * (u - 3.0 * (rv / r2)[:, None] * d) /
                r3[:, None], axis=0)
    return f, fd


def neighbours(x, rs):
    """Sorted 0-based members within rs[i] of body i, the body excluded."""
    tree = cKDTree(x)
    out = []
    for i in range(len(x)):
        nb = [j for j in tree.query_ball_point(x[i], rs[i]) if j != i]
        out.append(sorted(nb))
    return out


def build(n=2048, seed=5, lmax=600, nnbmax=500, nnbopt=50, dtr=1.0 / 64,
          time=1.0, change=0.15, ncm=0, smin=0.0, hrange=(1.0, 50.0),
          nb_flag=1, m_flag=0, tidal4=0.0):
    """A state for n bodies (the last ncm of them c.m. bodies I > N) whose
    lists change between the old radius rs and the new radius rs * (1 +/-
    change).  Returns (par_kw, state, block) with block = dict(idx, xi, vi,
    freg, fdr, nnb, off, lst, lists_new)."""
    rng = np.random.default_rng(seed)
    c = ics.plummer(n, seed=seed)
    x, v, m = c.x.copy(), c.v.copy(), c.m.copy()
    r = np.linalg.norm(x, axis=1)
    # radii giving roughly nnbopt neighbours from the local density
    tree = cKDTree(x)
    dk, _ = tree.query(x, k=nnbopt + 1)
    rs_old = dk[:, -1] * rng.uniform(0.85, 1.15, n)
    rs_new = rs_old * (1.0 + change * rng.uniform(-1, 1, n))
    old = neighbours(x, rs_old)
    new = neighbours(x, rs_new)
    for i in range(n):           # keep within nnbmax
        old[i] = old[i][:nnbmax]
        new[i] = new[i][:nnbmax]
    ntot = n
    N = n - ncm
    state = {k: np.zeros((ntot, 3)) for k in
             ("x0", "x0dot", "xn", "xndot", "fi", "fidot", "fr", "frdot",
              "d0", "d0r", "d1r", "d2r", "d3r")}
    LIST = np.zeros((ntot, lmax), dtype=np.int32)
    for i in range(ntot):
        LIST[i, 0] = len(old[i])
        LIST[i, 1:1 + len(old[i])] = np.asarray(old[i], dtype=np.int32) + 1
        state["fi"][i], state["fidot"][i] = pair_sums(x, v, m, i, old[i])
    F0 = rng.normal(0, 1.0, (ntot, 3))
    F1 = rng.normal(0, 2.0, (ntot, 3))
    F2 = rng.normal(0, 4.0, (ntot, 3))
    F3 = rng.normal(0, 8.0, (ntot, 3))
    state["fr"][:] = F0
    state["frdot"][:] = F1
    s = dtr
    freg = F0 + F1 * s + F2 * s * s / 2 + F3 * s ** 3 / 6
    fdr = F1 + F2 * s + F3 * s * s / 2
    state["x0"][:] = x + rng.normal(0, 1e-6, (ntot, 3))
    state["x0dot"][:] = v + rng.normal(0, 1e-6, (ntot, 3))
    state["t0r"] = np.full(ntot, time - dtr)
    state["stepr"] = np.full(ntot, dtr)
    state["step"] = dtr / 2.0 ** rng.integers(0, 6, ntot)
    state["t0"] = np.full(ntot, time)
    state["timenw"] = state["t0"] + state["step"]
    state["rs"] = rs_old.copy()
    state["body"] = m.copy()
    state["list"] = LIST
    state["h"] = -rng.uniform(hrange[0], hrange[1], max(ncm, 1))
    state["x"] = x
    state["xdot"] = v
    par_kw = dict(time=time, rscale=1.0, rdens=(0.01, -0.02, 0.005),
                  rc2=0.1 ** 2, rsfac=2.0, znbmin=0.2 * nnbopt, smax=1.0,
                  dtmin=1e-9, etar=0.02, smin=smin, bodym=float(m.mean()),
                  tidal4=tidal4, nnbopt=nnbopt, nnbmax=nnbmax, n=N, ntot=ntot,
                  ifirst=1, lmax=lmax, nb_flag=nb_flag, m_flag=m_flag)
    # new lists in the compact regular-force form (0-based j, self included)
    idx = np.arange(1, ntot + 1, dtype=np.int32)
    nnb = np.zeros(ntot, np.int32)
    off = np.zeros(ntot, np.int32)
    chunks, pos = [], 0
    for i in range(ntot):
        row = sorted(new[i] + [i])
        nnb[i] = len(row)
        off[i] = pos
        pos += len(row)
        chunks.append(np.asarray(row, dtype=np.int32))
    lst = np.concatenate(chunks)
    block = dict(idx=idx, xi=x.copy(), vi=v.copy(), freg=freg, fdr=fdr,
                 nnb=nnb, off=off, lst=lst, lists_new=new, lists_old=old,
                 F=(F0, F1, F2, F3))
    return par_kw, state, block


def nlist_rows(block, lmax, ifirst=1):
    """The compact lists as NLIST rows (util_gpu.F:103-110: add IFIRST,
    drop the body itself), shape (ni, lmax)."""
    ni = len(block["idx"])
    rows = np.zeros((ni, lmax), dtype=np.int32)
    for q in range(ni):
        I = int(block["idx"][q])
        o, c = int(block["off"][q]), int(block["nnb"][q])
        mem = [int(j) + ifirst for j in block["lst"][o:o + c]
               if int(j) + ifirst != I]
        rows[q, 0] = len(mem)
        rows[q, 1:1 + len(mem)] = mem
    return rows


def copy_state(state):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v)
            for k, v in state.items()}
