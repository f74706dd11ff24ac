"""The reference's Fortran host binds to this library unchanged.

CPU: the ISO_C_BINDING module and the implicit-interface calls both compile
with amdflang and resolve to exactly the symbols the library exports.
GPU: the Fortran driver runs gpunb_open_/send_/regf_/close_ and gpupot_ in
both binding styles, and the irregular-force library as intgrt.F drives it
(IRR_SIMD_SET_JP / SET_LIST for every particle inside an OpenMP parallel
loop, then FIRR_VEC); results are bit-identical to the same calls made from
Python through the C-ABI and meet the oracle bar.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from nbody6ppgpu_amd import gpunb, ics

HERE = os.path.dirname(os.path.abspath(__file__))
DRV = os.path.join(HERE, "fortran", "_build", "drv")


def have_flang():
    return bool(shutil.which("amdflang") or os.path.exists("/opt/rocm/bin/amdflang"))


@pytest.mark.skipif(not have_flang(), reason="amdflang not installed")
def test_fortran_driver_links_to_library_symbols():
    if not os.path.exists(DRV):
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "fortran")])
    und = subprocess.run(["nm", "-u", DRV], capture_output=True, text=True).stdout
    needed = {l.split()[-1] for l in und.splitlines()
              if "gpu" in l or "irr_simd" in l}
    assert needed == {"gpunb_open_", "gpunb_send_", "gpunb_regf_",
                      "gpunb_close_", "gpupot_", "irr_simd_open_",
                      "irr_simd_set_jp_", "irr_simd_set_list_",
                      "irr_simd_firr_vec_", "irr_simd_close_"}
    exported = subprocess.run(["nm", "-D", "--defined-only",
                               gpunb.library_path()], capture_output=True,
                              text=True).stdout
    names = {l.split()[-1] for l in exported.splitlines()}
    assert needed <= names


@pytest.mark.gpu
def test_fortran_driver_matches_c_abi(tmp_path):
    from oracle import oracle
    from oracle.parity import assert_parity
    assert os.path.exists(DRV), "build the Fortran driver (build())"
    c = ics.plummer(1500, seed=301, m=ics.kroupa(1500, seed=302))
    rs, dtr = ics.neighbour_radii(c, k=40)
    h2 = rs * rs
    n, lmax, nnbmax, mflag = c.n, 301, 300, 0
    fin = tmp_path / "in.bin"
    fout = tmp_path / "out.bin"
    # Irregular-force block: predictor state, lists (1-based, LIST(LMAX2,N)
    # rows: count then members) and the active particles.
    rng = np.random.default_rng(303)
    fa = rng.normal(0, 0.5, (n, 3))
    fdot = rng.normal(0, 2.0, (n, 3))
    t0 = -rng.uniform(0, 1 / 256, n)
    from scipy.spatial import cKDTree
    _, idx = cKDTree(c.x).query(c.x, k=33)
    lists = [np.sort(r[1:]) for r in idx]
    lists[7] = np.zeros(0, dtype=int)
    lmax2, tirr = 64, 0.002
    L2 = np.zeros((n, lmax2), dtype=np.int32)
    for i, l in enumerate(lists):
        L2[i, 0] = len(l)
        L2[i, 1:1 + len(l)] = l + 1
    act = np.arange(1, n + 1, 3, dtype=np.int32)
    with open(fin, "wb") as f:
        np.array([n, lmax, nnbmax, mflag], dtype=np.int32).tofile(f)
        for a in (c.m, c.x, c.v, h2, dtr):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        np.array([lmax2, act.size], dtype=np.int32).tofile(f)
        np.array([tirr], dtype=np.float64).tofile(f)
        for a in (fa, fdot, t0):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        L2.tofile(f)
        act.tofile(f)
    env = dict(os.environ, GPUNB_QUIET="1")
    subprocess.run([DRV, str(fin), str(fout)], check=True, env=env, timeout=120)
    raw = open(fout, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    outs = []
    for _ in range(2):
        acc = take(np.float64, 3 * n).reshape(n, 3)
        jrk = take(np.float64, 3 * n).reshape(n, 3)
        pot = take(np.float64, n)
        lst = take(np.int32, lmax * n).reshape(n, lmax)
        outs.append((acc, jrk, pot, lst))
    phi = take(np.float64, n)
    ni2 = act.size
    firr = take(np.float64, 3 * ni2).reshape(ni2, 3)
    fdirr = take(np.float64, 3 * ni2).reshape(ni2, 3)
    iminr = take(np.int32, ni2)
    assert off == len(raw)
    gpunb.gpunb_open(n + 10)
    try:
        gpunb.gpunb_send(c.m, c.x, c.v)
        py = gpunb.gpunb_regf(h2, dtr, c.x, c.v, lmax, nnbmax, mflag)
    finally:
        gpunb.gpunb_close()
    for fo in outs:          # module call, then implicit-interface call
        for a, b in zip(fo[:3], py[:3]):
            assert np.array_equal(a, b)
        for i in range(n):
            k = py[3][i, 0]
            assert fo[3][i, 0] == k
            assert np.array_equal(fo[3][i, 1:1 + k], py[3][i, 1:1 + k])
    o = oracle.regf(c.m, c.x, c.v, h2, dtr, c.x, c.v, lmax, nnbmax, mflag)
    nn = np.array([-1] * n, dtype=np.int32)
    assert_parity(outs[1] + (o[4],), o, label="Fortran driver",
                  inputs=(c.m, c.x, c.v, c.x, c.v))
    op = oracle.pot(1, n, c.m, c.x)
    assert np.max(np.abs(phi - op) / op) < 1e-6
    # Irregular force: the Fortran OpenMP staging gives the batched C-ABI's
    # bits, and both meet the oracle (identical nearest neighbours).
    gpunb.irr_open(n, lmax2)
    try:
        addr = np.arange(1, n + 1)
        gpunb.irr_set_jp(addr, c.x, c.v, fa, fdot, c.m, t0)
        gpunb.irr_set_lists(addr, [l + 1 for l in lists])
        pa, pj, pn = gpunb.irr_firr(tirr, act)
    finally:
        gpunb.irr_close()
    assert np.array_equal(firr, pa) and np.array_equal(fdirr, pj)
    assert np.array_equal(iminr, pn)
    oa, oj, on = oracle.irr(tirr, act - 1, t0, c.x, c.v, fa, fdot, c.m, lists)
    assert np.array_equal(iminr, on)
    scale = np.linalg.norm(oa, axis=1) + 1e-300
    assert np.all(np.linalg.norm(firr - oa, axis=1) <= 1e-10 * scale + 1e-12 * np.max(scale))
    del nn


HOST_LOOP = os.path.join(HERE, "fortran", "_build", "host_loop")


def run_host_loop(tmp_path, c, ifirst, rs, stepr, ireg, lmax, nnbmax, nnbopt,
                  m_flag, nb_flag):
    ntot = c.n
    bodym = float(np.mean(c.m))
    fin, fout = tmp_path / "hl_in.bin", tmp_path / "hl_out.bin"
    with open(fin, "wb") as f:
        np.array([ntot, ifirst, lmax, nnbmax, nnbopt, m_flag, nb_flag,
                  len(ireg)], dtype=np.int32).tofile(f)
        np.array([bodym], dtype=np.float64).tofile(f)
        for a in (c.m, c.x, c.v, rs, stepr):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        np.ascontiguousarray(ireg, dtype=np.int32).tofile(f)
    env = dict(os.environ, GPUNB_QUIET="1")
    subprocess.run([HOST_LOOP, str(fin), str(fout)], check=True, env=env,
                   timeout=120)
    raw = open(fout, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a
    rs_out = take(np.float64, ntot)
    phi = take(np.float64, ntot)
    freg = take(np.float64, 3 * ntot).reshape(ntot, 3)
    fdr = take(np.float64, 3 * ntot).reshape(ntot, 3)
    lst = take(np.int32, lmax * ntot).reshape(ntot, lmax)
    nretry = int(take(np.int32, 1)[0])
    assert off == len(raw)
    return rs_out, phi, freg, fdr, lst, nretry, bodym


@pytest.mark.gpu
@pytest.mark.parametrize("m_flag,nb_flag", [(0, 1), (1, 0)])
def test_fortran_host_loop_overflow_retry(tmp_path, m_flag, nb_flag):
    """The util_gpu.F:25-112 call sequence from Fortran: 1024-i chunks of
    an active list, overflow -> shrink RS and redo the chunk, IFIRST offset
    and self removal, PHIDBL.  Every accepted row equals the oracle's list
    for the body's final RS (plus IFIRST, without the body), and PHIDBL,
    FREG, FDR meet the plain bar against the oracle at the same radii."""
    from oracle import oracle
    from oracle.parity import FORCE_TOL, rel_err
    assert os.path.exists(HOST_LOOP), "build the Fortran drivers (build())"
    m = ics.kroupa(4000, seed=312) if m_flag else None
    c = ics.plummer(4000, seed=311, m=m)
    ifirst = 21                  # bodies 1..20 stand for KS components
    rs, stepr = ics.neighbour_radii(c, k=48)
    rng = np.random.default_rng(313)
    ireg = np.sort(rng.choice(np.arange(ifirst, c.n + 1), 2500,
                              replace=False)).astype(np.int32)
    lmax, nnbmax, nnbopt = 64, 60, 40
    rs_out, phi, freg, fdr, lst, nretry, bodym = run_host_loop(
        tmp_path, c, ifirst, rs, stepr, ireg, lmax, nnbmax, nnbopt, m_flag,
        nb_flag)
    assert nretry > 0, "the radii were chosen so that some lists overflow"
    shrunk = rs_out < rs
    assert np.all(rs_out <= rs) and shrunk.sum() > 0
    assert np.array_equal(rs_out[~shrunk], rs[~shrunk])
    js = slice(ifirst - 1, None)
    idx = ireg - 1
    h2 = rs_out[idx] ** 2
    if m_flag:
        h2 = h2 / bodym
    o = oracle.regf(c.m[js], c.x[js], c.v[js], h2, stepr[idx], c.x[idx],
                    c.v[idx], lmax, nnbmax, m_flag)
    assert np.all(o[3][:, 0] >= 0) and np.all(o[3][:, 0] <= nnbmax)
    for r, i in enumerate(idx):
        k = o[3][r, 0]
        want = o[3][r, 1:1 + k] + ifirst
        want = want[want != i + 1]
        assert lst[i, 0] == len(want), f"row of body {i + 1}"
        assert np.array_equal(lst[i, 1:1 + len(want)], want), f"body {i + 1}"
    for got, ref in ((freg[idx], o[0]), (fdr[idx], o[1]), (phi[idx], o[2])):
        assert rel_err(got, ref).max() < FORCE_TOL
    inactive = np.setdiff1d(np.arange(c.n), idx)
    assert np.all(lst[inactive] == 0) and np.all(phi[inactive] == 0)
