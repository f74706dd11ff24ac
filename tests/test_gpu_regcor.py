"""Device regular corrector (gpunb_regcor_dev) against the CPU restatement
of REGCOR_GPU (oracle/regcor.c).

Bar: every output identical, bit for bit.  Both sides evaluate the same FP64
operations in the same order without FMA contraction, and division and sqrt
are correctly rounded on both (pow enters only a comparison).  Also the
chain regf -> corrector entirely on the device at N = 131072.
"""
import numpy as np
import pytest

from nbody6ppgpu_amd import gpunb, regcor
from nbody6ppgpu_amd.hiprt import DeviceArray
from oracle import regcor as orc
from tests import regcor_cases as rc

pytestmark = pytest.mark.gpu

EXACT = ("step", "stepr", "timenw", "rs", "t0r")
FLOATS = ("xn", "xndot", "fi", "fidot", "fr", "frdot", "d0", "d0r", "d1r",
          "d2r", "d3r")


def oracle_run(par_kw, state, block, lmax):
    st = rc.copy_state(state)
    rows = rc.nlist_rows(block, lmax)
    freg, fdr, _, cnt = orc.regcor(orc.make_par(**par_kw), st, block["idx"],
                                   block["xi"], block["vi"], block["freg"],
                                   block["fdr"], rows)
    return st, freg, fdr, cnt


def device_run(par_kw, state, block):
    n = state["body"].shape[0]
    gpunb.gpunb_open(n + 16)
    try:
        gpunb.gpunb_send(state["body"], state["x"], state["xdot"])
        dev = regcor.DeviceState(state)
        freg, fdr, cnt = regcor.correct_host(
            regcor.make_par(**par_kw), dev, block["idx"], block["xi"],
            block["vi"], block["freg"], block["fdr"], block["nnb"],
            block["off"], block["lst"])
        out = dev.numpy()
    finally:
        gpunb.gpunb_close()
    return out, freg, fdr, cnt


def compare(got, want, tag):
    g, gf, gd, gc = got
    w, wf, wd, wc = want
    assert gc == wc, (tag, gc, wc)
    if not np.array_equal(g["list"], w["list"]):
        rows = np.nonzero(np.any(g["list"] != w["list"], axis=1))[0]
        info = []
        for r in rows[:4]:
            k = int(np.nonzero(g["list"][r] != w["list"][r])[0][0])
            info.append((int(r), int(w["list"][r, 0]), int(g["list"][r, 0]), k,
                         w["list"][r, max(k - 2, 0):k + 3].tolist(),
                         g["list"][r, max(k - 2, 0):k + 3].tolist()))
        raise AssertionError((tag, "LIST rows differ", len(rows), info))
    for k in EXACT:
        assert np.array_equal(g[k], w[k]), (tag, k)
    total = 0
    for k, a, b in [(k, g[k], w[k]) for k in FLOATS] + [("freg", gf, wf),
                                                        ("fdr", gd, wd)]:
        if not np.array_equal(a, b):
            rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
            raise AssertionError((tag, k, int(np.sum(a != b)), float(rel.max())))
        total += a.size
    print(f"{tag}: {total} FP64 outputs bit-identical, counters {gc}")


@pytest.mark.parametrize("variant", ["nb_flag1", "nb_flag0_mflag1_tidal"])
def test_block_matches_oracle(variant):
    if variant == "nb_flag1":
        par_kw, state, block = rc.build(n=4096, seed=31, ncm=64, smin=1e-4)
    else:
        # radii about the origin, mass-weighted lists (no short-list rule),
        # a tidal FIDOT term, and c.m. bodies with small |H| so that their
        # perturber-range branch (regcor_gpu.F:187-199) runs
        par_kw, state, block = rc.build(n=4096, seed=31, ncm=256, smin=1e-4,
                                        hrange=(1e-5, 1e-3), nb_flag=0,
                                        m_flag=1, tidal4=0.3)
    rng = np.random.default_rng(3)
    # small steps for a few bodies so that the retention path runs
    state["step"][rng.choice(4096, 200, replace=False)] = 1e-5
    # three bodies without neighbours (the list is the body itself)
    nnb, off, lst = block["nnb"].copy(), block["off"].copy(), block["lst"]
    rows = [lst[off[i]:off[i] + nnb[i]] for i in range(4096)]
    for i in (7, 900, 4000):
        rows[i] = np.array([i], np.int32)
    # lists that shrink while gaining a member beyond the last old one (the
    # reference's walk leaves its sentinel in the old row then)
    for i in range(11, 4096, 97):
        old = state["list"][i, 1:1 + state["list"][i, 0]]
        if len(old) < 6 or old[-1] >= 4096:
            continue
        new = [int(j) - 1 for j in old[:2]] + [int(old[-1])]   # 0-based
        rows[i] = np.array(sorted(set(new + [i])), np.int32)
    pos = 0
    for i, r in enumerate(rows):
        nnb[i], off[i] = len(r), pos
        pos += len(r)
    block.update(nnb=nnb, off=off, lst=np.concatenate(rows))
    want = oracle_run(par_kw, state, block, par_kw["lmax"])
    got = device_run(par_kw, state, block)
    assert want[3]["nbvoid"] == 3 and want[3]["nbsmin"] > 0
    assert want[3]["nbflux"] > 0 and want[3]["niconv"] > 0
    compare(got, want, f"n=4096 with c.m. bodies, {variant}")


def test_lds_and_hbm_workgroups_and_retention():
    """The library keeps a workgroup's 64 NLIST rows in LDS when they fit
    5120 ints (row of a body: new list incl. itself + 2) and in HBM
    otherwise.  Lists of ~64 give both kinds of workgroup in one call;
    bodies whose three nearest old neighbours have small steps while two
    members are gained keep those neighbours by the small-step retention
    (regcor_gpu.F:328-370), which grows NLIST past its LDS row (it then
    moves to HBM), in both kinds of workgroup."""
    from scipy.spatial import cKDTree
    n = 3072
    par_kw, state, block = rc.build(n=n, seed=34, nnbopt=64, smin=1e-4)
    x = block["xi"]
    tree = cKDTree(x)
    nnb, off, lst = block["nnb"].copy(), block["off"].copy(), block["lst"]
    rows = [lst[off[i]:off[i] + nnb[i]] for i in range(n)]
    grown = []
    for i in range(5, n, 32):
        old = [int(j) - 1 for j in state["list"][i, 1:1 + state["list"][i, 0]]]
        if len(old) < 8:
            continue
        d = np.linalg.norm(x[old] - x[i], axis=1)
        lost = [old[k] for k in np.argsort(d)[:3]]
        _, cand = tree.query(x[i], k=len(old) + 12)
        gain = [int(j) for j in cand if j != i and j not in old][:2]
        state["step"][lost] = 1e-5
        rows[i] = np.array(sorted(set(old) - set(lost) | set(gain) | {i}),
                           np.int32)
        grown.append(i)
    pos = 0
    for i, r in enumerate(rows):
        nnb[i], off[i] = len(r), pos
        pos += len(r)
    block.update(nnb=nnb, off=off, lst=np.concatenate(rows))
    want = oracle_run(par_kw, state, block, par_kw["lmax"])
    # the cases this call holds, from its inputs and the oracle's lists
    tot = (nnb + 2).reshape(-1, 64).sum(axis=1)
    lds = tot <= 5120
    grown = np.array(grown)
    kept = grown[want[0]["list"][grown, 0] > nnb[grown] - 1]
    assert 0 < lds.sum() < len(tot), tot
    assert np.any(lds[kept // 64]) and np.any(~lds[kept // 64]), kept
    got = device_run(par_kw, state, block)
    compare(got, want, f"n={n}, {int(lds.sum())} LDS and "
            f"{int((~lds).sum())} HBM workgroups, {len(kept)} of the "
            "modified lists grown by the retention")


def test_subset_block_and_repeat():
    """A block that is a strided subset of the bodies, corrected twice in a
    row (the second call sees the first call's state)."""
    par_kw, state, block = rc.build(n=3000, seed=32)
    sel = np.arange(5, 3000, 3)
    sub = dict(block)
    for k in ("xi", "vi", "freg", "fdr"):
        sub[k] = block[k][sel]
    sub["idx"] = block["idx"][sel]
    rows = [block["lst"][block["off"][i]:block["off"][i] + block["nnb"][i]]
            for i in sel]
    sub["nnb"] = np.array([len(r) for r in rows], np.int32)
    sub["off"] = np.concatenate([[0], np.cumsum(sub["nnb"])[:-1]]).astype(np.int32)
    sub["lst"] = np.concatenate(rows)
    st1, *_ = w1 = oracle_run(par_kw, state, sub, par_kw["lmax"])
    n = state["body"].shape[0]
    gpunb.gpunb_open(n + 16)
    try:
        gpunb.gpunb_send(state["body"], state["x"], state["xdot"])
        dev = regcor.DeviceState(state)
        par = regcor.make_par(**par_kw)
        g1 = regcor.correct_host(par, dev, sub["idx"], sub["xi"], sub["vi"],
                                 sub["freg"], sub["fdr"], sub["nnb"],
                                 sub["off"], sub["lst"])
        compare((dev.numpy(),) + g1, w1, "subset, first call")
        par_kw2 = dict(par_kw, time=par_kw["time"] + 1.0 / 64)
        w2 = oracle_run(par_kw2, st1, sub, par_kw["lmax"])
        g2 = regcor.correct_host(regcor.make_par(**par_kw2), dev, sub["idx"],
                                 sub["xi"], sub["vi"], sub["freg"],
                                 sub["fdr"], sub["nnb"], sub["off"],
                                 sub["lst"])
        compare((dev.numpy(),) + g2, w2, "subset, second call")
    finally:
        gpunb.gpunb_close()


def test_invalid_bodies_are_reported_and_left_unchanged():
    par_kw, state, block = rc.build(n=500, seed=33)
    n = 500
    gpunb.gpunb_open(n + 16)
    try:
        gpunb.gpunb_send(state["body"], state["x"], state["xdot"])
        dev = regcor.DeviceState(state)
        par = regcor.make_par(**par_kw)
        idx = block["idx"].copy()
        idx[10] = idx[11]                       # body 12 listed twice
        with pytest.raises(gpunb.GpunbError, match="regular corrector"):
            regcor.correct_host(par, dev, idx, block["xi"], block["vi"],
                                block["freg"], block["fdr"], block["nnb"],
                                block["off"], block["lst"])
        out = dev.numpy()
        for k in ("fi", "fr", "xn", "rs", "step", "stepr"):
            assert np.array_equal(out[k][11], state[k][11]), k
        # an unsorted new list
        lst = block["lst"].copy()
        o = int(block["off"][20])
        k = next(k for k in range(int(block["nnb"][20]) - 1)
                 if 20 not in (lst[o + k], lst[o + k + 1]))
        lst[o + k], lst[o + k + 1] = lst[o + k + 1], lst[o + k]
        dev2 = regcor.DeviceState(state)
        with pytest.raises(gpunb.GpunbError, match="regular corrector"):
            regcor.correct_host(par, dev2, block["idx"], block["xi"],
                                block["vi"], block["freg"], block["fdr"],
                                block["nnb"], block["off"], lst)
        assert np.array_equal(dev2.numpy()["fi"][20], state["fi"][20])
        # an old LIST row out of order, and one with a member outside the j
        # set: both reported, both bodies unchanged (LIST row included)
        bad_state = rc.copy_state(state)
        r30 = bad_state["list"][30]
        if r30[0] >= 3:
            r30[2], r30[3] = r30[3], r30[2]
        bad_state["list"][31, 1 + bad_state["list"][31, 0] // 2] = n + 7
        dev3 = regcor.DeviceState(bad_state)
        with pytest.raises(gpunb.GpunbError, match="regular corrector"):
            regcor.correct_host(par, dev3, block["idx"], block["xi"],
                                block["vi"], block["freg"], block["fdr"],
                                block["nnb"], block["off"], block["lst"])
        out3 = dev3.numpy()
        for i in (30, 31):
            for k in ("fi", "fr", "xn", "rs", "step", "stepr", "list"):
                assert np.array_equal(out3[k][i], bad_state[k][i]), (i, k)
        # host-side preconditions
        bad = regcor.make_par(**dict(par_kw, lmax=par_kw["nnbmax"] + 2))
        with pytest.raises(gpunb.GpunbError, match="lmax"):
            regcor.correct_host(bad, dev2, block["idx"], block["xi"],
                                block["vi"], block["freg"], block["fdr"],
                                block["nnb"], block["off"], block["lst"])
        bad = regcor.make_par(**dict(par_kw, ntot=n + 100))
        with pytest.raises(gpunb.GpunbError, match="j-particles"):
            regcor.correct_host(bad, dev2, block["idx"], block["xi"],
                                block["vi"], block["freg"], block["fdr"],
                                block["nnb"], block["off"], block["lst"])
        # Options whose REGCOR_GPU branches are not on the device are
        # refused, nothing changed.
        before = dev2.numpy()
        for kz in ("kz13", "kz14", "kz37", "kz38"):
            bad = regcor.make_par(**dict(par_kw, **{kz: 1}))
            with pytest.raises(gpunb.GpunbError, match="KZ"):
                regcor.correct_host(bad, dev2, block["idx"], block["xi"],
                                    block["vi"], block["freg"],
                                    block["fdr"], block["nnb"], block["off"],
                                    block["lst"])
        after = dev2.numpy()
        for k in before:
            assert np.array_equal(after[k], before[k]), k
    finally:
        gpunb.gpunb_close()


def test_list_rows_not_16_byte_aligned():
    """LMAX not a multiple of 4: the old LIST rows are not 16-byte aligned,
    so the merge loads them one entry at a time."""
    par_kw, state, block = rc.build(n=1500, seed=35, lmax=603, smin=1e-4)
    rng = np.random.default_rng(5)
    state["step"][rng.choice(1500, 80, replace=False)] = 1e-5
    want = oracle_run(par_kw, state, block, par_kw["lmax"])
    got = device_run(par_kw, state, block)
    assert want[3]["nbflux"] > 0
    compare(got, want, "n=1500, lmax 603")


def test_regf_to_corrector_chain_at_131072():
    """gpunb_regf_dev lists feed gpunb_regcor_dev without leaving the
    device (King W0=6, N = 131072, every body active); checked against the
    oracle on the same lists."""
    import bench
    c, h2, stepr = bench.workload()
    n = c.n
    lmax, nnbmax = 600, 550
    rng = np.random.default_rng(7)
    dtr = 1.0 / 64
    gpunb.gpunb_open(n + 16)
    try:
        gpunb.gpunb_send(c.m, c.x, c.v)
        d = {k: DeviceArray.from_numpy(a) for k, a in
             (("x", c.x), ("v", c.v), ("dtr", np.full(n, dtr)))}
        outs = {}
        for tag, fac in (("old", 0.9), ("new", 1.0)):
            dh2 = DeviceArray.from_numpy(h2 * fac * fac)
            acc = DeviceArray((n, 3), np.float64)
            jrk = DeviceArray((n, 3), np.float64)
            pot = DeviceArray((n,), np.float64)
            nnb = DeviceArray((n,), np.int32)
            off = DeviceArray((n,), np.int32)
            lst = DeviceArray((n * nnbmax,), np.int32)
            tot = DeviceArray((1,), np.int32)
            rc_ = gpunb.load().gpunb_regf_dev(
                n, dh2.ptr, d["dtr"].ptr, d["x"].ptr, d["v"].ptr, acc.ptr,
                jrk.ptr, pot.ptr, nnbmax, 0, nnb.ptr, 0, lst.ptr, off.ptr,
                tot.ptr, n * nnbmax)
            assert rc_ == 0
            gpunb.gpunb_sync()
            outs[tag] = dict(acc=acc, jrk=jrk, nnb=nnb, off=off, lst=lst)
        o_nnb, o_off, o_lst = (outs["old"][k].numpy() for k in ("nnb", "off", "lst"))
        assert np.all(o_nnb > 0)
        LIST = np.zeros((n, lmax), np.int32)
        for i in range(n):
            row = o_lst[o_off[i]:o_off[i] + o_nnb[i]] + 1
            row = row[row != i + 1]
            LIST[i, 0] = len(row)
            LIST[i, 1:1 + len(row)] = row
        state = {k: rng.normal(0, 1, (n, 3)) for k in
                 ("x0", "x0dot", "fi", "fidot", "fr", "frdot")}
        for k in ("xn", "xndot", "d0", "d0r", "d1r", "d2r", "d3r"):
            state[k] = np.zeros((n, 3))
        state["x0"] = c.x.copy()
        state["x0dot"] = c.v.copy()
        state.update(t0r=np.full(n, 1.0 - dtr), stepr=np.full(n, dtr),
                     step=dtr / 2.0 ** rng.integers(0, 5, n),
                     t0=np.full(n, 1.0), rs=np.sqrt(h2) * 0.9,
                     body=c.m.copy(), list=LIST, x=c.x, xdot=c.v)
        state["timenw"] = state["t0"] + state["step"]
        par_kw = dict(time=1.0, rscale=1.0, rdens=(0.0, 0.0, 0.0),
                      rc2=0.05 ** 2, rsfac=2.0, znbmin=10.0, smax=1.0,
                      dtmin=1e-9, etar=0.02, smin=0.0, bodym=float(c.m.mean()),
                      tidal4=0.0, nnbopt=56, nnbmax=nnbmax, n=n, ntot=n,
                      ifirst=1, lmax=lmax, nb_flag=1, m_flag=0)
        nw = outs["new"]
        freg_d, fdr_d = nw["acc"], nw["jrk"]
        freg0, fdr0 = freg_d.numpy(), fdr_d.numpy()
        dev = regcor.DeviceState(state)
        idx = DeviceArray.from_numpy(np.arange(1, n + 1, dtype=np.int32))
        cnt = DeviceArray.from_numpy(np.zeros(4, np.int32))
        regcor.correct(regcor.make_par(**par_kw), dev, n, idx.ptr,
                       d["x"].ptr, d["v"].ptr, freg_d.ptr, fdr_d.ptr,
                       nw["nnb"].ptr, nw["off"].ptr, nw["lst"].ptr,
                       n * nnbmax, cnt.ptr)
        gpunb.gpunb_sync()
        got = (dev.numpy(), freg_d.numpy(), fdr_d.numpy(),
               dict(zip(gpunb.REGCOR_COUNTERS, cnt.numpy().tolist())))
        block = dict(idx=np.arange(1, n + 1, dtype=np.int32), xi=c.x, vi=c.v,
                     freg=freg0, fdr=fdr0, nnb=nw["nnb"].numpy(),
                     off=nw["off"].numpy(), lst=nw["lst"].numpy())
    finally:
        gpunb.gpunb_close()
    want = oracle_run(par_kw, state, block, lmax)
    assert want[3]["nbflux"] > n          # the lists did change
    compare(got, want, "King N=131072 chain")
