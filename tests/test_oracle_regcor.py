"""Pins of the regular-corrector restatement (oracle/regcor.c, following
src/Main/regcor_gpu.F) against what the mathematics fixes, not against a
retyping of its formulas:

* the new LIST equals the new neighbour set and NBFLUX counts the symmetric
  difference (the merge of regcor_gpu.F:266-326 is a set difference);
* FI / FIDOT after the correction equal direct sums over the new list when
  they were direct sums over the old one (the gain / loss terms of
  regcor_gpu.F:400-438 have the right signs and members);
* for a regular force that is a cubic in time the corrector's D3R, D2R, XN,
  XNDOT are its exact derivatives and Taylor terms (regcor_gpu.F:440-551);
* hand-derived cases of the radius rule (regcor_gpu.F:146-253), including a
  REAL*4 literal (A1 = 1.05 is the float 1.05), and of the step rule;
* the small-step retention keeps the neighbour and moves its force from the
  regular to the irregular part (regcor_gpu.F:328-370).
"""
import numpy as np
import pytest

from oracle import regcor as orc
from tests import regcor_cases as rc


def run(par_kw, state, block, lmax):
    st = rc.copy_state(state)
    par = orc.make_par(**par_kw)
    rows = rc.nlist_rows(block, lmax)
    freg, fdr, rows_out, cnt = orc.regcor(par, st, block["idx"], block["xi"],
                                          block["vi"], block["freg"],
                                          block["fdr"], rows)
    return st, freg, fdr, cnt


def test_merge_is_set_difference_and_irregular_force_follows_the_list():
    par_kw, state, block = rc.build(n=1500, seed=21)
    st, freg, fdr, cnt = run(par_kw, state, block, 600)
    x, v, m = state["x"], state["xdot"], state["body"]
    flux = 0
    for i in range(len(x)):
        old, new = set(block["lists_old"][i]), set(block["lists_new"][i])
        flux += len(old ^ new)
        if old != new:
            n = st["list"][i, 0]
            assert list(st["list"][i, 1:1 + n] - 1) == sorted(new)
        f, fd = rc.pair_sums(x, v, m, i, sorted(new))
        sa = np.sum(m[sorted(new | old)] /
                    np.sum((x[sorted(new | old)] - x[i]) ** 2, axis=1)) \
            if new | old else 0.0
        assert np.allclose(st["fi"][i], f, rtol=0, atol=1e-12 * (sa + 1e-300))
        assert np.array_equal(st["d0"][i], st["fi"][i])
    assert cnt["nbflux"] == flux and flux > 0
    assert cnt["nbvoid"] == 0 and cnt["nbsmin"] == 0


def test_corrector_is_exact_for_a_cubic_regular_force():
    par_kw, state, block = rc.build(n=600, seed=22, change=0.0)
    dtr = 1.0 / 64
    st, freg, fdr, cnt = run(par_kw, state, block, 600)
    F0, F1, F2, F3 = block["F"]
    assert cnt["nbflux"] == 0
    assert np.allclose(st["d3r"], F3, rtol=1e-8, atol=1e-8 * np.abs(F3).max())
    assert np.allclose(st["d2r"], F2 + F3 * dtr, rtol=1e-10, atol=1e-10)
    assert np.array_equal(st["d1r"], block["fdr"])
    assert np.array_equal(st["d0r"], block["freg"])
    assert np.array_equal(st["fr"], block["freg"])
    # XN - X0 = F''(t0r) dt^4/24 + F''' dt^5/120 (the 0.6 of the formula is
    # the float 0.6: relative 4e-8 on the second term)
    dx = st["xn"] - state["x0"]
    want = F2 * dtr ** 4 / 24 + F3 * dtr ** 5 / 120
    # (plus the rounding of X0 + correction: half an ulp of X0)
    assert np.all(np.abs(dx - want) <= 1e-7 * np.abs(F3) * dtr ** 5 / 120
                  + 1e-9 * np.abs(want) + 2.3e-16 * np.abs(state["x0"]))
    dv = st["xndot"] - state["x0dot"]
    wantv = F2 * dtr ** 3 / 6 + F3 * dtr ** 4 / 24
    assert np.allclose(dv, wantv, rtol=1e-9, atol=1e-15)   # FP64 noise
    assert np.all(st["t0r"] == par_kw["time"])


def one_body_case(nnb, nnb0, nnbopt=50, znbmin=5.0, ri=0.05, rs=0.02,
                  rsfac_dtr=0.03, ridot_sign=0.0, nnbmax=500):
    """A single body at distance ri from the density centre (inside the
    core unless ri is large) with nnb new and nnb0 old neighbours placed far
    enough that no other rule fires."""
    n = max(nnb, nnb0) + 1
    lmax = nnbmax + 50
    dtr = 1.0 / 64
    x = np.zeros((n, 3))
    x[0] = (ri, 0.0, 0.0)
    ang = np.linspace(0, 2 * np.pi, n - 1, endpoint=False) if n > 1 else []
    for k in range(1, n):
        x[k] = x[0] + 0.5 * rs * np.array([np.cos(ang[k - 1]),
                                           np.sin(ang[k - 1]), 0.3])
    v = np.zeros((n, 3))
    v[0] = (ridot_sign * 0.1, 0.0, 0.0)
    m = np.full(n, 1.0 / n)
    state = {k: np.zeros((n, 3)) for k in
             ("x0", "x0dot", "xn", "xndot", "fi", "fidot", "fr", "frdot", "d0",
              "d0r", "d1r", "d2r", "d3r")}
    state["fr"][:] = 1.0
    state["frdot"][:] = 0.5
    state["x0"][:] = x
    state["x0dot"][:] = v
    state["t0r"] = np.full(n, 1.0 - dtr)
    state["stepr"] = np.full(n, dtr)
    state["step"] = np.full(n, dtr)
    state["t0"] = np.full(n, 1.0)
    state["timenw"] = state["t0"] + dtr
    state["rs"] = np.full(n, rs)
    state["body"] = m
    LIST = np.zeros((n, lmax), np.int32)
    LIST[0, 0] = nnb0
    LIST[0, 1:1 + nnb0] = np.arange(2, 2 + nnb0)
    state["list"] = LIST
    state["h"] = np.full(1, -1.0)
    state["x"], state["xdot"] = x, v
    par_kw = dict(time=1.0, rscale=1.0, rdens=(0.0, 0.0, 0.0), rc2=0.1 ** 2,
                  rsfac=rsfac_dtr / dtr, znbmin=znbmin, smax=1.0, dtmin=1e-9,
                  etar=0.02, smin=0.0, bodym=float(m.mean()), tidal4=0.0,
                  nnbopt=nnbopt, nnbmax=nnbmax, n=n, ntot=n, ifirst=1,
                  lmax=lmax, nb_flag=1, m_flag=0)
    members = list(range(1, 1 + nnb))
    block = dict(idx=np.array([1], np.int32), xi=x[:1].copy(),
                 vi=v[:1].copy(), freg=np.full((1, 3), 1.0),
                 fdr=np.full((1, 3), 0.5),
                 nnb=np.array([nnb + 1], np.int32),
                 off=np.array([0], np.int32),
                 lst=np.array(sorted(members + [0]), np.int32))
    return par_kw, state, block, lmax


def test_radius_unchanged_at_the_target_count():
    # NNB = NNB0 = NNBOPT inside the core: FAC = 1, volume ratio 1, A1 = 1.
    par_kw, state, block, lmax = one_body_case(50, 50, nnbopt=50)
    st, *_ = run(par_kw, state, block, lmax)
    assert st["rs"][0] == state["rs"][0]


def test_radius_growth_capped_by_rsfac():
    # NNB = NNB0 = 10 against NNBOPT = 50: volume ratio 1.4, capped at
    # A5 = RSFAC*DTR = 0.03 -> A1 = 1 + 0.01 - 0.01^2 (second-order cube root).
    par_kw, state, block, lmax = one_body_case(10, 10, nnbopt=50)
    st, *_ = run(par_kw, state, block, lmax)
    a3 = 0.03 / 3.0
    assert st["rs"][0] == pytest.approx(state["rs"][0] * (1.0 + a3 - a3 * a3),
                                        rel=1e-15)


def test_radius_floor_uses_the_float_literal_1_05():
    # Far outside (RI2 > 100 RH2, no radial factor), NNB = 4 < ZNBMIN = 10
    # while NBP = INT(4 * FAC) = 23 > ZNBMIN: a shrinking A1 is replaced by
    # the REAL*4 constant 1.05 (regcor_gpu.F:209-211).
    par_kw, state, block, lmax = one_body_case(4, 4, nnbopt=200, znbmin=10.0,
                                               ri=12.0, rs=0.5)
    st, *_ = run(par_kw, state, block, lmax)
    assert st["rs"][0] == state["rs"][0] * float(np.float32(1.05))
    assert st["rs"][0] != state["rs"][0] * 1.05


@pytest.mark.parametrize("sign", [+1.0, -1.0])
def test_no_neighbours(sign):
    # NNB = 0: RS -> max(0.75 RS, 0.1 RSCALE) moving out, 1.1 RS moving in
    # (both literals REAL*4 where single), FI gains the centre-mass term.
    par_kw, state, block, lmax = one_body_case(0, 0, ri=2.0, rs=0.2,
                                               ridot_sign=sign)
    st, freg, fdr, cnt = run(par_kw, state, block, lmax)
    rs = state["rs"][0]
    if sign > 0:
        assert st["rs"][0] == max(0.75 * rs, float(np.float32(0.1)) * 1.0)
    else:
        assert st["rs"][0] == float(np.float32(1.1)) * rs
    assert cnt["nbvoid"] == 1
    xi = block["xi"][0]
    r = np.linalg.norm(xi)
    want = -float(np.float32(0.01)) * par_kw["bodym"] * xi / r ** 3
    assert np.allclose(st["fi"][0] - state["fi"][0], want, rtol=1e-14)


def test_small_step_neighbour_is_retained():
    # Old list {2, 3}, new list {3}: body 2 has a step below SMIN and lies
    # inside 2 RS, so it stays a neighbour; its pair force leaves FREG.
    par_kw, state, block, lmax = one_body_case(2, 2)
    block["lst"] = np.array([0, 2], np.int32)        # j = 2 -> body 3
    block["nnb"] = np.array([2], np.int32)
    state["step"][1] = 1e-7
    par_kw["smin"] = 1e-6
    st, freg, fdr, cnt = run(par_kw, state, block, lmax)
    assert list(st["list"][0, :3]) == [2, 2, 3]
    assert cnt["nbsmin"] == 1
    f, fd = rc.pair_sums(state["x"], state["xdot"], state["body"], 0, [1])
    assert np.allclose(freg[0], block["freg"][0] - f, rtol=1e-14)
    assert np.allclose(fdr[0], block["fdr"][0] - fd, rtol=1e-14)
    # no loss is left, so FI is unchanged
    assert np.array_equal(st["fi"][0], state["fi"][0])


def cubic_step_case(dt0_over_dtr, time=1.0, stepr_mult=1.0):
    """A body whose cubic regular force gives a chosen DT0 (the composite
    step of regcor_gpu.F:618-628): F = F0 (along x), F' = F'' = 0 at T0R and
    F''' = f3: at TIME, F' = f3 s^2/2, F'' = f3 s, and
    W0 = ETAR (|F| |F''| + |F'|^2) / (|F'| |F'''| + |F''|^2)."""
    par_kw, state, block, lmax = one_body_case(10, 10, nnbopt=10)
    dtr = 1.0 / 64
    state["stepr"][:] = dtr * stepr_mult
    f3 = 1.0
    # choose F0 so that DT0 = dt0_over_dtr * dtr
    s = dtr
    fp, fpp = f3 * s * s / 2, f3 * s
    target = (dt0_over_dtr * dtr) ** 2 / par_kw["etar"]
    F0 = (target * (fp * f3 + fpp ** 2) - fp ** 2) / fpp
    # The cubic through T0R with F = F0, F' = fp, F'' = fpp at TIME:
    # F(s') = a + b s' + c s'^2/2 + f3 s'^3/6 (b = c = 0 here).
    a = F0 - fp * s + fpp * s * s / 2 - f3 * s ** 3 / 6
    b = fp - fpp * s + f3 * s * s / 2
    c = fpp - f3 * s
    state["fr"][0] = (a, 0.0, 0.0)
    state["frdot"][0] = (b, 0.0, 0.0)
    block["freg"] = np.array([[a + b * s + c * s * s / 2 + f3 * s ** 3 / 6,
                               0.0, 0.0]])
    block["fdr"] = np.array([[b + c * s + f3 * s * s / 2, 0.0, 0.0]])
    par_kw["time"] = time
    state["t0r"][:] = time - dtr
    return par_kw, state, block, lmax, dtr


@pytest.mark.parametrize("ratio,time,want", [
    # DT0 > 2 STEPR with TIME a multiple of 2 DTR: 2 DTR, doubled while
    # DT0 > 2 TTMP and TIME is a multiple of 2 TTMP: 7 > 4 -> 4 DTR, 7 > 8 no.
    (7.0, 1.0, 4.0),
    (1.5, 1.0, 1.0),        # DTR <= DT0 <= 2 STEPR: unchanged
    (0.75, 1.0, 0.5),       # DT0 < DTR: halved
    (0.3, 1.0, 0.25),       # DT0 < DTR / 2: halved twice
    (7.0, 1.0 + 1.0 / 64, 1.0),   # TIME not a multiple of 2 DTR: unchanged
    # 2 -> 4 -> 8 -> 16 DTR = 0.25 > 0.1: Makino's test (regcor_gpu.F:714-722)
    # DTJ = DTR (1e-4 DTR^2 |F|^2 / |v - v_new|^2)^0.1 decides (None: below).
    (20.0, 1.0, None),
])
def test_step_rule(ratio, time, want):
    par_kw, state, block, lmax, dtr = cubic_step_case(ratio, time=time)
    st, *_ = run(par_kw, state, block, lmax)
    if want is None:
        dv2 = np.sum((block["vi"][0] - st["xndot"][0]) ** 2)
        dtj = dtr * (1e-4 * dtr ** 2 * np.sum(block["freg"][0] ** 2) / dv2) ** 0.1
        want = 1.0 if dtj < 16 * dtr else 16.0
        assert want == 1.0        # this case has a tiny velocity change
    assert st["stepr"][0] == want * dtr
    # the irregular step is halved until it is below the new regular step
    s = state["step"][0]
    while want * dtr < s:
        s *= 0.5
    assert st["step"][0] == s
