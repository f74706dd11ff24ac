"""Device energy sum (gpunb_energy_dev) against the CPU restatement of the
GPU branch of ENERGY with PHICOR (oracle/energy.c).

Bar: with the same PHIDBL input, the corrected PHIDBL of every single and
KS component identical bit for bit (same FP64 operations in the same order);
POT and ZKIN within 1e-13 relative (the device sums in a fixed tree order,
the reference serially).  With the device's own potential (the FP32-pair
gpupot kernels) PHIDBL and POT within the force bar, 1e-6 relative.
"""
import numpy as np
import pytest

from nbody6ppgpu_amd import energy, gpunb
from oracle import energy as orc
from tests import energy_cases as ec

pytestmark = pytest.mark.gpu


def oracle_run(s, cut, phi):
    return orc.energy(s["n"], s["ntot"], s["lmax"], cut, s["body"], s["x"],
                      s["xdot"], s["list"], phi)


def device_run(s, cut, phi):
    return energy.potential_energy(s["n"], s["ntot"], s["lmax"], cut,
                                   s["body"], s["x"], s["xdot"], s["list"],
                                   phi)


@pytest.mark.parametrize("cut", [0.02 ** 2, 0.1 ** 2, 1e30])
def test_corrections_match_oracle_bit_for_bit(cut):
    s = ec.build(nsingle=3000, npairs=200, seed=11, rnb=0.25, rpert=0.3)
    phi0 = ec.gpupot_phidbl(s)
    pw, potw, kinw = oracle_run(s, cut, phi0)
    pg, potg, king = device_run(s, cut, phi0)
    assert np.array_equal(pg, pw)
    assert potg == pytest.approx(potw, rel=1e-13)
    assert king == pytest.approx(kinw, rel=1e-13)
    assert not np.array_equal(pw, phi0)         # corrections were applied


def test_with_the_device_potential():
    s = ec.build(nsingle=3000, npairs=200, seed=12)
    cut = 0.1 ** 2
    pw, potw, _ = oracle_run(s, cut, ec.gpupot_phidbl(s))
    pg, potg, _ = device_run(s, cut, None)
    used = np.r_[0:s["ntot"]]
    assert np.all(np.abs(pg[used] - pw[used]) <= 1e-6 * np.abs(pw[used]))
    assert potg == pytest.approx(potw, rel=1e-6)


def test_no_pairs():
    s = ec.build(nsingle=4000, npairs=0, seed=13)
    phi0 = ec.gpupot_phidbl(s)
    pg, potg, king = device_run(s, 1e30, phi0)
    assert np.array_equal(pg, phi0)
    assert potg == pytest.approx(ec.resolved_potential(s), rel=1e-12)


def test_large_system_against_the_abi_potential():
    """65536 singles and 4096 pairs: the device potential equals gpupot_'s,
    and the corrections on top of it match the oracle bit for bit."""
    s = ec.build(nsingle=65536, npairs=4096, seed=14, rnb=0.05, rpert=0.08)
    ifirst, ntot = s["ifirst"], s["ntot"]
    phi_abi = np.zeros(ntot)
    phi_abi[ifirst - 1:] = gpunb.gpupot(1, ntot - ifirst + 1,
                                        s["body"][ifirst - 1:],
                                        s["x"][ifirst - 1:])
    cut = 0.03 ** 2
    pw, potw, kinw = oracle_run(s, cut, phi_abi)
    pg, potg, king = device_run(s, cut, None)
    assert np.array_equal(pg, pw)
    assert potg == pytest.approx(potw, rel=1e-13)
    assert king == pytest.approx(kinw, rel=1e-13)


def test_invalid_row_is_reported():
    s = ec.build(nsingle=200, npairs=5, seed=15)
    # the walk starts at the last member of a single's row
    cnt = max(int(s["list"][20, 0]), 1)
    s["list"][20, 0] = cnt
    s["list"][20, cnt] = s["ntot"] + 7
    with pytest.raises(gpunb.GpunbError, match="energy sum"):
        device_run(s, 1.0, ec.gpupot_phidbl(s))
    # host-side precondition: rejected before any device work
    par = energy.make_par(s["n"], s["ntot"], s["lmax"], 1.0)
    par.ifirst = 3
    with pytest.raises(gpunb.GpunbError, match="ifirst"):
        energy.energy_dev(par, 1, 1, 1, 1, False, 1, 1)
