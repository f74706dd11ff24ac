"""Writes the golden vectors of tests/golden/*.npz from the reference's own
CPU regular-force and potential code (src/Main/reg.avx.cpp, reg.sse.cpp,
pot.avx.cpp), compiled from its sources by oracle/Makefile into oracle/_ref/.

The reference tree holds no test vectors of its own; these fixtures are the
outputs of its CPU path on seeded synthetic inputs, stored with the inputs.
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nbody6ppgpu_amd import ics  # noqa: E402
from oracle import ref  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def regf_case(name, c, k, lmax, nnbmax, m_flag=0):
    rs, stepr = ics.neighbour_radii(c, k=k)
    h2 = rs * rs
    if m_flag:
        h2 = h2 / np.mean(c.m)   # RS^2/BODYM as util_gpu.F:46-48 passes it
    out = {"m": c.m, "x": c.x, "v": c.v, "h2": h2, "dtr": stepr,
           "lmax": lmax, "nnbmax": nnbmax, "m_flag": m_flag}
    for lib in ("reg_avx", "reg_sse"):
        R = ref.RegLib(lib, c.n + 10)
        R.send(c.m, c.x, c.v)
        a, j, p, l = R.regf(h2, stepr, c.x, c.v, lmax, nnbmax, m_flag)
        R.close()
        tag = lib.split("_")[1]
        out.update({f"acc_{tag}": a, f"jrk_{tag}": j, f"pot_{tag}": p,
                    f"list_{tag}": l})
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)
    print(name, "mean nnb", out["list_avx"][:, 0].mean())


def pot_case(name, c):
    p = ref.pot("pot_avx", c.m, c.x)
    q = ref.pot("pot_sse", c.m, c.x)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), m=c.m, x=c.x,
                        pot_avx=p, pot_sse=q)
    print(name)


def main():
    assert ref.available("reg_avx"), "build oracle/_ref first (make -C oracle ref)"
    regf_case("plummer256", ics.plummer(256, seed=21), 48, 201, 200)
    c = ics.with_binaries(ics.plummer(640, seed=22, m=ics.kroupa(640, seed=23)),
                          0.1, seed=24)
    regf_case("kroupa_bin_mflag", c, 40, 301, 300, m_flag=1)
    regf_case("overflow128", ics.plummer(128, seed=25), 40, 21, 20)
    pot_case("pot300", ics.plummer(300, seed=26, m=ics.kroupa(300, seed=27)))


if __name__ == "__main__":
    main()
