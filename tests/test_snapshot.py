"""conf.3 snapshots (nbody6ppgpu_amd/snapshot.py) against a Fortran runtime
and the reference's own file-name formatter.

* A Fortran program of ours with output.F:526-534's I/O list
  (tests/fortran/conf3_io.f90, amdflang) reads what the Python writer
  wrote, and the Python reader reads what it wrote: the same values.
* string_left (the file-name time stamp) equals the reference's
  src/Main/string_left.f, compiled from its source under oracle/_ref.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from nbody6ppgpu_amd import snapshot

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F_PROG = os.path.join(ROOT, "tests", "fortran", "_build", "conf3_io")
REF_SL = os.path.join(ROOT, "oracle", "_ref", "libref_string_left.so")

need_fortran = pytest.mark.skipif(not os.path.exists(F_PROG),
                                  reason="tests/fortran not built")


def pattern(ntot):
    j = np.arange(1, ntot + 1, dtype=np.float32)
    return dict(as_=0.5 * np.arange(1, 21, dtype=np.float32),
                body=np.float32(1.0) / j, rho=2.0 * j, xn=j + 0.25,
                x=np.stack([j, -j, 0.5 * j], axis=1),
                v=np.stack([0.125 * j, np.full_like(j, 3.0), -0.75 * j], axis=1),
                phi=-j, name=1000 + np.arange(1, ntot + 1))


@need_fortran
def test_fortran_reads_the_python_writer(tmp_path):
    p = pattern(777)
    f = str(tmp_path / "conf.3_1.5")
    snapshot.write_conf3(f, p["as_"], p["body"], p["rho"], p["xn"], p["x"],
                         p["v"], p["phi"], p["name"], model=3, nrun=7)
    out = subprocess.run([F_PROG, "read", f], check=True, capture_output=True,
                         text=True).stdout.split()
    assert [int(w) for w in out[:4]] == [777, 3, 7, 20]
    sums = [float(w) for w in out[4:11]]
    want = [np.sum(p[k].astype(np.float64)) for k in
            ("as_", "body", "rho", "xn", "x", "v", "phi")]
    assert np.allclose(sums, want, rtol=1e-12)
    assert int(out[11]) == int(np.sum(p["name"]))
    last = [float(w) for w in out[12:19]]
    assert last == list(p["x"][-1].astype(float)) + \
        list(p["v"][-1].astype(float)) + [float(p["body"][-1])]


@need_fortran
def test_python_reads_the_fortran_writer(tmp_path):
    f = str(tmp_path / "conf.3_0")
    subprocess.run([F_PROG, "write", f, "513"], check=True)
    d = snapshot.read_conf3(f)
    p = pattern(513)
    assert (d["NTOT"], d["MODEL"], d["NRUN"], d["NK"]) == (513, 3, 7, 20)
    assert np.array_equal(d["AS"], p["as_"])
    for k in ("body", "rho", "xn", "x", "v", "phi"):
        assert np.array_equal(d[k], np.asarray(p[k], np.float32)), k
    assert np.array_equal(d["name"], p["name"])
    assert d["TTOT"] == 0.5 and d["RSCALE"] == 9.0


def test_round_trip_and_damaged_files(tmp_path):
    p = pattern(64)
    f = str(tmp_path / "c")
    snapshot.write_conf3(f, p["as_"], p["body"], p["rho"], p["xn"], p["x"],
                         p["v"], p["phi"], p["name"])
    d = snapshot.read_conf3(f)
    assert np.array_equal(d["x"], np.asarray(p["x"], np.float32))
    raw = open(f, "rb").read()
    open(f, "wb").write(raw[:-10])
    with pytest.raises(ValueError):
        snapshot.read_conf3(f)


@pytest.mark.skipif(not os.path.exists(REF_SL),
                    reason="reference string_left not built (oracle/_ref)")
def test_file_names_follow_the_reference_string_left():
    lib = ctypes.CDLL(REF_SL)
    fn = lib.string_left_
    fn.restype = None
    fn.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                   ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]

    def ref(t, dt):
        buf = ctypes.create_string_buffer(b" " * 20, 20)
        fn(buf, ctypes.byref(ctypes.c_double(t)),
           ctypes.byref(ctypes.c_double(dt)), 20)
        return buf.raw.decode().rstrip()

    for dt in (1.0, 0.5, 0.25, 0.125, 0.0625, 0.1, 2.0, 10.0):
        for k in range(0, 40):
            t = k * dt
            assert snapshot.string_left(t, dt) == ref(t, dt), (t, dt)
    assert snapshot.conf3_name(0.125, 0.125) == "conf.3_0.125"
