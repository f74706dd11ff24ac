"""NBODY6++GPU's conf.3 snapshot format: reader, writer and file name.

A conf.3 file is a Fortran unformatted sequential file (src/Main/output.F:
510-535) of two records, each framed by 4-byte little-endian byte counts:

    record 1: NTOT, MODEL, NRUN, NK                       (INTEGER*4)
    record 2: AS(1..NK)                                   (REAL*4)
              BODYS(1..NTOT), RHOS(1..NTOT), XNS(1..NTOT) (REAL*4)
              XS(3,NTOT), VS(3,NTOT)                      (REAL*4)
              PHI(1..NTOT)                                (REAL*4, = -PHIDBL)
              NAME(1..NTOT)                               (INTEGER*4)

AS holds (output.F:488-507) TTOT, NPAIRS, RBAR, ZMBAR, RTIDE, TIDAL(4),
RDENS(1..3), TTOT/TCR0, TSCALE, VSTAR, RC, NC, VC, RHOM, CMAX, RSCALE,
RSMIN, DMIN1.  The file is named 'conf.3_' + string_left(TTOT, DELTAT)
(src/Main/string_left.f).  Records of 2 GiB or more (Fortran subrecords) are
not supported: NTOT < 48.8 million.
"""
from __future__ import annotations

import math
import struct

import numpy as np

AS_FIELDS = ("TTOT", "NPAIRS", "RBAR", "ZMBAR", "RTIDE", "TIDAL4", "RDENS1",
             "RDENS2", "RDENS3", "TTOT_TCR0", "TSCALE", "VSTAR", "RC", "NC",
             "VC", "RHOM", "CMAX", "RSCALE", "RSMIN", "DMIN1")


def string_left(time: float, dt: float) -> str:
    """TIME in F20.ip, ip the decimals DT needs, left-justified
    (string_left.f); for ip = 0 without the trailing decimal point."""
    ip = 0
    dtmp = float(dt)
    while dtmp - math.trunc(dtmp) != 0.0:
        ip += 1
        dtmp = dtmp * 10.0
        if ip > 40:
            raise ValueError(f"string_left: {dt!r} has no finite decimal "
                             "expansion within 40 digits")
    s = "%.*f" % (ip, float(time))
    if len(s) > 20:
        s = "*" * 20            # the F20.ip field overflows
    return s


def conf3_name(time: float, dt: float) -> str:
    return "conf.3_" + string_left(time, dt)


def _record(f, payload: bytes) -> None:
    n = len(payload)
    if n >= 2 ** 31:
        raise ValueError("conf.3 record of 2 GiB or more (Fortran subrecords)")
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def write_conf3(path: str, as_: np.ndarray, body, rho, xn, x, v, phi, name,
                model: int = 1, nrun: int = 1) -> None:
    """Write a conf.3 file; every per-body array has NTOT entries (x, v:
    (NTOT, 3)), phi is PHI = -PHIDBL; floats are stored as REAL*4."""
    ntot = len(body)
    as_ = np.asarray(as_, dtype="<f4")
    rec1 = struct.pack("<4i", ntot, int(model), int(nrun), len(as_))
    parts = [as_.tobytes()]
    for a in (body, rho, xn):
        parts.append(np.asarray(a, dtype="<f4").reshape(ntot).tobytes())
    for a in (x, v):
        parts.append(np.ascontiguousarray(np.asarray(a, dtype="<f4")
                                          .reshape(ntot, 3)).tobytes())
    parts.append(np.asarray(phi, dtype="<f4").reshape(ntot).tobytes())
    parts.append(np.asarray(name, dtype="<i4").reshape(ntot).tobytes())
    with open(path, "wb") as f:
        _record(f, rec1)
        _record(f, b"".join(parts))


def _read_record(f) -> bytes:
    head = f.read(4)
    if len(head) != 4:
        raise ValueError("conf.3: truncated record header")
    (n,) = struct.unpack("<i", head)
    if n < 0:
        raise ValueError("conf.3: Fortran subrecords are not supported")
    data = f.read(n)
    tail = f.read(4)
    if len(data) != n or len(tail) != 4 or struct.unpack("<i", tail)[0] != n:
        raise ValueError("conf.3: record length markers do not match")
    return data


def read_conf3(path: str) -> dict:
    """Read a conf.3 file into a dict (keys NTOT, MODEL, NRUN, NK, AS, AS
    fields by name, body, rho, xn, x, v, phi, name)."""
    with open(path, "rb") as f:
        r1 = _read_record(f)
        if len(r1) != 16:
            raise ValueError("conf.3: first record is not 4 integers")
        ntot, model, nrun, nk = struct.unpack("<4i", r1)
        r2 = _read_record(f)
    want = 4 * (nk + 11 * ntot)
    if len(r2) != want:
        raise ValueError(f"conf.3: second record has {len(r2)} bytes, "
                         f"expected {want} for NTOT={ntot}, NK={nk}")
    fl = np.frombuffer(r2, dtype="<f4", count=nk + 10 * ntot)
    o = nk
    out = {"NTOT": ntot, "MODEL": model, "NRUN": nrun, "NK": nk,
           "AS": fl[:nk].copy()}
    for k, a in zip(AS_FIELDS, fl[:nk]):
        out[k] = float(a)
    for k in ("body", "rho", "xn"):
        out[k] = fl[o:o + ntot].copy()
        o += ntot
    out["x"] = fl[o:o + 3 * ntot].reshape(ntot, 3).copy()
    o += 3 * ntot
    out["v"] = fl[o:o + 3 * ntot].reshape(ntot, 3).copy()
    o += 3 * ntot
    out["phi"] = fl[o:o + ntot].copy()
    out["name"] = np.frombuffer(r2, dtype="<i4", offset=4 * (nk + 10 * ntot),
                                count=ntot).copy()
    return out
