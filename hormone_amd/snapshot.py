"""Reader for the reference's binary snapshot format.

Layout (src/main/output.f90:697-795, io.F90:345-363; stream access,
little-endian, legacy int4(-1) record markers): [-1][tn i4][time f8][-1]
[-1] d v1 v2 v3 e [-1] then optional groups; for gravswitch >= 2 [-1]
grvphi [-1]; for gravswitch 3 [-1] grvpsi cgrav_old [-1]; for compswitch
>= 2 [-1] spn spc(1:spn,...) names [-1]; for mag_on [-1] b1 b2 b3 phi [-1].
Each field is nx*ny*nz float64, x fastest (the gravity fields over the
gravity grid gis..gie, which is the hydro grid for the Cartesian runs here,
checksetup.f90:171-179).

The grid file (gridfile.bin, output.f90:432-452) is one sequential
unformatted record: int4 byte count, x1 xi1 dx1 dxi1 over is-2..ie+2, the
same for x2 and x3, int4 byte count.
"""
from __future__ import annotations

import numpy as np


def read_bin(path: str, shape_zyx, mag_on: bool, eostype: int = 0,
             gravswitch: int = 0, compswitch: int = 0):
    raw = open(path, "rb").read()
    n = int(np.prod(shape_zyx))
    off = 0

    def i4():
        nonlocal off
        v = int(np.frombuffer(raw, "<i4", 1, off)[0])
        off += 4
        return v

    def f8(cnt):
        nonlocal off
        v = np.frombuffer(raw, "<f8", cnt, off).copy()
        off += 8 * cnt
        return v

    out = {}
    assert i4() == -1
    out["tn"] = i4()
    out["time"] = float(f8(1)[0])
    assert i4() == -1
    assert i4() == -1
    for nm in ("d", "v1", "v2", "v3", "e"):
        out[nm] = f8(n).reshape(shape_zyx)
    assert i4() == -1
    if eostype >= 1:
        assert i4() == -1
        out["T"] = f8(n).reshape(shape_zyx)
        assert i4() == -1
    if compswitch == 1:
        raise NotImplementedError("compswitch 1 (mudata) groups are not read here")
    if gravswitch >= 2:  # readbin.f90:39-43
        assert i4() == -1
        out["grvphi"] = f8(n).reshape(shape_zyx)
        assert i4() == -1
    if gravswitch == 3:  # readbin.f90:45-50
        assert i4() == -1
        out["grvpsi"] = f8(n).reshape(shape_zyx)
        out["cgrav_old"] = float(f8(1)[0])
        assert i4() == -1
    if compswitch >= 2:
        # [-1] spn, spc(1:spn, cells) with the species index fastest,
        # species names (character(10) each) [-1]  (output.f90:756-761)
        assert i4() == -1
        spn = i4()
        spc = f8(spn * n).reshape(n, spn)
        for q in range(spn):
            out[f"spc{q + 1}"] = spc[:, q].copy().reshape(shape_zyx)
        out["species"] = [raw[off + 10 * q:off + 10 * q + 10].decode().strip()
                          for q in range(spn)]
        off += 10 * spn
        assert i4() == -1
    if mag_on:
        assert i4() == -1
        for nm in ("b1", "b2", "b3", "phi"):
            out[nm] = f8(n).reshape(shape_zyx)
        assert i4() == -1
    assert off == len(raw), (off, len(raw))
    return out


def write_bin(snap: dict, mag_on: bool, compswitch: int = 0, gravswitch: int = 0) -> bytes:
    """The bytes write_bin would write for this state (output.f90:697-795,
    io.F90:345-363): stream access, little-endian, int4(-1) record markers.
    snap: tn, time, d v1 v2 v3 e (nz, ny, nx) arrays, grvphi when
    gravswitch >= 2, grvpsi and cgrav_old when gravswitch is 3, spc1..spcN
    and species names when compswitch >= 2, b1 b2 b3 phi when mag_on."""
    m1 = np.int32(-1).tobytes()

    def f8(a):
        return np.ascontiguousarray(a, dtype="<f8").tobytes()

    out = [m1, np.int32(snap["tn"]).tobytes(), f8(snap["time"]), m1, m1]
    out += [f8(snap[k]) for k in ("d", "v1", "v2", "v3", "e")]
    out.append(m1)
    if gravswitch >= 2:
        out += [m1, f8(snap["grvphi"]), m1]
    if gravswitch == 3:
        out += [m1, f8(snap["grvpsi"]), f8(snap["cgrav_old"]), m1]
    if compswitch >= 2:
        names = snap["species"]
        spn = len(names)
        spc = np.stack([np.asarray(snap[f"spc{q + 1}"]).reshape(-1) for q in range(spn)], 1)
        out += [m1, np.int32(spn).tobytes(), f8(spc)]
        out += [nm.encode().ljust(10)[:10] for nm in names]
        out.append(m1)
    if mag_on:
        out += [m1] + [f8(snap[k]) for k in ("b1", "b2", "b3", "phi")] + [m1]
    return b"".join(out)


def write_grid_bin(g) -> bytes:
    """The bytes of gridfile.bin (write_grid_bin, output.f90:432-452) for
    grid g: x, xi, dx, dxi of each axis over is-2..ie+2 in one sequential
    unformatted record."""
    body = b"".join(np.ascontiguousarray(getattr(g.ax[a], k), dtype="<f8").tobytes()
                    for a in range(3) for k in ("x", "xi", "dx", "dxi"))
    m = np.int32(len(body)).tobytes()
    return m + body + m


def read_grid_bin(raw: bytes, n) -> dict:
    """gridfile.bin of a grid with n = (nx, ny, nz) interior cells ->
    {"x1": ..., "xi1": ..., "dx1": ..., "dxi1": ..., "x2": ...}, each over
    the padded index range."""
    ln = int(np.frombuffer(raw, "<i4", 1, 0)[0])
    want = 8 * 4 * sum(m + 4 for m in n)
    assert ln == want and len(raw) == ln + 8, (ln, want, len(raw))
    assert int(np.frombuffer(raw, "<i4", 1, 4 + ln)[0]) == ln
    out, off = {}, 4
    for a in range(3):
        for k in ("x", "xi", "dx", "dxi"):
            out[f"{k}{a + 1}"] = np.frombuffer(raw, "<f8", n[a] + 4, off).copy()
            off += 8 * (n[a] + 4)
    return out
