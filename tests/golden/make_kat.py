#!/usr/bin/env python3
"""Known-answer vectors of the reference's Riemann solver, EOS and limiter
routines, computed by the reference's own compiled code.

Builds tests/golden/kat_reference.f90 against the object files of the
reference build (oracle/build_ref.py must have run: oracle/_ref/obj), feeds
it seeded inputs chosen to reach every branch (sub- and supersonic faces,
B_n = 0, B = 0, identical states, a contact at rest, near-degenerate Alfven
states, strong jumps; signed zeros; limiter extrema and flat data on
non-uniform grids; non-positive EOS inputs) and stores inputs and outputs in
tests/golden/kat_reference.npz.  Runs only where the reference tree and
amdflang are present; the fixture it writes is data.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")
FC = os.environ.get("FC", "amdflang")
FLAGS = ["-O3", "-fopenmp", "-ffp-contract=off"]
OUT = os.path.join(HERE, "kat_reference.npz")


def get_cf(d, cs, b1, b2, b3):
    asq = cs * cs
    bb1, bb2, bb3 = b1 * b1 / d, b2 * b2 / d, b3 * b3 / d
    bbsq = bb1 + bb2 + bb3
    return np.sqrt(0.5 * (asq + bbsq + np.sqrt(np.maximum((asq + bbsq) ** 2 - 4 * asq * bb1, 0))))


def riemann_rows(rng, n, kind):
    """Left/right states of a face in hlldflux's argument order."""
    gam = rng.choice([5.0 / 3.0, 1.4, 2.0], n)
    dl, dr = rng.lognormal(0, 1, n), rng.lognormal(0, 1, n)
    pl, pr = rng.lognormal(0, 1, n), rng.lognormal(0, 1, n)
    v = rng.uniform(-1.5, 1.5, (6, n))   # v1l v1r v2l v2r v3l v3r
    b = rng.uniform(-1.0, 1.0, (6, n))   # b1l b1r b2l b2r b3l b3r
    phi = rng.uniform(-0.5, 0.5, (2, n))
    erad = np.zeros((2, n))
    if kind == "bn0":
        b[0:2] = 0.0
    elif kind == "b0":
        b[:] = 0.0
        phi[:] = 0.0
    elif kind == "same":
        dr, pr = dl.copy(), pl.copy()
        for q in range(0, 6, 2):
            v[q + 1], b[q + 1] = v[q], b[q]
        phi[1] = phi[0]
    elif kind == "collide":  # mirror states: the contact is at rest
        dr, pr = dl.copy(), pl.copy()
        v[1] = -v[0]
        v[3], v[5] = v[2], v[4]
        b[1], b[3], b[5] = b[0], -b[2], -b[4]
        phi[1] = phi[0]
    elif kind == "alfven":  # transverse B nearly zero, B_n finite
        b[2:6] *= rng.choice([0.0, 1e-12, 1e-8], (4, n))
    elif kind == "jump":
        dr = dl * 10.0 ** rng.uniform(-6, 6, n)
        pr = pl * 10.0 ** rng.uniform(-6, 6, n)
    elif kind == "erad":
        erad = rng.lognormal(0, 1, (2, n))
    csl, csr = np.sqrt(gam * pl / dl), np.sqrt(gam * pr / dr)
    cfl = get_cf(dl, csl, b[0], b[2], b[4])
    cfr = get_cf(dr, csr, b[1], b[3], b[5])
    if kind == "super":  # both states supersonic, either way
        s = rng.choice([-1.0, 1.0], n) * rng.uniform(2.0, 8.0, n)
        v[0] = s * np.maximum(cfl, cfr) + v[0] * 0.1
        v[1] = s * np.maximum(cfl, cfr) + v[1] * 0.1
    bl2 = b[0] ** 2 + b[2] ** 2 + b[4] ** 2
    br2 = b[1] ** 2 + b[3] ** 2 + b[5] ** 2
    el = pl / (gam - 1) + 0.5 * dl * (v[0] ** 2 + v[2] ** 2 + v[4] ** 2) + 0.5 * bl2
    er = pr / (gam - 1) + 0.5 * dr * (v[1] ** 2 + v[3] ** 2 + v[5] ** 2) + 0.5 * br2
    ch = np.maximum(cfl, cfr) + rng.uniform(0.0, 2.0, n)
    return np.stack([cfl, cfr, v[0], v[1], v[2], v[3], v[4], v[5], dl, dr, el, er,
                     pl + 0.5 * bl2, pr + 0.5 * br2, b[0], b[1], b[2], b[3], b[4], b[5],
                     phi[0], phi[1], ch, erad[0], erad[1]], axis=1)


def inputs(seed=20261203):
    rng = np.random.default_rng(seed)
    kinds = ["gen"] * 6 + ["bn0", "b0", "same", "collide", "alfven", "jump", "super", "erad"]
    hlld = np.concatenate([riemann_rows(rng, 200, k) for k in kinds])

    n = 1500
    b = rng.uniform(-2, 2, (3, n)) * rng.choice([0.0, 1.0], (3, n), p=[0.2, 0.8])
    cf = np.stack([rng.lognormal(0, 1.5, n), rng.lognormal(0, 1, n), b[0], b[1], b[2]], axis=1)

    d = rng.lognormal(0, 2, n)
    ei = rng.lognormal(0, 2, n)
    bad = rng.random(n) < 0.08
    d[bad & (rng.random(n) < 0.5)] *= -1.0
    ei[bad] = np.where(rng.random(bad.sum()) < 0.5, 0.0, -ei[bad])
    erad = np.where(rng.random(n) < 0.5, 0.0, rng.lognormal(0, 1, n))
    eos = np.stack([d, ei, rng.choice([5.0 / 3.0, 1.4, 2.0, 1.1], n), erad], axis=1)

    v = rng.uniform(-3, 3, (3, n))
    bb = rng.uniform(-2, 2, (3, n))
    dd = rng.lognormal(0, 1, n)
    etot = (0.5 * dd * (v ** 2).sum(0) + 0.5 * (bb ** 2).sum(0)) * rng.uniform(0.6, 3.0, n)
    eint = np.stack([etot, dd, v[0], v[1], v[2], bb[0], bb[1], bb[2],
                     rng.choice([0.0, 1.0], n)], axis=1)

    # limiters on non-uniform grids: x1 < xi1 < x2 < xi2 < x3 (cell centres
    # and faces), a quarter of them uniform
    m = 2000
    h = rng.lognormal(0, 0.3, (3, m))
    uni = rng.random(m) < 0.25
    h[:, uni] = 0.01
    x2 = rng.uniform(-1, 1, m)
    xi1 = x2 - 0.5 * h[1]
    xi2 = x2 + 0.5 * h[1]
    x1 = xi1 - 0.5 * h[0]
    x3 = xi2 + 0.5 * h[2]
    u = rng.normal(0, 1, (3, m))
    flat = rng.random(m) < 0.1
    u[1, flat] = u[0, flat]
    u[2, flat] = u[0, flat]
    ext = rng.random(m) < 0.2  # local extremum at the centre
    u[1, ext] = np.maximum(u[0, ext], u[2, ext]) + rng.uniform(0, 1, ext.sum())
    zer = rng.random(m) < 0.05
    u[:, zer] = rng.choice([0.0, -0.0], (3, zer.sum()))
    mc = np.stack([u[0], u[1], u[2], x1, x2, x3, xi1, xi2], axis=1)
    idx = 1.0 / h[:2]
    minmod = np.stack([u[0], u[1], u[2], idx[0], idx[1]], axis=1)
    return {"hlld": hlld, "cf": cf, "eos": eos, "eint": eint, "mc": mc, "minmod": minmod}


def build(tmp):
    objs = sorted(os.path.join(OBJ, f) for f in os.listdir(OBJ)
                  if f.endswith(".o") and f != "hormone.o")
    exe = os.path.join(tmp, "kat_reference")
    subprocess.run([FC, *FLAGS, "-I", OBJ, "-J", tmp, os.path.join(HERE, "kat_reference.f90"),
                    *objs, "-o", exe], check=True)
    return exe


def run(exe, tmp, kind, a):
    fin, fout = os.path.join(tmp, kind + ".in"), os.path.join(tmp, kind + ".out")
    a = np.ascontiguousarray(a, dtype="<f8")
    with open(fin, "wb") as f:
        np.array([a.shape[0], a.shape[1]], dtype="<i4").tofile(f)
        a.tofile(f)
    subprocess.run([exe, kind, fin, fout], check=True)
    with open(fout, "rb") as f:
        n, nout = np.fromfile(f, dtype="<i4", count=2)
        o = np.fromfile(f, dtype="<f8", count=n * nout).reshape(n, nout)
    return o


def main():
    if not os.path.isdir(OBJ):
        sys.exit("oracle/_ref/obj missing: run oracle/build_ref.py first")
    ins = inputs()
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        exe = build(tmp)
        for kind, a in ins.items():
            out[kind + "_in"] = a
            out[kind + "_out"] = run(exe, tmp, kind, a)
    np.savez_compressed(OUT, **out)
    print(OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
