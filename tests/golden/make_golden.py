"""Writes tests/golden/*.npz from the reference's OWN code and data:

* FFTPACK (cft_buff.f, rft_buff.f) and the tridiagonal solvers
  (tridLU_3D.f90) compiled unmodified from the reference sources into
  oracle/_ref/libref.so by oracle/build_ref.sh, run on seeded inputs (one
  channel, and a banded channel whose short pencils live in the wall bands);
* the y-grid, flow rate and dtv of the reference's 32x32x153 case, read from
  its input fixture (input/u1_0032x0032x0153_t00000.dat, record 3) and its
  history file (output/hist_0032x0032x0153.dat, iter 0/1 records).

Run in the container that has /root/reference:  python tests/golden/make_golden.py
The outputs are data (inputs and expected outputs), not reference source.
"""
import ctypes
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference"

from oracle import pyoracle as po  # noqa: E402
from rnl_multiblock_nofft_amd import fieldio  # noqa: E402

R = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref.so"))


def ip(v):
    return ctypes.byref(ctypes.c_int(v))


def dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ref_plane(P, N1, Ng1, Ng2, direction):
    """four_to_phys_du (direction +1) / phys_to_four_du (-1) composed from the
    reference's cft/rft exactly as FOU3D.f90:766-784, 2253-2272 calls them."""
    P = P.copy()
    igal = Ng1 + 2
    wc = np.zeros(10 * Ng2 + 19)
    R.cfti_(ip(Ng2), dp(wc))
    wr = np.zeros(2 * Ng1 + 18)
    R.rfti_(ip(Ng1), dp(wr))
    if direction > 0:
        P.reshape(Ng2, igal)[Ng2 // 2, :] = 0
        R.cft_(dp(P), ip(igal), ip(2), ip((N1 + 2) // 2), ip(1), dp(wc))
        R.rft_(dp(P), ip(igal), ip(Ng2), ip(1), dp(wr))
    else:
        R.rft_(dp(P), ip(igal), ip(Ng2), ip(-1), dp(wr))
        R.cft_(dp(P), ip(igal), ip(2), ip((N1 + 2) // 2), ip(-1), dp(wc))
        P.reshape(Ng2, igal)[Ng2 // 2, :] = 0
    return P


def make_fft():
    out = {}
    for N in (8, 32):
        Ng = 3 * N // 2
        rng = np.random.default_rng(100 + N)
        P = rng.standard_normal(Ng * (Ng + 2))
        Q = rng.standard_normal(Ng * (Ng + 2))
        out["f2p_in_%d" % N] = P
        out["f2p_out_%d" % N] = ref_plane(P, N, Ng, Ng, +1)
        out["p2f_in_%d" % N] = Q
        out["p2f_out_%d" % N] = ref_plane(Q, N, Ng, Ng, -1)
    np.savez_compressed(os.path.join(HERE, "fftpack_planes.npz"), **out)


def make_tridlu():
    out = {}
    for tag, cfg in (("poly", dict(N1=4, N2=4, ny=17, Re=3250.0, grid_kind=0, dyq=10.0, ppp=5, dt=2e-3)),
                     ("tanh", dict(N1=4, N2=4, ny=21, Re=4200.0, grid_kind=1, gamma=2.0, dt=1e-3))):
        o = po.Oracle(**cfg)
        g = o.grid()
        ny = cfg["ny"]
        ncol = (cfg["N1"] // 2 + 1) * cfg["N2"]
        yu, yv = g["yu"], g["yv"]
        gw1 = -(yv[0] - yu[0]) / (yv[0] - yu[1])
        gw2 = (yu[ny] - yv[ny - 1]) / (yv[ny - 1] - yu[ny - 1])
        gw = np.array([gw1, gw2])
        dyu2i = np.ascontiguousarray(g["dyu2i"])
        dyv2i = np.ascontiguousarray(g["dyv2i"])
        R.ref_setup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, ctypes.c_double] + [ctypes.c_void_p] * 7
        R.ref_setup(cfg["N1"], cfg["N2"], ny, 1.0, 2.0, cfg["Re"], cfg["dt"],
                    *[a.ctypes.data for a in (g["yu"], g["yv"], g["dthdyu"], g["dthdyv"], dyu2i, dyv2i, gw)])
        rng = np.random.default_rng(7)
        out[tag + "_yu"] = g["yu"]
        out[tag + "_yv"] = g["yv"]
        out[tag + "_dthdyu"] = g["dthdyu"]
        out[tag + "_dthdyv"] = g["dthdyv"]
        for kRK in (1, 2, 3):
            ru = rng.standard_normal((ncol, ny + 1)) + 1j * rng.standard_normal((ncol, ny + 1))
            rw = rng.standard_normal((ncol, ny + 1)) + 1j * rng.standard_normal((ncol, ny + 1))
            rv = rng.standard_normal((ncol, ny)) + 1j * rng.standard_normal((ncol, ny))
            u = np.zeros_like(ru)
            w = np.zeros_like(rw)
            v = np.zeros_like(rv)
            R.ref_lusol_uw(ctypes.c_int(kRK), ctypes.c_void_p(u.ctypes.data), ctypes.c_void_p(ru.ctypes.data),
                           ctypes.c_void_p(w.ctypes.data), ctypes.c_void_p(rw.ctypes.data))
            R.ref_lusol_v(ctypes.c_int(kRK), ctypes.c_void_p(v.ctypes.data), ctypes.c_void_p(rv.ctypes.data))
            x0 = rng.standard_normal(ny + 1)
            x0c = x0.copy()
            R.ref_lusol0(ctypes.c_int(kRK), ctypes.c_void_p(x0c.ctypes.data))
            out["%s_k%d_ru" % (tag, kRK)] = ru
            out["%s_k%d_rw" % (tag, kRK)] = rw
            out["%s_k%d_rv" % (tag, kRK)] = rv
            out["%s_k%d_u" % (tag, kRK)] = u
            out["%s_k%d_w" % (tag, kRK)] = w
            out["%s_k%d_v" % (tag, kRK)] = v
            out["%s_k%d_r0" % (tag, kRK)] = x0
            out["%s_k%d_x0" % (tag, kRK)] = x0c
        rp = rng.standard_normal((ncol, ny - 1)) + 1j * rng.standard_normal((ncol, ny - 1))
        xp = rp.copy()
        R.ref_lusol_p(ctypes.c_void_p(xp.ctypes.data))
        out[tag + "_rp"] = rp
        out[tag + "_xp"] = xp
    np.savez_compressed(os.path.join(HERE, "tridlu.npz"), **out)


def band_limits(ny, nb):
    """N(3,q), N(4,q) of ygrid after the band split (start.f90:990-1060;
    dny = 0, N(3,nband) = ny - 2) and the row limits jlim(1:2, grid, band)
    of proc_lims_columns (start.f90:1677-1711), grid = vgrid 1, ugrid 2,
    pgrid 3 (declaration.f90:15-17)."""
    N3 = [0, nb, ny - 2 - nb, ny - 2]
    N4 = [0, N3[1] + 1, N3[2], N3[3] + 1]
    lims = np.zeros((3, 3, 2), np.int32)  # [band][grid][lo/hi] (C order of jlim(2,3,3))
    lims[0][1] = (N4[0], N4[1] + 1)
    lims[1][1] = (N4[0], N4[3] + 1)
    lims[2][1] = (N4[2], N4[3] + 1)
    lims[0][0] = (N3[0], N3[1] + 1)
    lims[1][0] = (N3[0], N3[3] + 1)
    lims[2][0] = (N3[2], N3[3] + 1)
    lims[0][2] = (N4[0] + 1, N4[1])
    lims[1][2] = (N4[0] + 1, N4[3])
    lims[2][2] = (N4[2] + 1, N4[3])
    return N3, N4, lims


def make_tridlu_bands():
    """The reference's LUsolU_W / LUsolV / LUsolP on a banded channel: the
    short pencils of the bottom and top bands over their own rows with the
    band-edge weights of start.f90:539-544, the long pencils over the whole
    channel."""
    out = {}
    cfg = dict(N1=8, N2=8, ny=33, Re=3250.0, grid_kind=0, dyq=10.0, ppp=5, dt=2e-3)
    N1m, N2m = 4, 4
    B = po.BandOracle(N1m, N2m, band_h=0.25, **cfg)
    nb = B.nb
    ny = cfg["ny"]
    g = B.band(1).grid()
    yu, yv = g["yu"], g["yv"]
    N3, N4, lims = band_limits(ny, nb)
    gw = np.zeros((3, 2))
    gw[0, 0] = -(yv[N3[0]] - yu[N4[0]]) / (yv[N3[0]] - yu[N4[0] + 1])
    gw[0, 1] = (yu[N4[1] + 1] - yv[N3[1] + 1]) / (yv[N3[1] + 1] - yu[N4[1]])
    gw[1, 0] = -(yv[N3[0]] - yu[N4[0]]) / (yv[N3[0]] - yu[N4[0] + 1])
    gw[1, 1] = (yu[N4[3] + 1] - yv[N3[3] + 1]) / (yv[N3[3] + 1] - yu[N4[3]])
    gw[2, 0] = -(yv[N3[2]] - yu[N4[2]]) / (yv[N3[2]] - yu[N4[2] + 1])
    gw[2, 1] = (yu[N4[3] + 1] - yv[N3[3] + 1]) / (yv[N3[3] + 1] - yu[N4[3]])
    # columns: the middle band's long pencils (fine-grid k index), the wall
    # bands' short pencils, in the oracle's band column order
    ci, ck, ncb = [], [], []
    for q in range(3):
        b = B.band(q)
        nxh, N2 = b.cfg["N1"] // 2 + 1, b.cfg["N2"]
        own = B.owns(q)
        cols = [c for c in range(b.ncol) if own[c]]
        ncb.append(len(cols))
        for c in cols:
            i, k = c % nxh, c // nxh
            m = k if k < N2 // 2 else k - N2
            ci.append(i)
            ck.append((m % cfg["N2"]) + 1)
        out["b%d_cols" % q] = np.array(cols, np.int32)
    ci = np.array(ci, np.int32)
    ck = np.array(ck, np.int32)
    ncb = np.array(ncb, np.int32)
    dyu2i = np.ascontiguousarray(g["dyu2i"])
    dyv2i = np.ascontiguousarray(g["dyv2i"])
    gwf = np.ascontiguousarray(gw)  # C (band, side) = Fortran gw(side, band)
    R.ref_setup_bands.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double] + [ctypes.c_void_p] * 11
    R.ref_setup_bands(cfg["N1"], cfg["N2"], ny, 1.0, 2.0, cfg["Re"], cfg["dt"],
                      *[a.ctypes.data for a in (yu, yv, g["dthdyu"], g["dthdyv"], dyu2i, dyv2i, gwf,
                                                np.ascontiguousarray(lims), ncb, ci, ck)])
    out["nb"] = np.array(nb)
    out["lims"] = lims
    out["gw"] = gw
    rng = np.random.default_rng(11)
    for q in range(3):
        n = int(ncb[q])
        lu, lv, lp = lims[q][1], lims[q][0], lims[q][2]
        nu, nv, npr = lu[1] - lu[0] + 1, lv[1] - lv[0] + 1, lp[1] - lp[0] + 1
        for kRK in (1, 3):
            ru = rng.standard_normal((n, nu)) + 1j * rng.standard_normal((n, nu))
            rw = rng.standard_normal((n, nu)) + 1j * rng.standard_normal((n, nu))
            rv = rng.standard_normal((n, nv)) + 1j * rng.standard_normal((n, nv))
            u = np.zeros_like(ru)
            w = np.zeros_like(rw)
            v = np.zeros_like(rv)
            R.ref_lusol_uw_b(ctypes.c_int(q + 1), ctypes.c_int(kRK), ctypes.c_void_p(u.ctypes.data),
                             ctypes.c_void_p(ru.ctypes.data), ctypes.c_void_p(w.ctypes.data),
                             ctypes.c_void_p(rw.ctypes.data))
            R.ref_lusol_v_b(ctypes.c_int(q + 1), ctypes.c_int(kRK), ctypes.c_void_p(v.ctypes.data),
                            ctypes.c_void_p(rv.ctypes.data))
            for nm, a in (("ru", ru), ("rw", rw), ("rv", rv), ("u", u), ("w", w), ("v", v)):
                out["b%d_k%d_%s" % (q, kRK, nm)] = a
        rp = rng.standard_normal((n, npr)) + 1j * rng.standard_normal((n, npr))
        xp = rp.copy()
        R.ref_lusol_p_b(ctypes.c_int(q + 1), ctypes.c_void_p(xp.ctypes.data))
        out["b%d_rp" % q] = rp
        out["b%d_xp" % q] = xp
    np.savez_compressed(os.path.join(HERE, "tridlu_bands.npz"), **out)


def make_reference_case():
    d1 = fieldio.read_field(os.path.join(REF, "input", "u1_0032x0032x0153_t00000.dat"))
    d2 = fieldio.read_field(os.path.join(REF, "input", "u2_0032x0032x0153_t00000.dat"))
    buf = open(os.path.join(REF, "output", "hist_0032x0032x0153.dat"), "rb").read()
    recs = [r for r in fieldio._records(buf) if len(r) == 80]
    h0 = struct.unpack("<2i9d", recs[-3])
    h1 = struct.unpack("<2i9d", recs[-2])
    assert h0[1] == 0 and h1[1] == 1
    np.savez_compressed(os.path.join(HERE, "reference_case.npz"),
                        yu=d1["y"], dthdyu=d1["dthdy"], yv=d2["y"], dthdyv=d2["dthdy"],
                        N=d1["N"],
                        hist_iter0=np.array(h0[2:]), hist_iter1=np.array(h1[2:]))


if __name__ == "__main__":
    make_fft()
    make_tridlu()
    make_tridlu_bands()
    make_reference_case()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
