"""Statistics and spectra files in the reference's formats.

`write_stats` follows write_stats (stats.f90:254-441) and `write_spect`
follows write_spect / recvwrspec (spectra.f90:58-182) for a single process
(np = 1): Fortran unformatted sequential records with 4-byte length markers.

stats file:
  record 1: t, Re, alp, bet, mpgx (real*8), nband, iter (int*4), 88 x int*4 0
  record 2: Ngal(4, 0:nband+1) (int*4)
  record 3: yu(0:ny), dthetai, dthdyu(0:ny)
  record 4: yv(0:ny-1), dthetai, dthdyv(0:ny-1)
  record 5: istat
  records 6-18: Um U2m (u rows 0..ny), Vm V2m (v rows 0..ny-1), Wm W2m,
                Pm P2m (p rows 1..ny-1), UVm UWm, VWm wxm wx2m
The reference appends tile-conditional records for its wall bands
(conddown/condup, stats.f90:109-252); they are not produced here.

spectrum file (one per quantity, spU_/spV_/spW_/spUV_/spP_):
  records 1-5 as above with N(4, 0:nband+1) in record 2, then
  buffSP(0:N1/2, 1:N2/2+1, rows): the accumulated |f|^2 folded onto
  non-negative z wavenumbers.  recvwrspec assigns rank 0's columns in column
  order (a negative z mode overwrites its positive partner, k -> N2+2-k) and
  doubles i = 1..N1/2, k = 1..N2/2.
"""
import struct

import numpy as np

# profile order of the statistics file (and of lh_stats_get)
PROFILES = ("Um", "U2m", "Vm", "V2m", "Wm", "W2m", "Pm", "P2m", "UVm", "UWm", "VWm", "wxm", "wx2m")
GRID_OF = dict(Um="u", U2m="u", Vm="v", V2m="v", Wm="u", W2m="u", Pm="p", P2m="p",
               UVm="u", UWm="u", VWm="v", wxm="v", wx2m="v")
SPECTRA = ("U", "V", "W", "UV", "P")
SPEC_GRID = dict(U="u", V="v", W="u", UV="u", P="p")


def _rows(grid, ny):
    """Reference rows (first, count) of a profile on grid u, v or p."""
    return {"u": (0, ny + 1), "v": (0, ny), "p": (1, ny - 1)}[grid]


def _rec(f, payload):
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def _header(f, hdr, nb_array, istat):
    _rec(f, struct.pack("<5d", hdr["t"], hdr["Re"], hdr["alp"], hdr["bet"], hdr["mpgx"])
         + struct.pack("<2i", int(hdr["nband"]), int(hdr["iter"])) + bytes(4 * 88))
    _rec(f, np.asarray(nb_array, dtype="<i4").tobytes(order="F"))
    dti = float(hdr.get("dthetai", 1.0))
    _rec(f, np.asarray(hdr["yu"], "<f8").tobytes() + struct.pack("<d", dti)
         + np.asarray(hdr["dthdyu"], "<f8").tobytes())
    _rec(f, np.asarray(hdr["yv"], "<f8").tobytes() + struct.pack("<d", dti)
         + np.asarray(hdr["dthdyv"], "<f8").tobytes())
    _rec(f, struct.pack("<i", int(istat)))


def write_stats(path, prof, istat, hdr):
    """prof: (13, ny+1) array (lh_stats_get layout); hdr: dict with t, Re,
    alp, bet, mpgx, nband, iter, Ngal (4 x nband+2 ints), yu, yv, dthdyu,
    dthdyv and optionally dthetai (1.0)."""
    prof = np.asarray(prof, dtype=np.float64)
    ny = prof.shape[1] - 1
    with open(path, "wb") as f:
        _header(f, hdr, hdr["Ngal"], istat)
        for q, name in enumerate(PROFILES):
            j0, n = _rows(GRID_OF[name], ny)
            _rec(f, prof[q, j0:j0 + n].astype("<f8").tobytes())


def fold_spectrum(sp, N1, N2):
    """recvwrspec's buffSP(0:N1/2, 1:N2/2+1, rows) from sp (ncol, rows), for
    one process.  Returns an array of shape (rows, N2/2+1, N1/2+1) (C order
    of the Fortran array)."""
    nxh = N1 // 2 + 1
    ncol, nrows = sp.shape
    out = np.zeros((nrows, N2 // 2 + 1, nxh))
    for c in range(ncol):
        i, k = c % nxh, c // nxh + 1
        kk = N2 + 2 - k if k > N2 // 2 else k
        out[:, kk - 1, i] = sp[c, :]
    out[:, : N2 // 2, 1: N1 // 2 + 1] *= 2.0
    return out


def write_spect(path, sp, istat, hdr, N1, N2):
    """sp: (ncol, rows) accumulated spectrum (lh_spectrum_get layout); hdr as
    for write_stats with N (4 x nband+2 ints) in place of Ngal."""
    buf = fold_spectrum(np.asarray(sp, dtype=np.float64), N1, N2)
    with open(path, "wb") as f:
        _header(f, hdr, hdr["N"], istat)
        _rec(f, buf.astype("<f8").tobytes())


def _records(buf):
    off = 0
    while off < len(buf):
        (ln,) = struct.unpack_from("<i", buf, off)
        rec = buf[off + 4: off + 4 + ln]
        (ln2,) = struct.unpack_from("<i", buf, off + 4 + ln)
        if ln2 != ln:
            raise ValueError("corrupt record marker at byte %d" % off)
        yield rec
        off += ln + 8


def _read_header(recs):
    t, Re, alp, bet, mpgx = struct.unpack_from("<5d", recs[0], 0)
    nband, it = struct.unpack_from("<2i", recs[0], 40)
    nb = np.frombuffer(recs[1], dtype="<i4").reshape(nband + 2, 4).T.copy()
    r3 = np.frombuffer(recs[2], dtype="<f8")
    r4 = np.frombuffer(recs[3], dtype="<f8")
    nu, nv = (len(r3) - 1) // 2, (len(r4) - 1) // 2
    (istat,) = struct.unpack_from("<i", recs[4], 0)
    return dict(t=t, Re=Re, alp=alp, bet=bet, mpgx=mpgx, nband=nband, iter=it, N=nb,
                yu=r3[:nu].copy(), dthetai=float(r3[nu]), dthdyu=r3[nu + 1:].copy(),
                yv=r4[:nv].copy(), dthdyv=r4[nv + 1:].copy(), istat=istat)


def read_stats(path):
    """Header dict plus one array per profile name."""
    with open(path, "rb") as f:
        recs = list(_records(f.read()))
    out = _read_header(recs)
    for q, name in enumerate(PROFILES):
        out[name] = np.frombuffer(recs[5 + q], dtype="<f8").copy()
    return out


def read_spect(path, N1, N2):
    """Header dict plus buffSP as an array (rows, N2/2+1, N1/2+1)."""
    with open(path, "rb") as f:
        recs = list(_records(f.read()))
    out = _read_header(recs)
    a = np.frombuffer(recs[5], dtype="<f8")
    out["sp"] = a.reshape(-1, N2 // 2 + 1, N1 // 2 + 1).copy()
    return out


def write_sl_stats(path, maps, istat, hdr):
    """write_sl_stats (sl_stats2.f90:163-217): header (t, Re, alp, bet, mpgx,
    nband, iter, 88 zeros), N(4, 0:nband+1), istat_sl, then the four maps
    bslip_u1_M, bslip_du1dy_M, bslip_u3_M, bslip_du3dy_M, each (N1+2) x N2 in
    Fortran order, the two walls added as the reference's writer adds its
    first and last ranks.  maps: (2, 4, N2, N1+2) (Solver.slip_stats)."""
    m = np.asarray(maps, dtype=np.float64)
    with open(path, "wb") as f:
        _rec(f, struct.pack("<5d", hdr["t"], hdr["Re"], hdr["alp"], hdr["bet"], hdr["mpgx"])
             + struct.pack("<2i", int(hdr["nband"]), int(hdr["iter"])) + bytes(4 * 88))
        _rec(f, np.asarray(hdr["N"], dtype="<i4").tobytes(order="F"))
        _rec(f, struct.pack("<i", int(istat)))
        for q in range(4):
            _rec(f, (m[0, q] + m[1, q]).astype("<f8").tobytes())


def read_sl_stats(path):
    with open(path, "rb") as f:
        recs = list(_records(f.read()))
    t, Re, alp, bet, mpgx = struct.unpack_from("<5d", recs[0], 0)
    nband, it = struct.unpack_from("<2i", recs[0], 40)
    N = np.frombuffer(recs[1], dtype="<i4").reshape(nband + 2, 4).T.copy()
    (istat,) = struct.unpack_from("<i", recs[2], 0)
    nz, ig = int(N[1, 1]), int(N[0, 1]) + 2
    maps = [np.frombuffer(r, dtype="<f8").reshape(nz, ig).copy() for r in recs[3:7]]
    return dict(t=t, Re=Re, alp=alp, bet=bet, mpgx=mpgx, nband=nband, iter=it, N=N, istat=istat, maps=maps)
