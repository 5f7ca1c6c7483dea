"""Reader/writer for LittleHarsh field snapshots.

The on-disk format is the one `record_out` writes and `read_in` reads
(FOU3D.f90:2856-3073, start.f90:2444-2882): Fortran unformatted sequential
records with 4-byte length markers,

  record 1: t, Re, alp, bet, mpgx (real*8), nband, iter (int*4), 88 x int*4 0
  record 2: N(4, 0:nband+1) (int*4)
  record 3: y(:), dthetavi, dthdy(:)  (real*8)
  then one record per plane j: j, grid, nx, nz (int*4), y(j) (real*8),
            plane(nx, nz) (real*8, Fortran order)

A plane holds the spectral coefficients of that y-row: plane(2i+1, k) and
plane(2i+2, k) are the real and imaginary parts of x-mode i and z-index k
(k = 1..nz, the negative z modes in the upper half, `u_to_buff`
FOU3D.f90:3134-3163), which is exactly the solver's column order.
"""
import struct

import numpy as np


def _records(buf):
    off = 0
    n = len(buf)
    while off < n:
        (ln,) = struct.unpack_from("<i", buf, off)
        rec = buf[off + 4: off + 4 + ln]
        (ln2,) = struct.unpack_from("<i", buf, off + 4 + ln)
        if ln2 != ln:
            raise ValueError("corrupt record marker at byte %d" % off)
        yield rec
        off += ln + 8


def read_field(path):
    """Return dict(t, Re, alp, bet, mpgx, nband, iter, N, y, dthetavi, dthdy,
    rows: {j: (grid, yj, plane[nx, nz])}, field: complex array (ncol, nrows),
    jlo, jhi)."""
    with open(path, "rb") as f:
        buf = f.read()
    recs = list(_records(buf))
    h = recs[0]
    t, Re, alp, bet, mpgx = struct.unpack_from("<5d", h, 0)
    nband, it = struct.unpack_from("<2i", h, 40)
    N = np.frombuffer(recs[1], dtype="<i4").reshape(nband + 2, 4).T.copy()
    r3 = np.frombuffer(recs[2], dtype="<f8")
    ny3 = (len(r3) - 1) // 2
    y = r3[:ny3].copy()
    dthetavi = float(r3[ny3])
    dthdy = r3[ny3 + 1:].copy()
    rows = {}
    for rec in recs[3:]:
        j, grid, nx, nz = struct.unpack_from("<4i", rec, 0)
        (yj,) = struct.unpack_from("<d", rec, 16)
        plane = np.frombuffer(rec, dtype="<f8", offset=24).reshape(nz, nx).T.copy()
        rows[j] = (grid, yj, plane)
    js = sorted(rows)
    nx, nz = rows[js[0]][2].shape
    nxh = nx // 2
    field = np.zeros((nxh * nz, len(js)), dtype=np.complex128)
    for q, j in enumerate(js):
        pl = rows[j][2]
        c = pl[0::2, :] + 1j * pl[1::2, :]       # (nxh, nz)
        field[:, q] = c.T.reshape(-1)            # column = k*nxh + i
    return dict(t=t, Re=Re, alp=alp, bet=bet, mpgx=mpgx, nband=nband, iter=it,
                N=N, y=y, dthetavi=dthetavi, dthdy=dthdy, rows=rows,
                field=field, jlo=js[0], jhi=js[-1])


def write_field(path, field, jlo, grid, y, dthdy, t=0.0, Re=0.0, alp=1.0, bet=2.0,
                mpgx=0.0, iter=0, N=None, yplane=None, dthetavi=1.0):
    """Write `field` (ncol, nrows) for rows jlo.. in the record_out layout."""
    ncol, nrows = field.shape
    if N is None:
        raise ValueError("N (4 x nband+2 int array) is required")
    N = np.asarray(N, dtype="<i4")
    nband = N.shape[1] - 2
    nz = None
    # infer nxh, nz from N(1,1), N(2,1)
    nx = int(N[0, 1]) + 2
    nz = int(N[1, 1])
    nxh = nx // 2
    assert nxh * nz == ncol
    yp = y if yplane is None else yplane

    def rec(payload):
        return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))

    out = bytearray()
    out += rec(struct.pack("<5d", t, Re, alp, bet, mpgx) + struct.pack("<2i", nband, iter)
               + bytes(4 * 88))
    out += rec(N.T.astype("<i4").tobytes())
    out += rec(np.asarray(y, "<f8").tobytes() + struct.pack("<d", dthetavi)
               + np.asarray(dthdy, "<f8").tobytes())
    for q in range(nrows):
        j = jlo + q
        c = field[:, q].reshape(nz, nxh).T        # (nxh, nz)
        pl = np.zeros((nx, nz))
        pl[0::2, :] = c.real
        pl[1::2, :] = c.imag
        out += rec(struct.pack("<4i", j, grid, nx, nz) + struct.pack("<d", float(yp[j - (0 if yplane is None else 0)]))
                   + pl.T.astype("<f8").tobytes())
    with open(path, "wb") as f:
        f.write(bytes(out))


# ---------------------------------------------------------------------------
# History file (hist_NNNNxNNNNxNNNN.dat): start.f90:238-253 opens it and
# writes two header records, (t, Re, alp, bet, mpgx, nband) and N(4,
# 0:nband+1); record_out appends one record per output (FOU3D.f90:3063-3068):
#   flag_ctpress = 0: flag, iter, t, dtv, dtc, dt, err, Qx, QxT, mpgx, dgx, Umax, Uslip
#   flag_ctpress = 1: flag, iter, t, dtv, dtc, dt, err, Qx, mpgx, Umax, Uslip
HIST_KEYS = {0: ("t", "dtv", "dtc", "dt", "err", "Qx", "QxT", "mpgx", "dgx", "Umax", "Uslip"),
             1: ("t", "dtv", "dtc", "dt", "err", "Qx", "mpgx", "Umax", "Uslip")}


def read_hist(path):
    """dict(t, Re, alp, bet, mpgx, nband, N, records=[dict(flag, iter, ...)])."""
    with open(path, "rb") as f:
        recs = list(_records(f.read()))
    t, Re, alp, bet, mpgx = struct.unpack_from("<5d", recs[0], 0)
    (nband,) = struct.unpack_from("<i", recs[0], 40)
    N = np.frombuffer(recs[1], dtype="<i4").reshape(nband + 2, 4).T.copy()
    out = []
    for r in recs[2:]:
        flag, it = struct.unpack_from("<2i", r, 0)
        vals = struct.unpack_from("<%dd" % ((len(r) - 8) // 8), r, 8)
        d = dict(zip(HIST_KEYS[flag], vals))
        d.update(flag=flag, iter=it)
        out.append(d)
    return dict(t=t, Re=Re, alp=alp, bet=bet, mpgx=mpgx, nband=nband, N=N, records=out)


def hist_record(flag_ctpress, h):
    """One history record's bytes from a dict with the HIST_KEYS values and
    iter (Solver.history() gives them)."""
    payload = struct.pack("<2i", int(flag_ctpress), int(h["iter"])) + \
        struct.pack("<%dd" % len(HIST_KEYS[flag_ctpress]), *[float(h[k]) for k in HIST_KEYS[flag_ctpress]])
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def write_hist(path, t, Re, alp, bet, mpgx, N, records, flag_ctpress):
    """A history file with its two header records and `records` (dicts)."""
    N = np.asarray(N, dtype="<i4")
    nband = N.shape[1] - 2

    def rec(p):
        return struct.pack("<i", len(p)) + p + struct.pack("<i", len(p))

    with open(path, "wb") as f:
        f.write(rec(struct.pack("<5d", t, Re, alp, bet, mpgx) + struct.pack("<i", nband)))
        f.write(rec(N.T.astype("<i4").tobytes()))
        for h in records:
            f.write(hist_record(flag_ctpress, h))
