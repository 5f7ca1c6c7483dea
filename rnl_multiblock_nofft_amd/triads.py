"""Triad interaction lists of the reference's restricted nonlinear path.

The reference's active nonlinear term (`ops_in_planes`, FOU3D.f90:1724-1987)
does not transform to physical space: for every y plane it sums, in Fourier
space, the products of the mode pairs listed in a per-plane file
(`nonlinRead`, start.f90:2983-3069; `nonlinInter`, FOU3D.f90:1989-2030).

File layout (stream access, little endian):
  int32 nx, nz, j          nx = N1/2-1, nz = N2/2-1, j the plane index
  9 times: int32 type, length, then int8 list(length, 4) in Fortran order
           (all la, then all ka, then all lp, then all kp)
type 1..9 are the products uu, uv, uw, vu, vv, vw, wu, wv, ww, "the first
letter the passive field" (FOU3D.f90:1809).  A row (la, ka, lp, kp) adds
u_a(la, ka) u_p(lp, kp) to the target (la+lp, ka+kp), and its mirror in z
(-ka, -kp) to (la+lp, -(ka+kp)) unless ka+kp = 0; negative x indices are the
complex conjugates of the stored modes (iNeg, iLkup, kLkup, start.f90:
1139-1163).

Files exist for the lower half of the channel; plane j of the upper half
reads file Ngal3 + 1 - j (Ngal3 = ny - 2), and u planes above the centre
use the list of index j - 1 (FOU3D.f90:1811-1815, start.f90:3025-3030).
"""
import os
import struct

import numpy as np

# passive/active fields of each list type on its plane kind (FOU3D.f90:1821-1825, 1953-1956)
U_TYPES = (1, 3, 5, 7, 9)   # uu, uw, vv, wu, ww on u planes
V_TYPES = (2, 4, 6, 8)      # uv, vu, vw, wv on v planes


def read_list(path):
    """(nx, nz, j, {type: int8 array (length, 4) with columns la, ka, lp, kp})."""
    with open(path, "rb") as f:
        buf = f.read()
    nx, nz, j = struct.unpack_from("<3i", buf, 0)
    off = 12
    lists = {}
    for _ in range(9):
        t, n = struct.unpack_from("<2i", buf, off)
        off += 8
        a = np.frombuffer(buf, dtype=np.int8, count=4 * n, offset=off).reshape(4, n)
        off += 4 * n
        lists[t] = np.ascontiguousarray(a.T)
    if off != len(buf):
        raise ValueError("%s: %d trailing bytes" % (path, len(buf) - off))
    return nx, nz, j, lists


def write_list(path, nx, nz, j, lists):
    with open(path, "wb") as f:
        f.write(struct.pack("<3i", nx, nz, j))
        for t in range(1, 10):
            a = np.asarray(lists[t], dtype=np.int8).reshape(-1, 4)
            f.write(struct.pack("<2i", t, a.shape[0]))
            f.write(np.ascontiguousarray(a.T).tobytes())


def full_lists(N1, N2):
    """Every triad of the band |l| <= N1/2-1, |k| <= N2/2-1 with neither
    partner the mean mode and a target (lt >= 0, kt >= 0) other than the
    mean, one row per z-mirror pair -- the lists the reference's generator
    gives at threshold 0.  Targets that the x or z derivative removes are
    left out as in the reference's files: lt = 0 for uu, vu, wu (types 1, 4,
    7, differentiated in x) and kt = 0 for uw, vw, ww (3, 6, 9, in z)."""
    nx, nz = N1 // 2 - 1, N2 // 2 - 1
    modes = [(l, k) for l in range(-nx, nx + 1) for k in range(-nz, nz + 1) if (l, k) != (0, 0)]
    rows = []
    for la, ka in modes:
        for lp, kp in modes:
            lt, kt = la + lp, ka + kp
            if lt < 0 or lt > nx or abs(kt) > nz or (lt == 0 and kt == 0):
                continue
            # z-mirror pairs: (ka, kp) and (-ka, -kp) give the targets +kt and
            # -kt; keep the one with kt > 0 (for kt = 0 both rows are needed,
            # each is used mirrored only)
            if kt < 0:
                continue
            rows.append((la, ka, lp, kp))
    rows = np.array(rows, dtype=np.int8)
    lt = rows[:, 0].astype(int) + rows[:, 2]
    kt = rows[:, 1].astype(int) + rows[:, 3]
    out = {}
    for t in range(1, 10):
        keep = np.ones(len(rows), bool)
        if t in (1, 4, 7):
            keep &= lt != 0
        if t in (3, 6, 9):
            keep &= kt != 0
        out[t] = rows[keep]
    return out


def thinned(lists, fraction, seed):
    """A random subset of every list (tests of the restatement with
    asymmetric passive/active pairs)."""
    rng = np.random.default_rng(seed)
    return {t: a[rng.random(len(a)) < fraction] for t, a in lists.items()}


def plane_files(ny):
    """File index (1-based) of the list each plane uses: u planes j = 1..ny-1
    and v planes j = 1..ny-2, with the reference's mapping (jidx for the u
    planes above the centre, FOU3D.f90:1811-1815; the mirrored read,
    start.f90:3025-3030)."""
    g3 = ny - 2

    def mirror(jj):
        return g3 + 1 - jj if jj > (g3 + 1) // 2 else jj

    u = [mirror(j - 1 if j > (g3 + 1) // 2 else j) for j in range(1, ny)]
    v = [mirror(j) for j in range(1, ny - 1)]
    return u, v


def load_plane_lists(directory, ny, heading="full_y"):
    """Lists per plane from a directory of reference files; a plane whose
    file is absent gets empty lists.  Returns (files, u_map, v_map, present)
    with files[q] the lists of file q+1 (None when absent) and the maps
    giving each plane's file index (0-based)."""
    u, v = plane_files(ny)
    nfile = max(u + v)
    files, present = [], []
    for q in range(1, nfile + 1):
        p = os.path.join(directory, "%s%02d.dat" % (heading, q))
        if os.path.exists(p):
            files.append(read_list(p)[3])
            present.append(q)
        else:
            files.append(None)
    return files, [q - 1 for q in u], [q - 1 for q in v], present
