"""The restricted (triad-list) nonlinear path, ops_in_planes + nonlinInter
(FOU3D.f90:1724-2030), and its list files (nonlinRead, start.f90:2983-3069).

CPU: the reference's list files parse to the sizes it records and are
subsets of the threshold-0 lists the generator here builds; the oracle's
list path with full lists equals its plain FFT path on band-limited fields
(the truncated convolution both compute); with thinned lists it equals an
independent numpy evaluation of the same sums.  GPU: the device list path
against the oracle's, with generated lists and with the reference's own
files."""
import os

import numpy as np
import pytest

from helpers import oracle_for, relerr, small_cfg
from rnl_multiblock_nofft_amd import initial, triads

REFDIR = "/root/reference/3e-9"


def band_limited(cfg, grid, seed):
    """Fields with modes |k| < N/4 only: every product stays inside the
    list band |l| <= N1/2-1, |k| <= N2/2-1."""
    return initial.perturbed_poiseuille(cfg, grid, seed, amp=0.1, kmax=cfg["N1"] // 4 - 1)


# ---------------------------------------------------------------- numpy restatement


def full_spectrum(col, N1, N2):
    """Stored columns (i = 0..N1/2, stored z index) -> dict-free dense array
    F[l + N1/2, k + N2/2] for l = -N1/2..N1/2, k = -N2/2..N2/2-1, with the
    negative x modes the conjugates of the stored ones (a real field)."""
    nxh = N1 // 2 + 1
    C = col.reshape(N2, nxh)
    F = np.zeros((N1 + 1, N2), complex)
    for kk in range(N2):
        k = kk if kk < N2 // 2 else kk - N2
        for i in range(nxh):
            F[i + N1 // 2, k + N2 // 2] = C[kk, i]
            if i > 0:
                km = -k if -k < N2 // 2 else -k - N2
                F[-i + N1 // 2, km + N2 // 2] = np.conj(C[kk, i])
    return F


def list_sums(lst, A, P, N1, N2):
    """sum over rows (la, ka, lp, kp) of A(la, ka) P(lp, kp) into (la+lp, ka+kp)
    and of the z-mirror A(la, -ka) P(lp, -kp) into (la+lp, -(ka+kp)) unless
    ka+kp = 0; returns the stored-column array of the targets."""
    nxh = N1 // 2 + 1
    out = np.zeros(nxh * N2, complex)
    h1, h2 = N1 // 2, N2 // 2
    for la, ka, lp, kp in lst.astype(int):
        # p = 0 (FOU3D.f90:2004-2010, pn = -1): the z-mirrored pair into
        # -(ka+kp); p = 1: the row itself, skipped when ka + kp = 0
        for sgn in ((-1,) if ka + kp == 0 else (-1, 1)):
            a = A[la + h1, sgn * ka + h2]
            b = P[lp + h1, sgn * kp + h2]
            kt = sgn * (ka + kp)
            out[la + lp + nxh * (kt if kt >= 0 else kt + N2)] += a * b
    return out


def numpy_list_nonlinear(cfg, grid, f, files, mu, mv):
    """Nu1, Nu2, Nu3 of the list path with numpy (FOU3D.f90:1724-1987 and the
    y-derivative of the fluxes, FOU3D.f90:266-283)."""
    N1, N2, ny = cfg["N1"], cfg["N2"], cfg["ny"]
    nxh = N1 // 2 + 1
    ncol = nxh * N2
    yu, yv, du, dv = grid["yu"], grid["yv"], grid["dthdyu"], grid["dthdyv"]
    u1, u2, u3 = f["u1"], f["u2"], f["u3"]
    u2i = np.zeros_like(u1)
    for j in range(1, ny):
        u2i[:, j] = ((yu[j] - yv[j - 1]) * u2[:, j] + (yv[j] - yu[j]) * u2[:, j - 1]) / (yv[j] - yv[j - 1])
    u1i = np.zeros_like(u2)
    u3i = np.zeros_like(u2)
    for j in range(ny):
        w1, w2, d = yv[j] - yu[j], yu[j + 1] - yv[j], yu[j + 1] - yu[j]
        u1i[:, j] = (w1 * u1[:, j + 1] + w2 * u1[:, j]) / d
        u3i[:, j] = (w1 * u3[:, j + 1] + w2 * u3[:, j]) / d
    kx = cfg["alp"] * np.arange(nxh)
    kz = np.array([cfg["bet"] * k if k < N2 // 2 else -cfg["bet"] * (N2 - k) for k in range(N2)])
    KX = np.tile(kx, N2)
    KZ = np.repeat(kz, nxh)
    live = np.repeat(np.arange(N2) != N2 // 2, nxh)
    Nu1 = np.zeros((ncol, ny + 1), complex)
    Nu3 = np.zeros((ncol, ny + 1), complex)
    Nu2 = np.zeros((ncol, ny), complex)
    vv = np.zeros((ncol, ny + 1), complex)
    uv = np.zeros((ncol, ny), complex)
    wv = np.zeros((ncol, ny), complex)

    def zeroth(a, b, kneg):
        # sum over i, k of a(i, k) b(-i, -k), FOU3D.f90:1763-1778
        A, B = full_spectrum(a, N1, N2), full_spectrum(b, N1, N2)
        s = 0
        for i in range(-N1 // 2, N1 // 2 + 1):
            for k in range(-kneg, N2 // 2):
                kb = -k if -k < N2 // 2 else -k - N2
                s += A[i + N1 // 2, k + N2 // 2] * B[-i + N1 // 2, kb + N2 // 2]
        return s

    def lists_of(q):
        return files[q] if q is not None and q >= 0 and files[q] is not None else None

    for j in range(1, ny):
        a1, a2, a3 = u1[:, j], u2i[:, j], u3[:, j]
        m1, m2, m3 = a1[0].real, a2[0].real, a3[0].real
        uu = 2 * m1 * a1
        ww = 2 * m3 * a3
        vvj = 2 * m2 * a2
        vvj[0] = 0
        uw = m1 * a3 + m3 * a1
        wu = uw.copy()
        vvj[0] += zeroth(a2, a2, N2 // 2)
        L = lists_of(mu[j - 1])
        if L is not None:
            F1, F2, F3 = (full_spectrum(x, N1, N2) for x in (a1, a2, a3))
            uu = uu + list_sums(L[1], F1, F1, N1, N2)
            uw = uw + list_sums(L[3], F3, F1, N1, N2)
            vvj = vvj + list_sums(L[5], F2, F2, N1, N2)
            wu = wu + list_sums(L[7], F1, F3, N1, N2)
            ww = ww + list_sums(L[9], F3, F3, N1, N2)
        Nu1[:, j] = np.where(live, 1j * KX * uu + 1j * KZ * uw, 0)
        Nu3[:, j] = np.where(live, 1j * KX * wu + 1j * KZ * ww, 0)
        vv[:, j] = vvj
    for j in range(1, ny - 1):
        a1, a2, a3 = u1i[:, j], u2[:, j], u3i[:, j]
        m1, m2, m3 = a1[0].real, a2[0].real, a3[0].real
        vu = m1 * a2 + m2 * a1
        uvj = vu.copy()
        uvj[0] = 0
        vw = m2 * a3 + m3 * a2
        wvj = vw.copy()
        wvj[0] = 0
        uvj[0] += zeroth(a1, a2, N2 // 2 - 1)
        wvj[0] += zeroth(a3, a2, N2 // 2 - 1)
        L = lists_of(mv[j - 1])
        if L is not None:
            F1, F2, F3 = (full_spectrum(x, N1, N2) for x in (a1, a2, a3))
            uvj = uvj + list_sums(L[2], F2, F1, N1, N2)
            vu = vu + list_sums(L[4], F1, F2, N1, N2)
            vw = vw + list_sums(L[6], F3, F2, N1, N2)
            wvj = wvj + list_sums(L[8], F2, F3, N1, N2)
        Nu2[:, j] = np.where(live, 1j * KX * vu + 1j * KZ * vw, 0)
        uv[:, j] = uvj
        wv[:, j] = wvj
    for j in range(1, ny):
        Nu1[:, j] += (uv[:, j] - uv[:, j - 1]) * du[j]
        Nu3[:, j] += (wv[:, j] - wv[:, j - 1]) * du[j]
    for j in range(1, ny - 1):
        Nu2[:, j] += (vv[:, j + 1] - vv[:, j]) * dv[j]
    return Nu1[:, 1:ny], Nu2[:, 1:ny - 1], Nu3[:, 1:ny]


# ---------------------------------------------------------------- CPU


@pytest.mark.skipif(not os.path.isdir(REFDIR), reason="reference list files not present")
def test_reference_list_files():
    """The reference's own files: header (nx, nz, j), the sizes fullsize.txt
    records, and every row one of the threshold-0 generator's."""
    sizes = {}
    with open(os.path.join(REFDIR, "fullsize.txt")) as f:
        for line in f:
            a = line.split()
            if len(a) == 2:
                sizes[int(a[0])] = int(a[1])
    full = triads.full_lists(32, 32)
    fullsets = {t: set(map(tuple, full[t].tolist())) for t in range(1, 10)}
    present = [q for q in range(1, 77) if os.path.exists(os.path.join(REFDIR, "full_y%02d.dat" % q))]
    assert present, "no list files"
    for q in present:
        nx, nz, j, lists = triads.read_list(os.path.join(REFDIR, "full_y%02d.dat" % q))
        assert (nx, nz, j) == (15, 15, q)
        assert sum(len(lists[t]) for t in lists) == sizes[q]
        for t in (1, 5, 9):
            assert set(map(tuple, lists[t].tolist())) <= fullsets[t], (q, t)


def test_list_round_trip(tmp_path):
    lists = triads.thinned(triads.full_lists(8, 8), 0.5, seed=3)
    p = str(tmp_path / "full_y01.dat")
    triads.write_list(p, 3, 3, 1, lists)
    nx, nz, j, back = triads.read_list(p)
    assert (nx, nz, j) == (3, 3, 1)
    for t in range(1, 10):
        assert np.array_equal(back[t], lists[t])


def test_plane_file_map():
    """Planes above the centre read the mirrored file; u planes there use
    index j - 1 (FOU3D.f90:1811-1815, start.f90:3025-3030)."""
    u, v = triads.plane_files(153)
    assert len(u) == 152 and len(v) == 151
    assert u[:3] == [1, 2, 3] and v[:3] == [1, 2, 3]
    assert u[75] == 76 and u[76] == 76 and u[151] == 1      # j = 76, 77, 152
    assert v[75] == 76 and v[76] == 75 and v[150] == 1      # j = 76, 77, 151
    assert max(u + v) == 76


@pytest.mark.parametrize("N,ny", [(8, 13), (16, 17)])
def test_full_lists_equal_plain_path(N, ny):
    """Threshold-0 lists: the list path and the plain FFT path compute the
    same truncated convolution on band-limited fields."""
    cfg = small_cfg(N=N, ny=ny)
    a, b = oracle_for(cfg), oracle_for(cfg)
    f = band_limited(cfg, a.grid(), 7)
    for k, v in f.items():
        a.set(k, v)
        b.set(k, v)
    full = triads.full_lists(N, N)
    b.set_triads([full], [0] * (ny - 1), [0] * (ny - 2))
    a.phase("nonlinear", 1)
    b.phase("nonlinear", 1)
    for nm in ("Nu1", "Nu2", "Nu3"):
        e = relerr(b.get(nm), a.get(nm))
        assert e < 1e-13, (nm, e)
    assert a.scalars()["dtc"] == pytest.approx(b.scalars()["dtc"], rel=1e-14)


def test_thinned_lists_match_numpy():
    """Thinned lists (passive/active pairs no longer symmetric), planes with
    different files and planes without: the oracle against numpy."""
    N, ny = 8, 9
    cfg = small_cfg(N=N, ny=ny)
    o = oracle_for(cfg)
    rng = np.random.default_rng(4)
    f = {}
    for name in ("u1", "u2", "u3", "p"):
        lo, hi = o.rows(name)
        f[name] = 0.1 * (rng.standard_normal((o.ncol, hi - lo + 1)) + 1j * rng.standard_normal((o.ncol, hi - lo + 1)))
        o.set(name, f[name])
    full = triads.full_lists(N, N)
    files = [triads.thinned(full, 0.6, seed=1), triads.thinned(full, 0.3, seed=2)]
    mu = [0, 1, -1, 0, 1, 0, 1, 0][: ny - 1]
    mv = [1, 0, 0, -1, 1, 0, 1][: ny - 2]
    o.set_triads(files, mu, mv)
    o.phase("nonlinear", 2)
    n1, n2, n3 = numpy_list_nonlinear(cfg, o.grid(), f, files, mu, mv)
    for nm, ref in (("Nu1", n1), ("Nu2", n2), ("Nu3", n3)):
        e = relerr(o.get(nm), ref)
        assert e < 1e-13, (nm, e)


def test_list_outside_band_rejected():
    cfg = small_cfg(N=8, ny=9)
    o = oracle_for(cfg)
    bad = {t: np.array([[3, 0, 1, 0]], np.int8) for t in range(1, 10)}   # la + lp = 4 > nx = 3
    with pytest.raises(ValueError):
        o.set_triads([bad], [0] * 8, [0] * 7)


def _golden_lists():
    import importlib.util
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "make_triad_golden.py")
    spec = importlib.util.spec_from_file_location("make_triad_golden", p)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m.load()


@pytest.mark.skipif(not os.path.isdir(REFDIR), reason="reference list files not present")
def test_golden_lists_rebuild_reference_files():
    g = _golden_lists()
    for q, lists in g.items():
        ref = triads.read_list(os.path.join(REFDIR, "full_y%02d.dat" % q))[3]
        for t in range(1, 10):
            assert len(lists[t]) == len(ref[t])
            assert set(map(tuple, lists[t].tolist())) == set(map(tuple, ref[t].tolist()))


# ---------------------------------------------------------------- GPU


def _gpu_list_pair(cfg, files, mu, mv, seed, fields=None):
    from rnl_multiblock_nofft_amd import Solver
    s = Solver([cfg])
    o = oracle_for(cfg)
    f = fields if fields is not None else band_limited(cfg, s.grid(0), seed)
    for k, v in f.items():
        s.upload(0, k, v)
        o.set(k, v)
    s.init_flowrate()
    o.init_flowrate()
    s.set_triads(0, files, mu, mv)
    o.set_triads(files, mu, mv)
    return s, o


def _compare(s, o, tol, what):
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        print(what, nm, "%.2e" % e, flush=True)
        assert e <= tol, (what, nm, e)


@pytest.mark.gpu
def test_gpu_full_lists_match_oracle():
    """Threshold-0 lists on every plane: one substep's nonlinear terms, one
    step and three steps against the oracle's list path."""
    cfg = small_cfg(N=16, ny=17)
    full = triads.full_lists(16, 16)
    from rnl_multiblock_nofft_amd import _lib
    s, o = _gpu_list_pair(cfg, [full], [0] * 16, [0] * 15, seed=8)
    s.phase(_lib.PHASE_NONLINEAR, 1)
    s.phase(_lib.PHASE_RHS, 1)
    o.phase("rhs0", 1)
    o.phase("nonlinear", 1)
    for nm in ("Nu1", "Nu2", "Nu3"):
        e = relerr(s.download(0, nm), o.get(nm))
        print("nonlinear", nm, "%.2e" % e, flush=True)
        assert e <= 1e-13, (nm, e)
    s2, o2 = _gpu_list_pair(cfg, [full], [0] * 16, [0] * 15, seed=8)
    s2.step(1)
    o2.step(1)
    _compare(s2, o2, 1e-12, "1 step")
    s2.step(2)
    o2.step(2)
    _compare(s2, o2, 1e-11, "3 steps")


@pytest.mark.gpu
def test_gpu_thinned_lists_match_oracle():
    """Asymmetric (thinned) lists, two files and planes without lists."""
    cfg = small_cfg(N=8, ny=13)
    full = triads.full_lists(8, 8)
    files = [triads.thinned(full, 0.6, seed=1), triads.thinned(full, 0.3, seed=2), None]
    mu = [0, 1, -1, 0, 1, 2, 1, 0, 0, 1, 0, 1]
    mv = [1, 0, 0, -1, 1, 0, 1, 2, 0, 1, 1]
    s, o = _gpu_list_pair(cfg, files, mu, mv, seed=3)
    s.step(2)
    o.step(2)
    _compare(s, o, 1e-11, "thinned")


@pytest.mark.gpu
def test_gpu_reference_lists_32():
    """The reference's own lists (files 1, 6, 9 of its 32x32x153 case, from
    the committed fixture) on the planes that read them; the other planes
    keep the mean-mode terms only.  One step against the oracle."""
    from rnl_multiblock_nofft_amd import configs
    cfg = dict(configs.smooth_32()[0])
    g = _golden_lists()
    ny = cfg["ny"]
    u, v = triads.plane_files(ny)
    files = [g.get(q) for q in range(1, max(u + v) + 1)]
    mu = [q - 1 if g.get(q) is not None else -1 for q in u]
    mv = [q - 1 if g.get(q) is not None else -1 for q in v]
    assert sum(m >= 0 for m in mu) >= 6 and sum(m >= 0 for m in mv) >= 6
    s, o = _gpu_list_pair(cfg, files, mu, mv, seed=21)
    s.step(1)
    o.step(1)
    _compare(s, o, 1e-12, "reference lists")


@pytest.mark.gpu
def test_gpu_lists_need_fixed_dt():
    from rnl_multiblock_nofft_amd import Solver, _lib
    cfg = small_cfg(N=8, ny=9, adaptive_dt=1, CFL=0.5)
    s = Solver([cfg])
    with pytest.raises(_lib.LhError):
        s.set_triads(0, [triads.full_lists(8, 8)], [0] * 8, [0] * 7)
