"""Statistics and spectra (stats.f90:1-107, spectra.f90:1-182).

CPU: the oracle's restatement is pinned against an independent numpy
evaluation of the same planes (numpy FFTs, not the oracle's), against closed
forms, and the Parseval form the device kernel uses is checked against the
oracle's physical-space sums.  GPU: the device accumulators against the
oracle, including non-Hermitian i = 0 columns, textured walls, ensembles and
the reset.  File writers: round trips of the reference's record layout."""
import numpy as np
import pytest

from helpers import init_fields, oracle_for, profile_errors, small_cfg
from rnl_multiblock_nofft_amd import statsio

NAMES = statsio.PROFILES


def rand_fields(o, rng, amp=1.0):
    out = {}
    for name in ("u1", "u2", "u3", "p"):
        lo, hi = o.rows(name)
        a = amp * (rng.standard_normal((o.ncol, hi - lo + 1)) + 1j * rng.standard_normal((o.ncol, hi - lo + 1)))
        o.set(name, a)
        out[name] = a
    return out


def to_phys(col, cfg, Ng1, Ng2):
    """The reference's four_to_phys on one row with numpy: stored column
    (i, k) at z row k (k < N2/2) or k - N2 + Ng2; unscaled inverse complex
    transform in z, real inverse in x (imaginary parts of modes 0 and Ng1/2
    dropped, as rft does).  Returns (Ng2, Ng1)."""
    N1, N2 = cfg["N1"], cfg["N2"]
    nxh = N1 // 2 + 1
    C = np.zeros((Ng2, Ng1 // 2 + 1), complex)
    cc = col.reshape(N2, nxh)
    kr = np.array([k if k < N2 // 2 else k - N2 + Ng2 for k in range(N2)])
    if Ng2 == N2:
        kr = np.arange(N2)
    C[kr, :nxh] = cc
    gz = np.fft.ifft(C, axis=0) * Ng2
    return np.fft.irfft(gz, n=Ng1, axis=1) * Ng1


def numpy_stats(cfg, grid, f):
    """stats.f90:1-107 evaluated with numpy on fields f (reference layout)."""
    ny, N1, N2 = cfg["ny"], cfg["N1"], cfg["N2"]
    G1, G2 = cfg["Ngal1"], cfg["Ngal2"]
    nxh = N1 // 2 + 1
    yu, yv, dthdyv = grid["yu"], grid["yv"], grid["dthdyv"]
    u1, u2, u3 = f["u1"], f["u2"], f["u3"]
    p = np.zeros((u1.shape[0], ny + 1), complex)
    p[:, 1:ny] = f["p"]
    # interp_v to the u rows, interp_u to the v rows (FOU3D.f90:427-474)
    u2i = np.zeros_like(u1)
    for j in range(1, ny):
        u2i[:, j] = ((yu[j] - yv[j - 1]) * u2[:, j] + (yv[j] - yu[j]) * u2[:, j - 1]) / (yv[j] - yv[j - 1])
    gwi1 = (yu[0] - yu[1]) / (yv[0] - yu[1])
    gwi2 = (yu[ny] - yu[ny - 1]) / (yv[ny - 1] - yu[ny - 1])
    u2i[:, 0] = gwi1 * (u2[:, 0] - u2i[:, 1]) + u2i[:, 1]
    u2i[:, ny] = gwi2 * (u2[:, ny - 1] - u2i[:, ny - 1]) + u2i[:, ny - 1]
    u3i = np.zeros_like(u2)
    for j in range(ny):
        u3i[:, j] = ((yv[j] - yu[j]) * u3[:, j + 1] + (yu[j + 1] - yv[j]) * u3[:, j]) / (yu[j + 1] - yu[j])
    kz = np.array([cfg["bet"] * k if k < N2 // 2 else -cfg["bet"] * (N2 - k) for k in range(N2)])
    kzc = np.repeat(kz, nxh)

    def spec_sum(a):
        return np.sum(a.real + a.imag)

    def spec_dot(a, b):
        return np.sum(a.real * b.real + a.imag * b.imag)

    prof = np.zeros((13, ny + 1))
    q = {n: i for i, n in enumerate(NAMES)}
    for j in range(ny + 1):
        if 1 <= j <= ny - 1:
            a, b, c = (to_phys(x[:, j], cfg, G1, G2) for x in (u1, u3, u2i))
            prof[q["Um"], j], prof[q["U2m"], j] = a.sum(), (a * a).sum()
            prof[q["Wm"], j], prof[q["W2m"], j] = b.sum(), (b * b).sum()
            prof[q["UVm"], j], prof[q["UWm"], j] = (a * c).sum(), (a * b).sum()
        else:
            prof[q["Um"], j], prof[q["U2m"], j] = spec_sum(u1[:, j]), spec_dot(u1[:, j], u1[:, j])
            prof[q["Wm"], j], prof[q["W2m"], j] = spec_sum(u3[:, j]), spec_dot(u3[:, j], u3[:, j])
            prof[q["UVm"], j], prof[q["UWm"], j] = spec_dot(u1[:, j], u2i[:, j]), spec_dot(u1[:, j], u3[:, j])
        if j <= ny - 1:
            if 1 <= j <= ny - 2:
                a, c = to_phys(u2[:, j], cfg, G1, G2), to_phys(u3i[:, j], cfg, G1, G2)
                prof[q["Vm"], j], prof[q["V2m"], j], prof[q["VWm"], j] = a.sum(), (a * a).sum(), (a * c).sum()
            else:
                prof[q["Vm"], j], prof[q["V2m"], j] = spec_sum(u2[:, j]), spec_dot(u2[:, j], u2[:, j])
                prof[q["VWm"], j] = spec_dot(u2[:, j], u3i[:, j])
            wx = -(1j * kzc) * u2[:, j] + (u3[:, j + 1] - u3[:, j]) * dthdyv[j]
            w = to_phys(wx, cfg, N1, N2)
            prof[q["wxm"], j], prof[q["wx2m"], j] = w.sum(), (w * w).sum()
        if 1 <= j <= ny - 1:
            prof[q["Pm"], j], prof[q["P2m"], j] = spec_sum(p[:, j]), spec_dot(p[:, j], p[:, j])
    sp = dict(U=np.abs(u1) ** 2, V=np.abs(u2) ** 2, W=np.abs(u3) ** 2,
              UV=u1.real * u2i.real + u1.imag * u2i.imag, P=np.abs(f["p"]) ** 2)
    return prof, sp


def parseval_row(cfg, a, b, ngrid):
    """The device kernel's evaluation of sum_grid(A B) from the stored
    columns of one row: weight 2 for i > 0 (and, on the N grid, i < N1/2);
    the i = 0 (N grid: also i = N1/2) column through its Hermitian part; on
    the Ngal grid the z index N2/2 with weight 1/2."""
    N1, N2 = cfg["N1"], cfg["N2"]
    nxh = N1 // 2 + 1
    A, B = a.reshape(N2, nxh), b.reshape(N2, nxh)
    partner = (N2 - np.arange(N2)) % N2
    s = 0.0
    proj = [0] + ([N1 // 2] if ngrid == "N" else [])
    for i in range(nxh):
        if i in proj:
            ha = 0.5 * (A[:, i] + np.conj(A[partner, i]))
            hb = 0.5 * (B[:, i] + np.conj(B[partner, i]))
            t = ha.real * hb.real + ha.imag * hb.imag
            if ngrid == "G":
                t[N2 // 2] = 0.5 * (A[N2 // 2, i].real * B[N2 // 2, i].real + A[N2 // 2, i].imag * B[N2 // 2, i].imag)
            s += t.sum()
        else:
            s += 2.0 * np.sum(A[:, i].real * B[:, i].real + A[:, i].imag * B[:, i].imag)
    n = cfg["Ngal1"] * cfg["Ngal2"] if ngrid == "G" else N1 * N2
    return n * s


def scale(x):
    return max(np.max(np.abs(x)), 1e-300)


# ---------------------------------------------------------------- CPU


def test_oracle_stats_against_numpy():
    cfg = small_cfg(N=8, ny=11)
    o = oracle_for(cfg)
    f = rand_fields(o, np.random.default_rng(7))
    prof, sp = o.stats()
    ref, refsp = numpy_stats(cfg, o.grid(), f)
    for name, err in profile_errors(prof, ref, cfg).items():
        assert err < 1e-13, (name, err)
    nyu = cfg["ny"] + 1
    for w, name in enumerate(statsio.SPECTRA):
        got = sp[w][:, : refsp[name].shape[1]] if name != "P" else sp[w][:, 1:nyu - 1]
        assert np.allclose(got, refsp[name], rtol=1e-14, atol=0), name
    # a second sample adds
    prof2, _ = o.stats(prof.copy(), sp.copy())
    assert np.allclose(prof2, 2 * prof, rtol=1e-14, atol=1e-12)


def test_oracle_stats_closed_forms():
    """Mean-mode and single-mode fields: sum = n Re f00, sum of squares =
    2 n |A|^2 for one i > 0 mode; omega_x of a u3 mean profile is its y
    derivative."""
    cfg = small_cfg(N=8, ny=11)
    o = oracle_for(cfg)
    ny, ncol = cfg["ny"], o.ncol
    n = cfg["Ngal1"] * cfg["Ngal2"]
    U = np.linspace(0.3, 1.7, ny + 1)
    u1 = np.zeros((ncol, ny + 1), complex)
    u1[0] = U
    A = 0.25 - 0.4j
    u3 = np.zeros((ncol, ny + 1), complex)
    u3[2 + 5 * 3] = A                       # i = 2, k = 3
    W = np.linspace(-1.0, 2.0, ny + 1)
    u3[0] = W
    o.set("u1", u1)
    o.set("u3", u3)
    o.set("u2", np.zeros((ncol, ny), complex))
    o.set("p", np.zeros((ncol, ny - 1), complex))
    prof, _ = o.stats()
    q = {nm: i for i, nm in enumerate(NAMES)}
    j = np.arange(1, ny)
    assert np.allclose(prof[q["Um"], j], n * U[j], rtol=1e-14)
    assert np.allclose(prof[q["U2m"], j], n * U[j] ** 2, rtol=1e-13)
    assert np.allclose(prof[q["Wm"], j], n * W[j], rtol=1e-13, atol=1e-12)
    assert np.allclose(prof[q["W2m"], j], n * (W[j] ** 2 + 2 * abs(A) ** 2), rtol=1e-13)
    assert np.allclose(prof[q["UWm"], j], n * U[j] * W[j], rtol=1e-13, atol=1e-12)
    # ghost rows keep the spectral plane: sum(re + im) of the stored modes
    assert np.isclose(prof[q["Um"], 0], U[0]) and np.isclose(prof[q["Wm"], ny], W[ny] + A.real + A.imag)
    dthdyv = o.grid()["dthdyv"]
    jv = np.arange(ny)
    assert np.allclose(prof[q["wxm"], jv], cfg["N1"] * cfg["N2"] * (W[jv + 1] - W[jv]) * dthdyv[jv], rtol=1e-13)


@pytest.mark.parametrize("seed", [3, 11])
def test_parseval_form_matches_oracle(seed):
    """The spectral-space evaluation the kernel uses equals the physical
    sums, also for i = 0 columns that are not Hermitian and a non-zero z
    Nyquist index."""
    cfg = small_cfg(N=8, ny=9)
    o = oracle_for(cfg)
    f = rand_fields(o, np.random.default_rng(seed))
    prof, _ = o.stats()
    ny = cfg["ny"]
    q = {nm: i for i, nm in enumerate(NAMES)}
    g = o.grid()
    u1, u3 = f["u1"], f["u3"]
    for j in range(1, ny):
        assert np.isclose(parseval_row(cfg, u1[:, j], u1[:, j], "G"), prof[q["U2m"], j], rtol=1e-13)
        assert np.isclose(parseval_row(cfg, u1[:, j], u3[:, j], "G"), prof[q["UWm"], j], rtol=1e-12,
                          atol=1e-12 * scale(prof[q["U2m"]]))
    N2, nxh = cfg["N2"], cfg["N1"] // 2 + 1
    kz = np.array([cfg["bet"] * k if k < N2 // 2 else -cfg["bet"] * (N2 - k) for k in range(N2)])
    for j in range(ny):
        wx = -(1j * np.repeat(kz, nxh)) * f["u2"][:, j] + (u3[:, j + 1] - u3[:, j]) * g["dthdyv"][j]
        assert np.isclose(parseval_row(cfg, wx, wx, "N"), prof[q["wx2m"], j], rtol=1e-13)


def test_stats_file_round_trip(tmp_path):
    cfg = small_cfg(N=8, ny=11)
    o = oracle_for(cfg)
    rand_fields(o, np.random.default_rng(1))
    prof, sp = o.stats()
    g = o.grid()
    ny = cfg["ny"]
    Ngal = np.array([[12, 12, ny - 2, ny - 1]] * 3).T
    hdr = dict(t=1.5, Re=cfg["Re"], alp=cfg["alp"], bet=cfg["bet"], mpgx=0.1, nband=1, iter=20,
               Ngal=Ngal, N=Ngal, yu=g["yu"], yv=g["yv"], dthdyu=g["dthdyu"], dthdyv=g["dthdyv"])
    p = tmp_path / "stats.dat"
    statsio.write_stats(p, prof, 3, hdr)
    r = statsio.read_stats(p)
    assert r["istat"] == 3 and r["iter"] == 20 and r["t"] == 1.5
    assert np.array_equal(r["N"], Ngal)
    assert np.array_equal(r["yu"], g["yu"]) and np.array_equal(r["dthdyv"], g["dthdyv"])
    for qq, name in enumerate(NAMES):
        j0, n = {"u": (0, ny + 1), "v": (0, ny), "p": (1, ny - 1)}[statsio.GRID_OF[name]]
        assert np.array_equal(r[name], prof[qq, j0:j0 + n]), name
    # record layout: header, Ngal, two grid records, istat, 13 profiles
    rows = {"u": ny + 1, "v": ny, "p": ny - 1}
    size = (8 + 5 * 8 + 2 * 4 + 88 * 4) + (8 + 4 * Ngal.size) + (8 + 8 * (2 * (ny + 1) + 1)) \
        + (8 + 8 * (2 * ny + 1)) + (8 + 4) + sum(8 + 8 * rows[statsio.GRID_OF[n]] for n in NAMES)
    assert len(p.read_bytes()) == size


def test_spectrum_fold():
    """recvwrspec for one process: negative z modes overwrite their positive
    partner (k -> N2+2-k), i = 1..N1/2 and k = 1..N2/2 are doubled."""
    N1, N2 = 8, 8
    nxh = N1 // 2 + 1
    ncol = nxh * N2
    sp = np.arange(ncol, dtype=float)[:, None] * np.array([[1.0, 10.0]])
    out = statsio.fold_spectrum(sp, N1, N2)
    assert out.shape == (2, N2 // 2 + 1, nxh)
    for i in range(nxh):
        for kk in range(1, N2 // 2 + 2):
            k_src = kk if kk in (1, N2 // 2 + 1) else N2 + 2 - kk
            v = sp[i + nxh * (k_src - 1), 0]
            if 1 <= i <= N1 // 2 and kk <= N2 // 2:
                v *= 2
            assert out[0, kk - 1, i] == v


# ---------------------------------------------------------------- GPU


def _gpu_vs_oracle(cfgs, steps, seed, perturb0=True):
    """Device statistics after each of the step counts in `steps`, against
    (a) the oracle's statistics of the device's own fields (the kernel alone:
    rounding level) and (b) the oracle run independently from the same
    initial fields (the kernel plus the step's parity)."""
    from rnl_multiblock_nofft_amd import Solver
    s = Solver(cfgs)
    oracles, mirrors, indep, mirr = [], [], [], []
    rng = np.random.default_rng(seed)
    for q, cfg in enumerate(cfgs):
        o, m = oracle_for(cfg), oracle_for(cfg)
        f = init_fields(cfg, s.grid(q), seed=seed + q)
        if perturb0:
            # i = 0 columns that are not Hermitian in z, and a z-Nyquist mode
            nxh, N2 = cfg["N1"] // 2 + 1, cfg["N2"]
            for name in ("u1", "u3"):
                a = f[name]
                a[::nxh, 1:-1] += 1e-3 * (rng.standard_normal((N2, a.shape[1] - 2))
                                          + 1j * rng.standard_normal((N2, a.shape[1] - 2)))
        for k, v in f.items():
            o.set(k, v)
            s.upload(q, k, v)
        o.init_flowrate()
        oracles.append(o)
        mirrors.append(m)
        indep.append(None)
        mirr.append(None)
    s.init_flowrate()
    s.stats_reset()
    for n in steps:
        s.step(n)
        for q, o in enumerate(oracles):
            o.step(n)
            indep[q] = o.stats(*indep[q]) if indep[q] is not None else o.stats()
            m = mirrors[q]
            for name in ("u1", "u2", "u3", "p"):
                m.set(name, s.download(q, name))
            mirr[q] = m.stats(*mirr[q]) if mirr[q] is not None else m.stats()
        s.stats_accumulate()
    for q, cfg in enumerate(cfgs):
        prof, istat = s.stats(q)
        assert istat == len(steps)
        ny = cfg["ny"]
        for ref, tol, what in ((mirr[q], 1e-13, "same fields"), (indep[q], 1e-11, "independent run")):
            errs = profile_errors(prof, ref[0], cfg, istat)
            print(q, what, " ".join("%s %.1e" % kv for kv in errs.items()), flush=True)
            for name, err in errs.items():
                assert err < tol, (q, what, name, err)
            for w, name in enumerate(statsio.SPECTRA):
                got = s.spectrum(q, name)
                exp = ref[1][w][:, 1:ny] if name == "P" else ref[1][w][:, : got.shape[1]]
                err = np.max(np.abs(got - exp)) / scale(exp)
                assert err < tol, (q, what, name, err)
    return s


@pytest.mark.gpu
def test_gpu_stats_match_oracle():
    cfg = small_cfg(N=16, ny=21, Re=1000.0)
    _gpu_vs_oracle([cfg], [1, 2], seed=21)


@pytest.mark.gpu
def test_gpu_stats_textured_ensemble():
    from rnl_multiblock_nofft_amd import textures
    a = small_cfg(N=16, ny=21)
    b = small_cfg(N=8, ny=17, Re=2000.0)
    m = textures.stripes(16, 16, 4, 0.5)
    a["texture"] = (m, m)
    _gpu_vs_oracle([a, b], [1, 1, 1], seed=5)


@pytest.mark.gpu
def test_gpu_stats_reset():
    from rnl_multiblock_nofft_amd import Solver
    cfg = small_cfg(N=8, ny=13)
    s = Solver([cfg])
    prof, n = s.stats(0)
    assert n == 0 and not prof.any()
    for k, v in init_fields(cfg, s.grid(0), seed=2).items():
        s.upload(0, k, v)
    s.stats_accumulate()
    s.stats_accumulate()
    prof2, n2 = s.stats(0)
    assert n2 == 2 and prof2.any()
    s.stats_reset()
    prof3, n3 = s.stats(0)
    assert n3 == 0 and not prof3.any() and not s.spectrum(0, "U").any()
    s.stats_accumulate()
    prof4, n4 = s.stats(0)
    assert n4 == 1 and np.allclose(prof4, prof2 / 2, rtol=1e-15, atol=0)


# ---------------------------------------------------------------- sl_stats


def test_oracle_slstats_closed_forms():
    """Wall maps of sl_stats (sl_stats2.f90:33-161) for a profile linear
    near the walls: the interpolated wall velocity is the linear
    extrapolation, du/dy the slope; a profile symmetric about the centre
    gives equal bottom and (sign-flipped) top gradients."""
    cfg = small_cfg(N=8, ny=13)
    o = oracle_for(cfg)
    g = o.grid()
    yu = g["yu"]
    ny, ncol = cfg["ny"], o.ncol
    u1 = np.zeros((ncol, ny + 1), complex)
    u1[0] = 1.0 - yu ** 2                      # symmetric about y = 0
    u3 = np.zeros((ncol, ny + 1), complex)
    u3[0] = 0.5 + 2.0 * yu                     # linear
    o.set("u1", u1)
    o.set("u3", u3)
    o.set("u2", np.zeros((ncol, ny), complex))
    o.set("p", np.zeros((ncol, ny - 1), complex))
    m = o.slstats()
    yv = g["yv"]
    # linear u3: value at the wall faces and slope (bottom +2, top -2 after the flip)
    assert m[0, 2, 0, 0] == pytest.approx(0.5 + 2.0 * yv[0], rel=1e-13)
    assert m[1, 2, 0, 0] == pytest.approx(0.5 + 2.0 * yv[ny - 1], rel=1e-13)
    # the staggered difference across the wall face, d/dtheta times dtheta/dy
    dv = g["dthdyv"]
    assert m[0, 3, 0, 0] == pytest.approx(2.0 * (yu[1] - yu[0]) * dv[0], rel=1e-13)
    assert m[1, 3, 0, 0] == pytest.approx(-2.0 * (yu[ny] - yu[ny - 1]) * dv[ny - 1], rel=1e-13)
    assert m[0, 3, 0, 0] == pytest.approx(2.0, rel=0.05)        # the slope, to O(dy^2)
    # symmetric u1: equal gradients at both walls after the flip, equal values
    assert m[0, 1, 0, 0] == pytest.approx(m[1, 1, 0, 0], rel=1e-12)
    assert m[0, 0, 0, 0] == pytest.approx(m[1, 0, 0, 0], rel=1e-12)
    # only the mean mode is set; the imaginary slot and other modes stay 0
    assert np.count_nonzero(m[:, :, :, 1:]) == 0


def test_sl_stats_file_layout(tmp_path):
    import os
    cfg = small_cfg(N=8, ny=13)
    o = oracle_for(cfg)
    rand_fields(o, np.random.default_rng(2))
    m = o.slstats()
    N = np.array([[0, 8, 8, 8, -2], [0, 8, 8, 8, 0], [0, 10, 10, 10, 0], [0, 11, 11, 11, 0]], np.int32)
    hdr = dict(t=0.5, Re=cfg["Re"], alp=1.0, bet=2.0, mpgx=0.0, nband=3, iter=7, N=N)
    p = str(tmp_path / "sl.dat")
    statsio.write_sl_stats(p, m, 1, hdr)
    r = statsio.read_sl_stats(p)
    assert r["istat"] == 1 and r["iter"] == 7 and np.array_equal(r["N"], N)
    for q in range(4):
        assert np.array_equal(r["maps"][q], m[0, q] + m[1, q])
    ref = "/root/reference/output/sl_stats_0032x0032x0153_00052.dat"
    if os.path.exists(ref):
        rr = statsio.read_sl_stats(ref)
        assert [mm.shape for mm in rr["maps"]] == [(32, 34)] * 4 and rr["istat"] == 1
        assert os.path.getsize(ref) == 8 + 400 + 8 + 80 + 8 + 4 + 4 * (8 + 8 * 34 * 32)


@pytest.mark.gpu
def test_gpu_slip_stats_match_oracle():
    from rnl_multiblock_nofft_amd import Solver
    cfg = small_cfg(N=16, ny=21)
    s = Solver([cfg])
    o, m = oracle_for(cfg), oracle_for(cfg)
    f = init_fields(cfg, s.grid(0), seed=13)
    for k, v in f.items():
        s.upload(0, k, v)
        o.set(k, v)
    s.init_flowrate()
    o.init_flowrate()
    s.stats_reset()
    ref = mir = None
    for _ in range(2):
        s.step(1)
        o.step(1)
        ref = o.slstats(ref)
        for name in ("u1", "u2", "u3", "p"):
            m.set(name, s.download(0, name))
        mir = m.slstats(mir)
        s.stats_accumulate()
    got = s.slip_stats(0)
    for what, r, tol in (("same fields", mir, 1e-14), ("independent run", ref, 1e-11)):
        for w in range(2):
            # the wall velocity of a no-slip wall is round-off: every map is
            # measured against the largest map of its wall (the gradients)
            sc = max(np.abs(r[w]).max(), 1e-300)
            for q in range(4):
                e = np.abs(got[w, q] - r[w, q]).max() / sc
                print(what, w, q, "%.1e" % e, flush=True)
                assert e < tol, (what, w, q, e)
