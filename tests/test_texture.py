"""Textured walls (free-shear / no-slip masks): the CPU oracle against a
dense direct solve of the mode-coupled Helmholtz system, and the limits that
fix the formulation (all no-slip = smooth wall, all free-shear = zero wall
gradient), plus the mask generators."""
import numpy as np

from helpers import oracle_for, small_cfg
from rnl_multiblock_nofft_amd import textures


def T_inv(cols, N1, N2):
    """four_to_phys_N with numpy: columns (ncol,) complex, c = i + nxh k ->
    physical (N2, N1) real.  z: unscaled inverse DFT; x: real inverse with
    the imaginary parts of modes 0 and N1/2 ignored."""
    nxh = N1 // 2 + 1
    cc = cols.reshape(N2, nxh)
    zz = np.fft.ifft(cc, axis=0) * N2
    return np.fft.irfft(zz, n=N1, axis=1) * N1


def T_fwd(phys, N1, N2):
    """phys_to_four_N with numpy: forward real x transform / N1 (imaginary
    parts of modes 0 and N1/2 zero), forward z transform / N2."""
    cc = np.fft.rfft(phys, axis=1) / N1
    cc[:, 0] = cc[:, 0].real
    cc[:, -1] = cc[:, -1].real
    return (np.fft.fft(cc, axis=0) / N2).reshape(-1)


def dense_textured_solve(o, cfg, kRK, rhs, mb, mt):
    """Real-linear dense solve of A(k) u = rhs with the wall rows replaced
    point by point: free-shear u0 - u1 = rhs0 (u_ny - u_{ny-1} at the top),
    no-slip the smooth rows."""
    N1, N2, ny = cfg["N1"], cfg["N2"], cfg["ny"]
    nyu, nxh = ny + 1, N1 // 2 + 1
    ncol = nxh * N2
    g = o.grid()
    dyu2i = g["dyu2i"]
    aRK = [4 / 15, 1 / 15, 1 / 6]
    beta = aRK[kRK - 1] * cfg["dt"] / cfg["Re"]
    yu, yv = g["yu"], g["yv"]
    gw1 = -(yv[0] - yu[0]) / (yv[0] - yu[1])
    gw2 = (yu[ny] - yv[ny - 1]) / (yv[ny - 1] - yu[ny - 1])
    alp, bet = cfg["alp"], cfg["bet"]
    kx = alp * np.arange(nxh)
    kz = np.array([bet * k if k < N2 // 2 else -bet * (N2 - k) for k in range(N2)])
    n = 2 * ncol * nyu
    M = np.zeros((n, n))
    b = np.zeros(n)

    def idx(c, j, part):
        return 2 * (c * nyu + j) + part

    for c in range(ncol):
        i, k = c % nxh, c // nxh
        k2 = -(kx[i] ** 2) - kz[k] ** 2
        for j in range(1, nyu - 1):
            a1, a2, a3 = -beta * dyu2i[j, 0], 1 - beta * (dyu2i[j, 1] + k2), -beta * dyu2i[j, 2]
            for part in (0, 1):
                r = idx(c, j, part)
                M[r, idx(c, j - 1, part)] += a1
                M[r, idx(c, j, part)] += a2
                M[r, idx(c, j + 1, part)] += a3
        for part in (0, 1):
            M[idx(c, 0, part), idx(c, 0, part)] += 1.0
            M[idx(c, 0, part), idx(c, 1, part)] += gw1
            M[idx(c, nyu - 1, part), idx(c, nyu - 1, part)] += 1.0
            M[idx(c, nyu - 1, part), idx(c, nyu - 2, part)] += gw2
        for j in range(nyu):
            b[idx(c, j, 0)] = rhs[c, j].real
            b[idx(c, j, 1)] = rhs[c, j].imag
    # wall rows: minus T[P . T^-1 u_adj], P = 1 + gw on free-shear points
    for w, mask, gw in ((0, mb, gw1), (1, mt, gw2)):
        if mask is None:
            continue
        P = np.where(mask == 2, 1.0 + gw, 0.0)
        wall, adj = (0, 1) if w == 0 else (nyu - 1, nyu - 2)
        for c in range(ncol):
            for part in (0, 1):
                e = np.zeros(ncol, dtype=complex)
                e[c] = 1.0 if part == 0 else 1j
                out = T_fwd(P * T_inv(e, N1, N2), N1, N2)
                for c2 in range(ncol):
                    M[idx(c2, wall, 0), idx(c, adj, part)] -= out[c2].real
                    M[idx(c2, wall, 1), idx(c, adj, part)] -= out[c2].imag
    x = np.linalg.solve(M, b)
    return (x[0::2] + 1j * x[1::2]).reshape(ncol, nyu)


def test_numpy_transforms_round_trip_and_oracle_conventions():
    N1, N2 = 8, 8
    rng = np.random.default_rng(3)
    phys = rng.standard_normal((N2, N1))
    assert np.allclose(T_inv(T_fwd(phys, N1, N2), N1, N2), phys, atol=1e-14)


def test_oracle_textured_solve_matches_dense_solve():
    cfg = small_cfg(8, 9)
    o = oracle_for(cfg)
    rng = np.random.default_rng(11)
    N1, N2 = cfg["N1"], cfg["N2"]
    mb = rng.choice([1, 2], size=(N2, N1)).astype(np.int32)
    mt = textures.stripes(N1, N2, 2, 0.5)
    o.set_texture(mb, mt)
    ncol = (N1 // 2 + 1) * N2
    nyu = cfg["ny"] + 1
    # right-hand sides of a real field (Hermitian-consistent columns):
    # transform random physical rows
    def real_rows():
        r = np.zeros((ncol, nyu), dtype=complex)
        for j in range(nyu):
            r[:, j] = T_fwd(rng.standard_normal((N2, N1)), N1, N2)
        return r
    ru, rw = real_rows(), real_rows()
    # a large beta makes the wall coupling strong enough to matter
    for kRK in (1, 3):
        u, w = o.lusol_uw_tex(kRK, ru, rw)
        du = dense_textured_solve(o, cfg, kRK, ru, mb, mt)
        dw = dense_textured_solve(o, cfg, kRK, rw, mb, mt)
        assert np.abs(u - du).max() <= 1e-12 * np.abs(du).max()
        assert np.abs(w - dw).max() <= 1e-12 * np.abs(dw).max()


def test_all_no_slip_mask_is_the_smooth_solve():
    cfg = small_cfg(8, 17)
    o = oracle_for(cfg)
    rng = np.random.default_rng(5)
    ncol = 5 * 8
    r = rng.standard_normal((ncol, 18)) + 1j * rng.standard_normal((ncol, 18))
    smooth = o.lusol_helm(2, 0, r)
    o.set_texture(np.ones((8, 8), np.int32), np.ones((8, 8), np.int32))
    u, _ = o.lusol_uw_tex(2, r, r)
    assert np.array_equal(u, smooth)


def test_all_free_shear_mask_gives_zero_wall_gradient():
    cfg = small_cfg(8, 17)
    o = oracle_for(cfg)
    rng = np.random.default_rng(6)
    N1 = N2 = 8
    ncol = 5 * 8
    r = np.zeros((ncol, 18), dtype=complex)
    for j in range(1, 17):
        r[:, j] = T_fwd(rng.standard_normal((N2, N1)), N1, N2)
    fs = np.full((8, 8), 2, np.int32)
    o.set_texture(fs, fs)
    u, _ = o.lusol_uw_tex(1, r, r)
    scale = np.abs(u).max()
    assert np.abs(u[:, 0] - u[:, 1]).max() <= 1e-13 * scale
    assert np.abs(u[:, -1] - u[:, -2]).max() <= 1e-13 * scale


def test_textured_step_keeps_the_ghost_rule_and_continuity():
    """After a full RK3 step the ghost rows follow the masks point by point
    and the velocity stays divergence-free."""
    from helpers import init_fields
    cfg = small_cfg(16, 33)
    o = oracle_for(cfg)
    N1, N2 = 16, 16
    mb = textures.stripes(N1, N2, 4, 0.5)
    mt = textures.posts(N1, N2, 8, 4)
    o.set_texture(mb, mt)
    f = init_fields(cfg, o.grid(), seed=3)
    for nm in ("u1", "u2", "u3", "p"):
        o.set(nm, f[nm])
    o.init_flowrate()
    o.step(1)
    g = o.grid()
    yu, yv = g["yu"], g["yv"]
    gwbc = (yu[0] - yv[0]) / (yu[1] - yv[0])  # no slip length
    for nm in ("u1", "u3"):
        u = o.get(nm)
        for wall, adj, mask in ((0, 1, mb), (-1, -2, mt)):
            p0 = T_inv(u[:, wall], N1, N2)
            p1 = T_inv(u[:, adj], N1, N2)
            ratio = np.where(mask == 2, 1.0, gwbc)
            assert np.abs(p0 - ratio * p1).max() <= 1e-12 * np.abs(p1).max()
    assert np.abs(o.divergence()).max() <= 1e-12 * np.abs(o.get("u1")).max()


def test_mask_generators():
    m = textures.stripes(32, 32, 8, 0.5)
    assert m.shape == (32, 32)
    assert (m[:, 0] == m[:, 5]).all()                  # streamwise uniform
    assert (m == 2).mean() == 0.5
    p = textures.posts(256, 256, 16, 8)
    assert abs((p == 2).mean() - 0.75) < 1e-12
    # planeBC with one tile per period and Lfrac giving 4 points of post
    bc = textures.planebc(16, 16, 1, 4, Lfracx=0.0, Lfracz=0.0)
    assert (bc == 1).all()                             # posts fill the tiles
    bc = textures.planebc(16, 16, 2, 2, Lfracx=2.0, Lfracz=8.0 / 3.0)
    assert bc[0, 0] == 1 and bc.dtype == np.int32


def test_oracle_N_grid_transforms_match_numpy():
    """four_to_phys_N / phys_to_four_N of the oracle (FOU3D.f90:786-804,
    2274-2293) against numpy's transforms with the FFTPACK conventions."""
    cfg = small_cfg(N=8, ny=9)
    o = oracle_for(cfg)
    N1, N2 = cfg["N1"], cfg["N2"]
    rng = np.random.default_rng(12)
    nxh = N1 // 2 + 1
    P = rng.standard_normal((N2, N1 + 2))
    cols = (P[:, 0::2] + 1j * P[:, 1::2]).reshape(-1)
    phys = o.four_to_phys_N(P)
    assert np.allclose(phys[:, :N1], T_inv(cols, N1, N2), rtol=0, atol=1e-13)
    back = o.phys_to_four_N(phys)
    fwd = T_fwd(phys[:, :N1], N1, N2).reshape(N2, nxh)
    assert np.allclose(back[:, 0::2], fwd.real, atol=1e-14) and np.allclose(back[:, 1::2], fwd.imag, atol=1e-14)
