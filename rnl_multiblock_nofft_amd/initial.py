"""Initial fields: laminar Poiseuille flow plus a seeded, discretely
divergence-free random perturbation, in the reference's column layout.

Fills the role of getini/mblock_ini_parabolic_profile (start.f90:2144-2440)
for synthetic starts.  Each channel's draws come from a counter-based Philox
generator keyed by its seed (default 1234 + channel index).
"""
import numpy as np


def _kz_index(k, N2):
    """signed z wavenumber index of column z-index k (0-based), def_k order"""
    return k if k < N2 // 2 else k - N2


def perturbed_poiseuille(cfg, grid, seed, amp=0.1, kmax=8):
    """Return dict u1 (ncol, ny+1), u2 (ncol, ny), u3 (ncol, ny+1),
    p (ncol, ny-1) complex128 for one channel.

    grid: dict with yu, yv, dthdyu, dthdyv (as lh_get_grid returns).
    u1 mean mode = 1 - yu^2; perturbation modes 0 < |(i, kz)| <= kmax with
    u, w = even and odd wall-normal shapes (1-y^2)^2, y(1-y^2)^2, the even
    part projected so that kx a + kz b = 0, and v from discrete continuity
    (divergence, littleharsh.f90:237-272) integrated from the lower wall.
    Closing v at the top wall leaves the last row divergent, so the
    perturbation is then projected once, as solveP + v_corr do
    (littleharsh.f90:581-633, 956-1052): (D_y G_y + k^2) phi = div with the
    Neumann rows of LU_buildP (tridLU_3D.f90:333-407), u -= grad phi; the
    discrete divergence is then zero to rounding on every pressure row.
    Ghost rows follow the no-slip ghost rule u(0) = gw u(1)
    (v_corr, littleharsh.f90:992-1044)."""
    N1, N2, ny = cfg["N1"], cfg["N2"], cfg["ny"]
    alp, bet = cfg["alp"], cfg["bet"]
    nxh = N1 // 2 + 1
    ncol = nxh * N2
    nyu = ny + 1
    yu, yv, dyu = grid["yu"], grid["yv"], grid["dthdyu"]
    ge = (1.0 - yu ** 2) ** 2
    go = yu * (1.0 - yu ** 2) ** 2
    rng = np.random.Generator(np.random.Philox(key=int(seed)))
    u1 = np.zeros((ncol, nyu), np.complex128)
    u3 = np.zeros((ncol, nyu), np.complex128)
    u2 = np.zeros((ncol, ny), np.complex128)
    coef = {}
    for k in range(N2):
        kk = _kz_index(k, N2)
        for i in range(nxh):
            if i == 0 and kk < 0:
                continue
            if i * i + kk * kk == 0 or i * i + kk * kk > kmax * kmax or abs(kk) >= N2 // 2:
                continue
            z = rng.standard_normal(8)
            a, b = z[0] + 1j * z[1], z[2] + 1j * z[3]
            a2, b2 = z[4] + 1j * z[5], z[6] + 1j * z[7]
            kx, kz = alp * i, bet * kk
            s = (kx * a + kz * b) / (kx * kx + kz * kz)
            a, b = a - s * kx, b - s * kz
            coef[(i, kk)] = (a, b, a2, b2)
    for (i, kk), (a, b, a2, b2) in coef.items():
        k = kk if kk >= 0 else kk + N2
        c = k * nxh + i
        u1[c] = a * ge + a2 * go
        u3[c] = b * ge + b2 * go
        if i == 0:  # Hermitian partner of the x-mean modes
            kn = (-kk) % N2
            cn = kn * nxh
            u1[cn] = np.conj(u1[c])
            u3[cn] = np.conj(u3[c])
    # scale to max |u'| = amp on the physical grid (3/2 padded)
    up = _phys_max(u1, N1, N2, cfg["Ngal1"], cfg["Ngal2"])
    scale = amp / up if up > 0 else 0.0
    u1 *= scale
    u3 *= scale
    cols = np.arange(ncol)
    kxs = alp * (cols % nxh)
    kzs = bet * np.array([_kz_index(k, N2) for k in cols // nxh], dtype=float)
    # v from discrete continuity: v(j) = v(j-1) - (ikx u + ikz w)(j) / dthdyu(j)
    for j in range(1, ny):
        D = 1j * kxs * u1[:, j] + 1j * kzs * u3[:, j]
        u2[:, j] = u2[:, j - 1] - D / dyu[j]
    u2[:, ny - 1] = 0.0
    u2[:, 0] = 0.0
    _project(u1, u2, u3, kxs, kzs, grid)
    # mean flow
    u1[0, 1:nyu - 1] += 1.0 - yu[1:nyu - 1] ** 2
    gw = (yu[0] - yv[0]) / (yu[1] - yv[0])
    u1[:, 0] = gw * u1[:, 1]
    u1[:, nyu - 1] = gw * u1[:, nyu - 2]
    u3[:, 0] = gw * u3[:, 1]
    u3[:, nyu - 1] = gw * u3[:, nyu - 2]
    p = np.zeros((ncol, ny - 1), np.complex128)
    return dict(u1=u1, u2=u2, u3=u3, p=p)


def divergence(u1, u2, u3, kxs, kzs, dthdyu):
    """divergence on the pressure rows j = 1..ny-1 (littleharsh.f90:237-272)"""
    ny = u2.shape[1]
    j = np.arange(1, ny)
    return (1j * kxs[:, None] * u1[:, j] + 1j * kzs[:, None] * u3[:, j]
            + dthdyu[j][None, :] * (u2[:, j] - u2[:, j - 1]))


def _project(u1, u2, u3, kxs, kzs, grid):
    """u -= grad phi with (D_y G_y + k^2) phi = div, v fixed at both walls
    (Neumann rows); every column but the mean mode (whose divergence is
    zero: its v vanishes).  Batched Thomas sweep over the columns."""
    du, dv = grid["dthdyu"], grid["dthdyv"]
    ny = u2.shape[1]
    ne = ny - 1                                   # pressure rows 1..ne
    rhs = divergence(u1, u2, u3, kxs, kzs, du)    # (ncol, ne)
    k2 = -(kxs * kxs + kzs * kzs)
    j = np.arange(1, ne + 1)
    lo = du[j] * dv[j - 1]                        # phi(j-1)
    up = du[j] * dv[np.minimum(j, ny - 1)]        # phi(j+1)
    lo[0] = 0.0                                   # v(0) fixed
    up[-1] = 0.0                                  # v(ny-1) fixed
    sel = np.arange(1, len(kxs))                  # skip the mean mode
    diag = (-lo - up)[None, :] + k2[sel, None]
    r = rhs[sel]
    n = ne
    cp = np.zeros((len(sel), n))
    dp = np.zeros((len(sel), n), np.complex128)
    cp[:, 0] = up[0] / diag[:, 0]
    dp[:, 0] = r[:, 0] / diag[:, 0]
    for q in range(1, n):
        m = diag[:, q] - lo[q] * cp[:, q - 1]
        cp[:, q] = up[q] / m
        dp[:, q] = (r[:, q] - lo[q] * dp[:, q - 1]) / m
    phi = np.zeros((len(kxs), ne + 2), np.complex128)  # rows 0..ny (ends unused)
    x = dp[:, n - 1]
    phi[sel, n] = x
    for q in range(n - 2, -1, -1):
        x = dp[:, q] - cp[:, q] * x
        phi[sel, q + 1] = x
    jj = np.arange(1, ne + 1)
    u1[:, jj] -= 1j * kxs[:, None] * phi[:, jj]
    u3[:, jj] -= 1j * kzs[:, None] * phi[:, jj]
    jv = np.arange(1, ny - 1)
    u2[:, jv] -= dv[jv][None, :] * (phi[:, jv + 1] - phi[:, jv])


def _phys_max(f, N1, N2, Ng1, Ng2):
    """max over rows of the physical-space |f| on the Ng1 x Ng2 grid"""
    nxh = N1 // 2 + 1
    ncol, nr = f.shape
    spec = np.zeros((nr, Ng2, Ng1 // 2 + 1), np.complex128)
    c = f.T.reshape(nr, N2, nxh)
    spec[:, :N2 // 2, :nxh] = c[:, :N2 // 2, :]
    spec[:, Ng2 - N2 // 2:, :nxh] = c[:, N2 // 2:, :]
    phys = np.fft.irfft2(spec, s=(Ng2, Ng1), axes=(1, 2)) * (Ng1 * Ng2)
    return float(np.abs(phys).max())
