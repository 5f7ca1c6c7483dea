"""Wall texture masks on the N1 x N2 grid of a channel's Fourier modes.

A mask is an int32 array of shape (N2, N1), mask[z, x], holding the codes of
the reference's planeBC array (start.f90:1776-1829): 1 = no-slip point,
2 = free-shear point.  mask[z, x] is planeBC(x+1, z+1).
"""
import numpy as np

NO_SLIP, FREE_SHEAR = 1, 2


def planebc(N1, N2, ntilex, ntilez, Lfracx=0.0, Lfracz=0.0):
    """The reference's post geometry (start.f90:1776-1829): ntilex x ntilez
    tiles; each tile starts with a no-slip post of postpointsx x postpointsz
    points, postpoints = texturepoints/Lfrac + 1 (the whole tile when both
    Lfrac are 0); everything else free-shear."""
    texx, texz = N1 // ntilex, N2 // ntilez
    if Lfracx == 0 and Lfracz == 0:
        px, pz = texx, texz
    else:
        fx, fz = N1 / 2.0 / ntilex / Lfracx + 1, N2 / ntilez / Lfracz + 1
        if fx % 1 != 0 or fz % 1 != 0:
            raise ValueError("post size is not a whole number of points (start.f90:1794-1802)")
        px, pz = int(N1 / ntilex / Lfracx) + 1, int(N2 / ntilez / Lfracz) + 1
    pz = min(pz, N2)
    bc = np.full((N1 + 2, N2), FREE_SHEAR, dtype=np.int32)  # planeBC(0:N1+1, 1:N2)
    for ip in range(ntilex):
        for kp in range(ntilez):
            for i in range(1, px + 1):
                for k in range(1, pz + 1):
                    ii, kk = i + ip * texx, k + kp * texz
                    if ii <= N1 + 1 and kk <= N2:
                        bc[ii, kk - 1] = NO_SLIP
    return np.ascontiguousarray(bc[1:N1 + 1, :].T)


def stripes(N1, N2, nstripes, gas_fraction):
    """Streamwise ridges: nstripes periods across z, each with a free-shear
    band of round(gas_fraction * period) points followed by no-slip."""
    period = N2 // nstripes
    if period * nstripes != N2:
        raise ValueError("N2 must be a multiple of the number of stripes")
    ngas = int(round(gas_fraction * period))
    col = np.full(N2, NO_SLIP, dtype=np.int32)
    for s in range(nstripes):
        col[s * period:s * period + ngas] = FREE_SHEAR
    return np.ascontiguousarray(np.repeat(col[:, None], N1, axis=1))


def posts(N1, N2, spacing, size):
    """Square no-slip posts of size x size points on a square lattice of the
    given spacing (points) in x and z; the rest free-shear."""
    m = np.full((N2, N1), FREE_SHEAR, dtype=np.int32)
    for z0 in range(0, N2, spacing):
        for x0 in range(0, N1, spacing):
            m[z0:z0 + size, x0:x0 + size] = NO_SLIP
    return m


def posts_for_gas_fraction(N1, N2, spacing, gas_fraction):
    """Posts whose area fraction is as close as the grid allows to
    1 - gas_fraction."""
    size = max(1, min(spacing, int(round(spacing * np.sqrt(1.0 - gas_fraction)))))
    return posts(N1, N2, spacing, size)
