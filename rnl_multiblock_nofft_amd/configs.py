"""Channel configurations: the reference's input.in parameters plus the
knobs this build adds (fixed dt, grid kind, precision).

`Re` is the reference's `Re`, i.e. 1/nu (C1 = aRK/Re, littleharsh.f90:651).
A bulk Reynolds number Re_b = U_b h / nu with the laminar start u = 1 - y^2
(U_b = 2/3) maps to Re = 1.5 Re_b: Re_b = 2800 -> 4200, Re_b = 10000 -> 15000.
Sizes count Fourier modes, N = 2 (Ngal/3) (start.f90:112-115), so "32x32"
runs on a 48x48 physical grid.
"""
import copy

import numpy as np

from . import textures

BASE = dict(alp=1.0, bet=2.0, grid_kind=1, gamma=2.0, dyq=10.0, ppp=5,
            bslpu1=0.0, bslpu3=0.0, flag_ctpress=0, mpgx=0.0, dt=1e-3)


def channel(N1, N2, ny, Re, **kw):
    c = dict(BASE)
    c.update(N1=N1, N2=N2, Ngal1=3 * N1 // 2, Ngal2=3 * N2 // 2, ny=ny, Re=Re)
    c.update(kw)
    return c


def re_code(re_bulk):
    return 1.5 * re_bulk


# The workloads.  Textures follow the configurations they model: free-shear
# / no-slip masks on both walls (textures.py), kept in cfg["texture"].
def smooth_32():
    """Smooth-wall 32x32x153 channel, Re_b = 2800."""
    return [channel(32, 32, 153, re_code(2800.0))]


def ridges_64():
    """64x64x153, Re_b = 2800, both walls with streamwise ridges: alternating
    free-slip / no-slip z-stripes of period Lz/8, gas fraction 0.5."""
    m = textures.stripes(64, 64, 8, 0.5)
    return [channel(64, 64, 153, re_code(2800.0), texture=(m, m))]


def ensemble_64():
    """48 channels of 32x32x153 and 16 of 64x64x153, each with square posts
    at a spacing drawn from 2..8 points and a gas fraction from U(0.3, 0.9),
    seeded per channel (seed 1234 + channel)."""
    out = []
    for q in range(64):
        N = 32 if q < 48 else 64
        rng = np.random.default_rng(1234 + q)
        spacing = int(rng.integers(2, 9))
        phi = float(rng.uniform(0.3, 0.9))
        m = textures.posts_for_gas_fraction(N, N, spacing, phi)
        out.append(channel(N, N, 153, re_code(2800.0), texture=(m, m)))
    return out


def large_256():
    """256x256x257, Re_b = 10000, square posts of 8x8 points at spacing 16
    (gas fraction 0.75) on both walls."""
    m = textures.posts(256, 256, 16, 8)
    return [channel(256, 256, 257, re_code(10000.0), texture=(m, m))]


def fine_128():
    return [channel(128, 128, 153, re_code(2800.0))]


def banded_256():
    """The 256x256x257 channel (Re_b = 10000, smooth walls) with
    multi-resolution bands: 256x256 modes in wall bands of height 0.25
    (ygrid's split, start.f90:999-1011) and 128x128 modes in the middle
    band."""
    from .solver import band_rows
    wall = channel(256, 256, 257, re_code(10000.0))
    return banded(wall, 128, 128, band_rows(wall, 0.25))


def banded(wall, N1m, N2m, nb):
    """The three channel configurations of a banded channel (lh_set_bands):
    the bottom wall band, the middle band and the top wall band.  `wall` is
    the whole channel at the wall bands' resolution; the middle band takes
    N1m x N2m modes over the same ny, the wall bands the wall's modes over
    nb + 2 v rows each (v rows 0..nb+1 and ny-2-nb..ny-1; nb from
    solver.band_rows, ygrid's split of start.f90:999-1011) and no mean mode.
    Returns [bottom, middle, top]."""
    mid = dict(wall, N1=N1m, N2=N2m, Ngal1=3 * N1m // 2, Ngal2=3 * N2m // 2)
    w = dict(wall, ny=nb + 2, flag_ctpress=1, mpgx=0.0)
    for c in (mid, w):
        c.pop("texture", None)
    out = [dict(w), mid, dict(w)]
    for q, c in enumerate(out):
        c["band"] = (q, nb)  # Solver links the three (lh_set_bands)
    return out


def smooth(cfgs):
    """The same channels with smooth walls."""
    out = [dict(c) for c in cfgs]
    for c in out:
        c.pop("texture", None)
    return out


WORKLOADS = {"smooth_32x32x153": smooth_32, "ridges_64x64x153": ridges_64,
             "ensemble_48x32_16x64": ensemble_64, "large_256x256x257": large_256,
             "fine_128x128x153": fine_128, "banded_256x256x257_mid128": banded_256}


def reference_input_case():
    """The reference's own input.in case (input.in:10-38): Re=3250, alp=1,
    bet=2, 32x32 modes (Ngal 48), ny=153, dyq=10, pp=5, constant pressure
    gradient mpgx=-3.0675e-3; dt = 2.5 dtv = 5.287867162159505e-3 (the
    viscous cap the reference run took, slurm-17059240.out:132-135)."""
    return [channel(32, 32, 153, 3250.0, grid_kind=0, dyq=10.0, ppp=5,
                    flag_ctpress=1, mpgx=-3.0675e-3, dt=0.005287867162159505)]


def with_precision(cfgs, precision):
    out = copy.deepcopy(cfgs)
    for c in out:
        c["precision"] = precision
    return out
