"""GPU parity of multi-resolution bands (lh_set_bands): a banded channel is
three channels of one context (bottom wall band, middle band, top wall
band), checked through the C-ABI against the oracle's banded step
(oracle/lh_oracle.c lhob_*, pinned in tests/test_oracle_bands.py)."""
import sys

import numpy as np
import pytest

from helpers import relerr, small_cfg
from oracle import pyoracle as po
from rnl_multiblock_nofft_amd import Solver, _lib, configs, initial
from rnl_multiblock_nofft_amd.solver import band_rows

KEYS = {f for f, _ in po.Cfg._fields_}
FIELDS = ("u1", "u2", "u3", "p")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ocfg(cfg):
    return {k: v for k, v in cfg.items() if k in KEYS}


def banded_pair(wall, N1m, N2m, seed=1234, extra=()):
    """A Solver holding the banded channel (channels 0, 1, 2, then `extra`
    ordinary channels) and the oracle's BandOracle, with the same start: a
    perturbed Poiseuille field on the middle band and a perturbation of the
    short pencils."""
    nb = band_rows(wall, 0.25)
    cfgs = configs.banded(wall, N1m, N2m, nb) + list(extra)
    s = Solver(cfgs)  # links the three bands (configs.banded marks them)
    assert s.bands == [(0, 1, 2, nb)]
    B = po.BandOracle(N1m, N2m, nb=nb, **ocfg(wall))
    M = B.band(1)
    f = initial.perturbed_poiseuille(cfgs[1], s.grid(1), seed, kmax=min(3, N1m // 2 - 1))
    for nm in FIELDS:
        s.upload(1, nm, f[nm])
        M.set(nm, f[nm])
    # the short pencils: the wall resolution's high modes of a perturbed
    # field on the whole channel, over each wall band's rows
    # (the middle band's grid is the whole channel's; every field's rows
    # r0 .. r0+n-1 of the reference layout are the band's rows)
    fw = initial.perturbed_poiseuille(wall, s.grid(1), seed + 7,
                                      kmax=min(8, wall["N1"] // 2 - 1, wall["N2"] // 2 - 1))
    for q in (0, 2):
        own = B.owns(q)
        r0 = B.row0(q)
        for nm in FIELDS:
            n = _rows(cfgs[q], nm)
            a = np.zeros((len(own), n), complex)
            a[own] = fw[nm][own, r0:r0 + n]
            s.upload(q, nm, a)
            B.band(q).set(nm, a)
    for q, c in enumerate(extra):
        g = initial.perturbed_poiseuille(c, s.grid(3 + q), seed + 1 + q, kmax=3)
        for nm in FIELDS:
            s.upload(3 + q, nm, g[nm])
    s.init_flowrate()
    B.init_flowrate()
    return s, B, nb


def _rows(cfg, nm):
    ny = cfg["ny"]
    return {"u1": ny + 1, "u3": ny + 1, "u2": ny, "p": ny - 1}[nm]


# The wall bands' p is small (about 1e-4 of the field scale at these
# sizes) and ill-conditioned: in the oracle alone, 1e-16 relative noise on
# the inputs moves it by 4e-13 relative after one step (condition about
# 4e3), so rounding differences of the O(1) products reach it as a few
# 1e-12.  Its bound is 20 times the velocities'.
P_WALL = 20.0


def compare(s, B, tol, label):
    for q in range(3):
        own = B.owns(q)
        for nm in FIELDS:
            e = relerr(s.download(q, nm)[own], B.band(q).get(nm)[own])
            t = tol * (P_WALL if (nm == "p" and q != 1) else 1.0)
            say(label, "band", q, nm, e)
            assert e <= t, (label, q, nm, e)


@pytest.mark.gpu
def test_banded_grid_rows_and_edges():
    wall = small_cfg(16, 49)
    s, B, nb = banded_pair(wall, 8, 8)
    for q in (0, 2):
        gs, go = s.grid(q), B.band(q).grid()
        for k in ("yu", "yv", "dthdyu", "dthdyv"):
            assert np.array_equal(gs[k], go[k]), (q, k)
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cols", [-1, 0, 3])
def test_banded_step_matches_oracle(cols):
    """One step and three steps of a 16x16 wall / 8x8 middle channel,
    ny = 49, for the three column-solve forms; the wall bands are compared
    on their own pencils."""
    wall = small_cfg(16, 49)
    s, B, nb = banded_pair(wall, 8, 8)
    s.set_tuning(_lib.TUNE_SOLVE_COLS, cols)
    say("banded nb", nb, "solve cols", cols)
    s.step(1)
    B.step(1)
    compare(s, B, 1e-12, "1 step")
    s.step(2)
    B.step(2)
    compare(s, B, 1e-11, "3 steps")
    # the short pencils grew from the wall bands' products
    assert np.abs(s.download(0, "u1")[B.owns(0)]).max() > 0
    s.close()


@pytest.mark.gpu
def test_equal_resolution_bands_equal_one_channel():
    """With the middle band at the wall resolution the banded step equals the
    step of the same channel without bands."""
    wall = small_cfg(16, 49)
    s, B, nb = banded_pair(wall, 16, 16)
    p = Solver([wall])
    f = initial.perturbed_poiseuille(wall, p.grid(0), 1234, kmax=3)
    for nm in FIELDS:
        p.upload(0, nm, f[nm])
    p.init_flowrate()
    s.step(2)
    p.step(2)
    for nm in FIELDS:
        e = relerr(s.download(1, nm), p.download(0, nm))
        say("equal resolution", nm, e)
        assert e <= 1e-13, (nm, e)
    s.close()
    p.close()


@pytest.mark.gpu
def test_banded_channel_in_an_ensemble():
    """A banded channel beside an ordinary channel in one launch: each
    matches its own oracle."""
    wall = small_cfg(16, 33)
    other = small_cfg(8, 17)
    s, B, nb = banded_pair(wall, 8, 8, extra=[other])
    o = po.Oracle(**ocfg(other))
    g = initial.perturbed_poiseuille(other, s.grid(3), 1235, kmax=3)
    for nm in FIELDS:
        o.set(nm, g[nm])
    o.init_flowrate()
    s.step(2)
    B.step(2)
    o.step(2)
    compare(s, B, 1e-11, "ensemble")
    for nm in FIELDS:
        e = relerr(s.download(3, nm), o.get(nm))
        say("ordinary channel", nm, e)
        assert e <= 1e-11, (nm, e)
    s.close()


@pytest.mark.gpu
def test_set_bands_rejects_bad_links():
    """Host-side validation (no kernels run)."""
    import ctypes
    wall = small_cfg(16, 49)
    nb = band_rows(wall, 0.25)
    cfgs = configs.banded(wall, 8, 8, nb)
    L = _lib.load()
    arr = (_lib.LhConfig * 3)()
    from rnl_multiblock_nofft_amd.solver import _config
    for q, c in enumerate(cfgs):
        arr[q] = _config(c)
    h = ctypes.c_void_p()
    rc = L.lh_ctx_create(arr, 3, ctypes.byref(h))
    if rc != 0:
        pytest.skip("no device: %s" % L.lh_last_error().decode())
    try:
        assert L.lh_set_bands(h, 0, 1, 1, nb) != 0
        assert L.lh_set_bands(h, 0, 1, 2, nb + 1) != 0
        assert L.lh_set_bands(h, 1, 0, 2, nb) != 0
        assert L.lh_set_bands(h, 0, 1, 2, nb) == 0
        assert L.lh_set_bands(h, 0, 1, 2, nb) != 0  # linked already
    finally:
        L.lh_ctx_destroy(h)
