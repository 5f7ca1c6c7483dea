"""ctypes binding of the CPU FP64 oracle (oracle/liblhoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

Fields cross this binding in the reference's allocation layout
(`complex(8) f(jlim1:jlim2, ncol)`, y contiguous) as complex128 numpy arrays
of shape (ncol, nrows).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# $LHO_LIB selects another build of the same source (the pin tests load
# copies with planted mistakes this way, in child processes)
_LIB = os.environ.get("LHO_LIB") or os.path.join(_HERE, "liblhoracle.so")

U1, U2, U3, P, NU1, NU2, NU3, DU1, DU2, DU3, PSI, DIV = range(12)
FIELD_IDS = {"u1": U1, "u2": U2, "u3": U3, "p": P, "Nu1": NU1, "Nu2": NU2,
             "Nu3": NU3, "du1": DU1, "du2": DU2, "du3": DU3, "psi": PSI, "div": DIV}


class Cfg(ctypes.Structure):
    _fields_ = [("N1", ctypes.c_int), ("N2", ctypes.c_int),
                ("Ngal1", ctypes.c_int), ("Ngal2", ctypes.c_int),
                ("ny", ctypes.c_int),
                ("alp", ctypes.c_double), ("bet", ctypes.c_double),
                ("Re", ctypes.c_double),
                ("grid_kind", ctypes.c_int), ("dyq", ctypes.c_double),
                ("ppp", ctypes.c_int), ("gamma", ctypes.c_double),
                ("bslpu1", ctypes.c_double), ("bslpu3", ctypes.c_double),
                ("flag_ctpress", ctypes.c_int), ("mpgx", ctypes.c_double),
                ("fixed_dt", ctypes.c_int), ("dt", ctypes.c_double),
                ("CFL", ctypes.c_double), ("band_edge", ctypes.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.lho_create.restype = vp
        L.lho_create.argtypes = [ctypes.POINTER(Cfg)]
        L.lho_destroy.argtypes = [vp]
        L.lho_set_threads.argtypes = [ctypes.c_int]
        L.lho_set_threads.restype = ctypes.c_int
        L.lho_nrows_u.argtypes = [vp]
        L.lho_ncol.argtypes = [vp]
        L.lho_get_grid.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.lho_set_grid.argtypes = [vp, dp, dp, dp, dp]
        L.lho_get_scalars.argtypes = [vp, dp]
        L.lho_set_scalars.argtypes = [vp, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int]
        L.lho_field_rows.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int)]
        L.lho_set_field.argtypes = [vp, ctypes.c_int, dp]
        L.lho_get_field.argtypes = [vp, ctypes.c_int, dp]
        for name in ("lho_rhs0", "lho_nonlinear", "lho_solve_uw", "lho_solve_v",
                     "lho_meanflow_ctU", "lho_solve_p", "lho_v_corr", "lho_substep",
                     "lho_step"):
            getattr(L, name).argtypes = [vp, ctypes.c_int]
        L.lho_init_flowrate.argtypes = [vp]
        L.lho_divergence.argtypes = [vp]
        L.lho_four_to_phys_plane.argtypes = [vp, dp]
        L.lho_phys_to_four_plane.argtypes = [vp, dp]
        L.lho_four_to_phys_N_plane.argtypes = [vp, dp]
        L.lho_phys_to_four_N_plane.argtypes = [vp, dp]
        L.lho_cft.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int]
        L.lho_rft.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.lho_lusol0.argtypes = [vp, ctypes.c_int, dp]
        L.lho_lusol_helm.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.lho_lusolp.argtypes = [vp, dp]
        ip_ = ctypes.POINTER(ctypes.c_int)
        L.lho_set_texture.argtypes = [vp, ip_, ip_]
        L.lho_set_texture.restype = ctypes.c_int
        L.lho_lusol_uw_tex.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp]
        L.lho_stats.argtypes = [vp, dp, dp]
        L.lho_slstats.argtypes = [vp, dp]
        L.lho_set_triads.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.lho_set_triads.restype = ctypes.c_int
        L.lhob_create.restype = vp
        L.lhob_create.argtypes = [ctypes.POINTER(Cfg), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int]
        L.lhob_destroy.argtypes = [vp]
        L.lhob_band.restype = vp
        L.lhob_band.argtypes = [vp, ctypes.c_int]
        L.lhob_row0.argtypes = [vp, ctypes.c_int]
        L.lhob_owns.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.lhob_init_flowrate.argtypes = [vp]
        for name in ("lhob_substep", "lhob_step", "lhob_nonlinear"):
            getattr(L, name).argtypes = [vp, ctypes.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def cft(c, n, ise, isa, m, iopt):
    """FFTPACK cft semantics on a float64 array (modified in place)."""
    assert c.dtype == np.float64 and c.flags.c_contiguous
    lib().lho_cft(_dp(c), n, ise, isa, m, iopt)


def rft(c, n, isa, m, iopt):
    assert c.dtype == np.float64 and c.flags.c_contiguous
    lib().lho_rft(_dp(c), n, isa, m, iopt)


def set_threads(n):
    return lib().lho_set_threads(int(n))


class Oracle:
    """One channel on the CPU.  `cfg` is a dict of lho_cfg fields."""

    def __init__(self, **cfg):
        c = Cfg()
        defaults = dict(grid_kind=0, dyq=10.0, ppp=5, gamma=2.0, bslpu1=0.0,
                        bslpu3=0.0, flag_ctpress=0, mpgx=0.0, fixed_dt=1,
                        dt=1e-3, CFL=0.7, alp=1.0, bet=2.0)
        defaults.update(cfg)
        if "Ngal1" not in defaults:
            defaults["Ngal1"] = 3 * defaults["N1"] // 2
        if "Ngal2" not in defaults:
            defaults["Ngal2"] = 3 * defaults["N2"] // 2
        for k, v in defaults.items():
            setattr(c, k, v)
        self.cfg = defaults
        self._c = c
        self.h = lib().lho_create(ctypes.byref(c))
        self.ny = c.ny
        self.nyu = c.ny + 1
        self.ncol = (c.N1 // 2 + 1) * c.N2

    def __del__(self):
        try:
            if self.h:
                lib().lho_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # grid
    def grid(self):
        nyu, nyv = self.nyu, self.ny
        yu = np.zeros(nyu); yv = np.zeros(nyv)
        du = np.zeros(nyu); dv = np.zeros(nyv)
        a = np.zeros(3 * nyu); b = np.zeros(3 * nyv)
        lib().lho_get_grid(self.h, _dp(yu), _dp(yv), _dp(du), _dp(dv), _dp(a), _dp(b))
        return dict(yu=yu, yv=yv, dthdyu=du, dthdyv=dv,
                    dyu2i=a.reshape(nyu, 3), dyv2i=b.reshape(nyv, 3))

    def set_grid(self, yu, yv, dthdyu, dthdyv):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (yu, yv, dthdyu, dthdyv)]
        lib().lho_set_grid(self.h, *[_dp(x) for x in arrs])

    def scalars(self):
        o = np.zeros(8)
        lib().lho_get_scalars(self.h, _dp(o))
        return dict(t=o[0], dt=o[1], mpgx=o[2], QxT=o[3], Qx=o[4], dtv=o[5], dtc=o[6],
                    iter=int(o[7]))

    def set_scalars(self, t=0.0, mpgx=None, QxT=None, iter=0):
        s = self.scalars()
        lib().lho_set_scalars(self.h, t, s["mpgx"] if mpgx is None else mpgx,
                              s["QxT"] if QxT is None else QxT, iter)

    def rows(self, name):
        lo = ctypes.c_int(); hi = ctypes.c_int()
        lib().lho_field_rows(self.h, FIELD_IDS[name], ctypes.byref(lo), ctypes.byref(hi))
        return lo.value, hi.value

    def get(self, name):
        lo, hi = self.rows(name)
        out = np.zeros((self.ncol, hi - lo + 1), dtype=np.complex128)
        lib().lho_get_field(self.h, FIELD_IDS[name], _dp(out.view(np.float64)))
        return out

    def set(self, name, arr):
        lo, hi = self.rows(name)
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        assert a.shape == (self.ncol, hi - lo + 1), (a.shape, self.ncol, hi - lo + 1)
        lib().lho_set_field(self.h, FIELD_IDS[name], _dp(a.view(np.float64)))

    def init_flowrate(self):
        lib().lho_init_flowrate(self.h)

    def substep(self, kRK):
        lib().lho_substep(self.h, kRK)

    def step(self, n=1):
        lib().lho_step(self.h, n)

    def phase(self, name, kRK):
        getattr(lib(), "lho_" + name)(self.h, kRK)

    def divergence(self):
        lib().lho_divergence(self.h)
        return self.get("div")

    def four_to_phys(self, plane):
        p = np.ascontiguousarray(plane, dtype=np.float64).copy()
        lib().lho_four_to_phys_plane(self.h, _dp(p))
        return p

    def phys_to_four(self, plane):
        p = np.ascontiguousarray(plane, dtype=np.float64).copy()
        lib().lho_phys_to_four_plane(self.h, _dp(p))
        return p

    def four_to_phys_N(self, plane):
        p = np.ascontiguousarray(plane, dtype=np.float64).copy()
        lib().lho_four_to_phys_N_plane(self.h, _dp(p))
        return p

    def phys_to_four_N(self, plane):
        p = np.ascontiguousarray(plane, dtype=np.float64).copy()
        lib().lho_phys_to_four_N_plane(self.h, _dp(p))
        return p

    def lusol_helm(self, kRK, vgrid, rhs):
        r = np.ascontiguousarray(rhs, dtype=np.complex128)
        x = np.zeros_like(r)
        lib().lho_lusol_helm(self.h, kRK, int(vgrid), _dp(r.view(np.float64)), _dp(x.view(np.float64)))
        return x

    def set_texture(self, bottom=None, top=None):
        """Wall masks on the N1 x N2 grid, shape (N2, N1), planeBC codes
        (1 no-slip, 2 free-shear); None leaves a wall smooth."""
        keep = []

        def arg(m):
            if m is None:
                return None
            a = np.ascontiguousarray(m, dtype=np.int32)
            assert a.shape == (self.cfg["N2"], self.cfg["N1"]), a.shape
            keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        rc = lib().lho_set_texture(self.h, arg(bottom), arg(top))
        if rc != 0:
            raise ValueError("texture masks must hold 1 (no-slip) or 2 (free-shear)")

    def lusol_uw_tex(self, kRK, rhs_u, rhs_w):
        ru = np.ascontiguousarray(rhs_u, dtype=np.complex128)
        rw = np.ascontiguousarray(rhs_w, dtype=np.complex128)
        u = np.zeros_like(ru)
        w = np.zeros_like(rw)
        lib().lho_lusol_uw_tex(self.h, kRK, _dp(ru.view(np.float64)), _dp(rw.view(np.float64)),
                               _dp(u.view(np.float64)), _dp(w.view(np.float64)))
        return u, w

    def lusolp(self, rhs):
        x = np.ascontiguousarray(rhs, dtype=np.complex128).copy()
        lib().lho_lusolp(self.h, _dp(x.view(np.float64)))
        return x

    def lusol0(self, kRK, x):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().lho_lusol0(self.h, kRK, _dp(x))
        return x

    def stats(self, prof=None, sp=None):
        """Add one statistics sample (stats.f90:1-107, spectra.f90:1-56) to
        prof (13, ny+1) and sp (5, ncol, ny+1); new zero arrays when None.
        Returns (prof, sp)."""
        nyu, ncol = self.nyu, self.ncol
        if prof is None:
            prof = np.zeros((13, nyu))
        if sp is None:
            sp = np.zeros((5, ncol, nyu))
        assert prof.shape == (13, nyu) and prof.dtype == np.float64 and prof.flags.c_contiguous
        assert sp.shape == (5, ncol, nyu) and sp.dtype == np.float64 and sp.flags.c_contiguous
        lib().lho_stats(self.h, _dp(prof), _dp(sp))
        return prof, sp

    def set_triads(self, files, map_u, map_v):
        """files: list of {type: int8 (len, 4)} (None: no triads); map_u /
        map_v: file index of u planes 1..ny-1 and v planes 1..ny-2 (-1 or a
        None file: no triads).  Switches lho_nonlinear to the list path."""
        lens, chunks = [], []
        for f in files:
            for t in range(1, 10):
                a = np.zeros((0, 4), np.int8) if f is None else np.asarray(f[t], np.int8).reshape(-1, 4)
                lens.append(a.shape[0])
                chunks.append(np.ascontiguousarray(a.T).reshape(-1))
        rows = np.concatenate(chunks) if chunks else np.zeros(0, np.int8)
        mu = np.array([m if m is not None and m >= 0 and files[m] is not None else -1 for m in map_u], np.int32)
        mv = np.array([m if m is not None and m >= 0 and files[m] is not None else -1 for m in map_v], np.int32)
        ln = np.array(lens, np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        rc = lib().lho_set_triads(self.h, len(files), ln.ctypes.data_as(ip), rows.ctypes.data,
                                  mu.ctypes.data_as(ip), mv.ctypes.data_as(ip))
        if rc != 0:
            raise ValueError("triad list entry outside the band")

    def slstats(self, maps=None):
        """Add one sl_stats sample (sl_stats2.f90:33-161) to maps
        (2, 4, N2, N1+2); a new zero array when None."""
        N1, N2 = self.cfg["N1"], self.cfg["N2"]
        if maps is None:
            maps = np.zeros((2, 4, N2, N1 + 2))
        assert maps.shape == (2, 4, N2, N1 + 2) and maps.flags.c_contiguous
        lib().lho_slstats(self.h, _dp(maps))
        return maps


def band_split(yu, h):
    """N(3,1) of ygrid (start.f90:999-1011): the first row j with
    yu(j) >= yu(0) + h, the last v row of the bottom band."""
    j = 0
    while yu[j] < yu[0] + h:
        j += 1
    return j


class _BandView(Oracle):
    """One band of a BandOracle: an Oracle over that band's rows."""

    def __init__(self, parent, q, h, cfg):
        self.parent = parent
        self.q = q
        self.h = h
        self.cfg = cfg
        self.ny = cfg["ny"]
        self.nyu = self.ny + 1
        self.ncol = (cfg["N1"] // 2 + 1) * cfg["N2"]

    def __del__(self):
        pass


class BandOracle:
    """A channel with multi-resolution bands on the CPU (the reference's
    three blocks): middle-band modes N1m x N2m over the whole height, the
    wall configuration's modes in the bottom and top bands (v rows 0..nb+1
    and ny-2-nb..ny-1).  band(q) is an Oracle view of band q (0 bottom,
    1 middle, 2 top) over its own rows; owns(q) flags its pencils (the wall
    bands' long-mode columns are not theirs and stay zero)."""

    def __init__(self, N1m, N2m, nb=None, band_h=0.25, **cfg):
        c = Cfg()
        defaults = dict(grid_kind=0, dyq=10.0, ppp=5, gamma=2.0, bslpu1=0.0,
                        bslpu3=0.0, flag_ctpress=0, mpgx=0.0, fixed_dt=1,
                        dt=1e-3, CFL=0.7, alp=1.0, bet=2.0, band_edge=0)
        defaults.update(cfg)
        if "Ngal1" not in defaults:
            defaults["Ngal1"] = 3 * defaults["N1"] // 2
        if "Ngal2" not in defaults:
            defaults["Ngal2"] = 3 * defaults["N2"] // 2
        for k, v in defaults.items():
            setattr(c, k, v)
        self.cfg = defaults
        if nb is None:
            tmp = Oracle(**defaults)
            nb = band_split(tmp.grid()["yu"], band_h)
            del tmp
        self.nb = nb
        self.N1m, self.N2m = N1m, N2m
        Ng1m, Ng2m = 3 * N1m // 2, 3 * N2m // 2
        self.h = lib().lhob_create(ctypes.byref(c), N1m, N2m, Ng1m, Ng2m, nb)
        if not self.h:
            raise ValueError("bands do not fit: nb=%d ny=%d" % (nb, defaults["ny"]))
        self.bands = []
        for q in range(3):
            bc = dict(defaults)
            if q == 1:
                bc.update(N1=N1m, N2=N2m, Ngal1=Ng1m, Ngal2=Ng2m)
            else:
                bc.update(ny=nb + 2, flag_ctpress=1, mpgx=0.0, band_edge=2 if q == 0 else 1)
            self.bands.append(_BandView(self, q, lib().lhob_band(self.h, q), bc))

    def __del__(self):
        try:
            if self.h:
                lib().lhob_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def band(self, q):
        return self.bands[q]

    def row0(self, q):
        return lib().lhob_row0(self.h, q)

    def owns(self, q):
        b = self.bands[q]
        return np.array([bool(lib().lhob_owns(self.h, q, c)) for c in range(b.ncol)])

    def init_flowrate(self):
        lib().lhob_init_flowrate(self.h)

    def substep(self, kRK):
        lib().lhob_substep(self.h, kRK)

    def step(self, n=1):
        lib().lhob_step(self.h, n)

    def nonlinear(self, kRK):
        lib().lhob_nonlinear(self.h, kRK)
