"""Host-side driver of liblhgpu: an ensemble of channels on one GPU.

Mirrors the reference's main program (littleharsh.f90:58-230): setup
(`start`), initial fields (`getini`), the RK3 loop, and field access in the
reference's column layout.  All compute goes through the C-ABI of
liblhgpu.so; there is no CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import LhConfig, check


def _rows(cfg, name):
    ny = cfg["ny"]
    return {"u1": ny + 1, "u3": ny + 1, "u2": ny, "p": ny - 1, "Nu1": ny - 1, "Nu3": ny - 1,
            "Nu2": ny - 2, "r1": ny + 1, "r3": ny + 1, "r2": ny, "psi": ny - 1,
            "div": ny - 1}[name]


def _ncol(cfg):
    return (cfg["N1"] // 2 + 1) * cfg["N2"]


def _config(c, precision=_lib.PREC_FP64):
    s = LhConfig()
    for f, _ in LhConfig._fields_:
        if f == "precision":
            s.precision = precision
        elif f in c:
            setattr(s, f, c[f])
    return s


def band_rows(cfg, band_h):
    """nb = N(3,1) of ygrid's band split (start.f90:999-1011) on the grid of
    the channel configuration `cfg`: the last v row of the bottom band."""
    L = _lib.load()
    nb = ctypes.c_int()
    check(L.lh_band_rows(ctypes.byref(_config(cfg)), float(band_h), ctypes.byref(nb)))
    return nb.value


class Solver:
    """`cfgs`: list of channel dicts (see configs.channel).  precision:
    "fp64" (plain) or "mixed"."""

    def __init__(self, cfgs, precision="fp64"):
        L = _lib.load()
        self.cfgs = [dict(c) for c in cfgs]
        n = len(cfgs)
        arr = (LhConfig * n)()
        prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
        self.precision = precision
        for q, c in enumerate(cfgs):
            s = arr[q]
            for f, _ in LhConfig._fields_:
                if f == "precision":
                    s.precision = prec
                elif f in c:
                    setattr(s, f, c[f])
        h = ctypes.c_void_p()
        check(L.lh_ctx_create(arr, n, ctypes.byref(h)))
        self.h = h
        self.L = L
        # wall textures carried by the configurations: cfg["texture"] =
        # (bottom mask, top mask), see textures.py
        for q, c in enumerate(self.cfgs):
            if c.get("texture") is not None:
                self.set_texture(q, *c["texture"])
        # banded channels from configs.banded: three consecutive channels
        # marked (0, nb), (1, nb), (2, nb)
        q = 0
        while q + 2 < n:
            b = [self.cfgs[q + r].get("band") for r in range(3)]
            if b[0] is not None and b[0][0] == 0 and b[1] == (1, b[0][1]) and b[2] == (2, b[0][1]):
                self.set_bands(q, q + 1, q + 2, b[0][1])
                q += 3
            else:
                q += 1

    def close(self):
        if getattr(self, "h", None):
            self.L.lh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nchan(self):
        return len(self.cfgs)

    def grid(self, ch=0):
        c = self.cfgs[ch]
        ny = c["ny"]
        yu, yv = np.zeros(ny + 1), np.zeros(ny)
        du, dv = np.zeros(ny + 1), np.zeros(ny)
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(self.L.lh_get_grid(self.h, ch, dp(yu), dp(yv), dp(du), dp(dv)))
        return dict(yu=yu, yv=yv, dthdyu=du, dthdyv=dv)

    def set_bands(self, ch_bottom, ch_middle, ch_top, nb):
        """Link three channels into one banded channel (lh_set_bands)."""
        check(self.L.lh_set_bands(self.h, ch_bottom, ch_middle, ch_top, nb))
        self.bands = getattr(self, "bands", []) + [(ch_bottom, ch_middle, ch_top, nb)]

    def set_grid(self, ch, yu, yv, dthdyu, dthdyv):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (yu, yv, dthdyu, dthdyv)]
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(self.L.lh_set_grid(self.h, ch, *[dp(a) for a in arrs]))

    def upload(self, ch, name, arr):
        c = self.cfgs[ch]
        a = np.ascontiguousarray(arr, np.complex128)
        if a.shape != (_ncol(c), _rows(c, name)):
            raise ValueError("field %s: shape %s, expected %s" % (name, a.shape, (_ncol(c), _rows(c, name))))
        check(self.L.lh_field_upload(self.h, ch, _lib.FIELDS[name],
                                     a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def download(self, ch, name):
        c = self.cfgs[ch]
        out = np.zeros((_ncol(c), _rows(c, name)), np.complex128)
        check(self.L.lh_field_download(self.h, ch, _lib.FIELDS[name],
                                       out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set_fields(self, ch, fields):
        for k, v in fields.items():
            self.upload(ch, k, v)

    def scalars(self, ch=0):
        o = np.zeros(8)
        check(self.L.lh_get_scalars(self.h, ch, o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return dict(t=o[0], dt=o[1], mpgx=o[2], QxT=o[3], Qx=o[4], dtv=o[5], dtc=o[6], iter=int(o[7]))

    def history(self, ch=0):
        """The record_out history values (FOU3D.f90:3040-3070), reduced on
        the device."""
        o = np.zeros(12)
        check(self.L.lh_history(self.h, ch, o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        keys = ("t", "dtv", "dtc", "dt", "err", "Qx", "QxT", "mpgx", "dgx", "Umax", "Uslip", "iter")
        d = dict(zip(keys, o))
        d["iter"] = int(d["iter"])
        return d

    def set_scalars(self, ch, t=0.0, mpgx=None, QxT=None, iter=0):
        s = self.scalars(ch)
        check(self.L.lh_set_scalars(self.h, ch, float(t), float(s["mpgx"] if mpgx is None else mpgx),
                                    float(s["QxT"] if QxT is None else QxT), int(iter)))

    def init_flowrate(self):
        check(self.L.lh_init_flowrate(self.h))

    def substep(self, kRK):
        check(self.L.lh_substep(self.h, kRK))

    def step(self, n=1):
        check(self.L.lh_step(self.h, n))

    def phase(self, phase, kRK):
        check(self.L.lh_run_phase(self.h, phase, kRK))

    def time_kernels(self, nsub):
        """HIP-event times (ms, summed over nsub substeps) of the kernels
        [zinv, xprod, zfwd, rhs, solve, texture]."""
        ms = np.zeros(6)
        check(self.L.lh_time_kernels(self.h, int(nsub), ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return ms

    def set_texture(self, ch, bottom=None, top=None):
        """Wall masks of channel ch, arrays of shape (N2, N1) with the codes
        1 (no-slip) / 2 (free-shear); None leaves a wall smooth."""
        keep = []

        def arg(m):
            if m is None:
                return None
            a = np.ascontiguousarray(m, dtype=np.int32)
            c = self.cfgs[ch]
            if a.shape != (c["N2"], c["N1"]):
                raise ValueError("mask shape %s, expected (N2, N1) = %s" % (a.shape, (c["N2"], c["N1"])))
            keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        check(self.L.lh_set_texture(self.h, int(ch), arg(bottom), arg(top)))

    def set_triads(self, ch, files, map_u, map_v):
        """Triad lists for channel ch (the reference's restricted nonlinear
        path): files[q] a {type: int8 array (len, 4)} dict or None, map_u /
        map_v the file index of u planes 1..ny-1 and v planes 1..ny-2 (-1:
        none).  files = [] returns the channel to the FFT path."""
        lens, chunks = [], []
        for f in files:
            for t in range(1, 10):
                a = np.zeros((0, 4), np.int8) if f is None else np.asarray(f[t], np.int8).reshape(-1, 4)
                lens.append(a.shape[0])
                chunks.append(np.ascontiguousarray(a.T).reshape(-1))
        rows = np.concatenate(chunks) if chunks else np.zeros(1, np.int8)
        mu = np.array([m if m is not None and m >= 0 and files[m] is not None else -1 for m in map_u], np.int32)
        mv = np.array([m if m is not None and m >= 0 and files[m] is not None else -1 for m in map_v], np.int32)
        ln = np.array(lens if lens else [0], np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        check(self.L.lh_set_triads(self.h, int(ch), len(files), ln.ctypes.data_as(ip), rows.ctypes.data,
                                   mu.ctypes.data_as(ip), mv.ctypes.data_as(ip)))

    def set_tuning(self, key, value):
        check(self.L.lh_set_tuning(self.h, int(key), int(value)))

    def get_tuning(self, key):
        v = ctypes.c_int()
        check(self.L.lh_get_tuning(self.h, int(key), ctypes.byref(v)))
        return v.value

    def autotune(self, reps=10):
        """Time every feasible launch shape on this context's channels and
        keep the fastest; the state is left unchanged."""
        check(self.L.lh_autotune(self.h, int(reps)))
        return self.tuning()

    def tuning(self):
        return {"zinv_points": self.get_tuning(_lib.TUNE_ZINV_POINTS),
                "xprod_points": self.get_tuning(_lib.TUNE_XPROD_POINTS),
                "zfwd_points": self.get_tuning(_lib.TUNE_ZFWD_POINTS),
                "solve_cols": self.get_tuning(_lib.TUNE_SOLVE_COLS),
                "solve_scan": self.get_tuning(_lib.TUNE_SOLVE_SCAN),
                "solve_waves": self.get_tuning(_lib.TUNE_SOLVE_WAVES),
                "xprod_inplace": self.get_tuning(_lib.TUNE_XPROD_INPLACE),
                "zfwd_inplace": self.get_tuning(_lib.TUNE_ZFWD_INPLACE),
                "zinv_inplace": self.get_tuning(_lib.TUNE_ZINV_INPLACE),
                "zfwd_fieldwise": self.get_tuning(_lib.TUNE_ZFWD_FIELDWISE),
                "xprod_threads": self.get_tuning(_lib.TUNE_XPROD_THREADS),
                "rhs_fused": self.get_tuning(_lib.TUNE_RHS_FUSED),
                "ghost_in_zinv": self.get_tuning(_lib.TUNE_GHOST_IN_ZINV)}

    def sync(self):
        check(self.L.lh_sync(self.h))

    # statistics and spectra (stats.f90:1-107, spectra.f90:1-56)
    def stats_accumulate(self):
        """Add one sample of every channel's current fields."""
        check(self.L.lh_stats_accumulate(self.h))

    def stats_reset(self):
        check(self.L.lh_stats_reset(self.h))

    def stats(self, ch=0):
        """(prof, istat): prof (13, ny+1) in the order of statsio.PROFILES,
        rows by the reference's row index."""
        c = self.cfgs[ch]
        prof = np.zeros((_lib.ST_N, c["ny"] + 1))
        n = ctypes.c_int()
        check(self.L.lh_stats_get(self.h, ch, prof.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  ctypes.byref(n)))
        return prof, n.value

    def slip_stats(self, ch=0):
        """sl_stats wall maps, (2 walls, 4 quantities, N2, N1+2): u1, du1dy,
        u3, du3dy at the bottom and top wall faces (plane layout of spectral
        coefficients)."""
        c = self.cfgs[ch]
        out = np.zeros((2, 4, c["N2"], c["N1"] + 2))
        for w in range(2):
            check(self.L.lh_slip_stats_get(self.h, ch, w, out[w].ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def spectrum(self, ch, which):
        """Accumulated spectrum `which` in U, V, W, UV, P as (ncol, rows),
        the reference's sp(jlo:jhi, ncol) layout."""
        c = self.cfgs[ch]
        rows = {"U": c["ny"] + 1, "W": c["ny"] + 1, "UV": c["ny"] + 1, "V": c["ny"], "P": c["ny"] - 1}[which]
        out = np.zeros((_ncol(c), rows))
        check(self.L.lh_spectrum_get(self.h, ch, _lib.SP_IDS[which],
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out
