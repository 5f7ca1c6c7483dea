"""Shared helpers for the parity tests (the oracle is the checker)."""
import numpy as np

from oracle import pyoracle as po
from rnl_multiblock_nofft_amd import initial

ORACLE_KEYS = {f for f, _ in po.Cfg._fields_}


def oracle_for(cfg):
    kw = {k: v for k, v in cfg.items() if k in ORACLE_KEYS}
    kw.setdefault("fixed_dt", 1)
    o = po.Oracle(**kw)
    if cfg.get("texture") is not None:
        o.set_texture(*cfg["texture"])
    return o


def relerr(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def small_cfg(N=8, ny=17, **kw):
    c = dict(N1=N, N2=N, Ngal1=3 * N // 2, Ngal2=3 * N // 2, ny=ny, alp=1.0, bet=2.0,
             Re=4200.0, grid_kind=1, gamma=2.0, dyq=10.0, ppp=5, bslpu1=0.0,
             bslpu3=0.0, flag_ctpress=0, mpgx=0.0, dt=1e-3)
    c.update(kw)
    return c


def init_fields(cfg, grid, seed=1234, amp=0.1):
    kmax = min(8, cfg["N1"] // 2 - 1, cfg["N2"] // 2 - 1)
    return initial.perturbed_poiseuille(cfg, grid, seed, amp=amp, kmax=kmax)


STAT_NAMES = ("Um", "U2m", "Vm", "V2m", "Wm", "W2m", "Pm", "P2m", "UVm", "UWm", "VWm", "wxm", "wx2m")


def profile_errors(prof, ref, cfg, istat=1):
    """Max error of each statistics profile (write_stats order, rows by
    reference index) relative to its natural magnitude: a second moment's
    own maximum; a first moment's sqrt(points * second moment); a cross
    moment's sqrt of the product of the two second moments' maxima.  (Mean
    profiles of v are zero up to rounding, so their own maxima are no
    scale.)"""
    q = {n: i for i, n in enumerate(STAT_NAMES)}
    n = cfg["Ngal1"] * cfg["Ngal2"] * istat
    nN = cfg["N1"] * cfg["N2"] * istat
    m2 = {k: max(np.max(np.abs(ref[q[k]])), 1e-300) for k in ("U2m", "V2m", "W2m", "P2m", "wx2m")}
    sc = dict(U2m=m2["U2m"], V2m=m2["V2m"], W2m=m2["W2m"], P2m=m2["P2m"], wx2m=m2["wx2m"],
              Um=np.sqrt(n * m2["U2m"]), Vm=np.sqrt(n * m2["V2m"]), Wm=np.sqrt(n * m2["W2m"]),
              Pm=np.sqrt(n * m2["P2m"]), wxm=np.sqrt(nN * m2["wx2m"]),
              UVm=np.sqrt(m2["U2m"] * m2["V2m"]), UWm=np.sqrt(m2["U2m"] * m2["W2m"]),
              VWm=np.sqrt(m2["V2m"] * m2["W2m"]))
    return {k: float(np.max(np.abs(prof[q[k]] - ref[q[k]])) / sc[k]) for k in STAT_NAMES}
