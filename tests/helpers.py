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
