"""Channel configurations: the reference's input.in parameters plus the
knobs this build adds (fixed dt, grid kind, precision).

`Re` is the reference's `Re`, i.e. 1/nu (C1 = aRK/Re, littleharsh.f90:651).
A bulk Reynolds number Re_b = U_b h / nu with the laminar start u = 1 - y^2
(U_b = 2/3) maps to Re = 1.5 Re_b: Re_b = 2800 -> 4200, Re_b = 10000 -> 15000.
Sizes count Fourier modes, N = 2 (Ngal/3) (start.f90:112-115), so "32x32"
runs on a 48x48 physical grid.
"""
import copy

BASE = dict(alp=1.0, bet=2.0, grid_kind=1, gamma=2.0, dyq=10.0, ppp=5,
            bslpu1=0.0, bslpu3=0.0, flag_ctpress=0, mpgx=0.0, dt=1e-3)


def channel(N1, N2, ny, Re, **kw):
    c = dict(BASE)
    c.update(N1=N1, N2=N2, Ngal1=3 * N1 // 2, Ngal2=3 * N2 // 2, ny=ny, Re=Re)
    c.update(kw)
    return c


def re_code(re_bulk):
    return 1.5 * re_bulk


# The five workloads (smooth walls; see DESIGN.md for textures)
def smooth_32():
    return [channel(32, 32, 153, re_code(2800.0))]


def ridges_64():
    return [channel(64, 64, 153, re_code(2800.0))]


def ensemble_64():
    return [channel(32, 32, 153, re_code(2800.0)) for _ in range(48)] + \
           [channel(64, 64, 153, re_code(2800.0)) for _ in range(16)]


def large_256():
    return [channel(256, 256, 257, re_code(10000.0))]


def fine_128():
    return [channel(128, 128, 153, re_code(2800.0))]


WORKLOADS = {"smooth_32x32x153": smooth_32, "ridges_64x64x153": ridges_64,
             "ensemble_48x32_16x64": ensemble_64, "large_256x256x257": large_256,
             "fine_128x128x153": fine_128}


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
