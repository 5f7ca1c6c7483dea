"""The reference-signature drop-in routines (rnl_multiblock_nofft_amd/
fortran/lh_dropin.f90), called by lh_dropin_check exactly as the reference's
call sites call them -- LUsolU_W(u,rhsu,w,rhsw,a,grid,iband,myid) on
type(cfield) band arrays, modes_to_planes_UVP(xPL,x,grid,myid,status,ierr),
cft(c,ise,isa,m,iopt,aux) after cfti, boundary with planeBC in the module
state, ... -- against the FP64 oracle on the same inputs."""
import os
import sys

import numpy as np
import pytest

from helpers import init_fields, oracle_for, relerr, small_cfg
from oracle import pyoracle as po
from rnl_multiblock_nofft_amd import textures

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_fortran_host import BIN, _nml, _run, say  # noqa: E402


def _rd(path, shape, dtype=np.float64):
    return np.fromfile(path, dtype=dtype).reshape(shape)


def test_dropin_routines_are_built_with_the_reference_names():
    """lh_dropin.f90 defines external subroutines with the reference's
    names and argument lists (checked on the source)."""
    import re
    src = open(os.path.join(os.path.dirname(BIN), "lh_dropin.f90")).read().lower()
    want = {
        "lusolu_w": "u, rhsu, w, rhsw, a, grid, iband, myid",
        "lusolv": "u, rhsu, a, grid, iband, myid",
        "lusolp": "x, myid, iband, nystart, nyend",
        "lusol0": "x, nystart, nyend",
        "lu_buildp": "nystart, nyend, myid, iband, a",
        "modes_to_planes_uvp": "xpl, x, grid, myid, status, ierr",
        "planes_to_modes_uvp": "x, xpl, grid, myid, status, ierr",
        "planes_to_modes_nuvp": "x, xpl, grid, myid, status, ierr",
        "four_to_phys_u": "u1, u2, u3, iband",
        "four_to_phys_du": "du, iband",
        "phys_to_four_du": "dupl, iband",
        "four_to_phys_n": "du, iband",
        "phys_to_four_n": "dupl, iband",
        "cft": "c, ise, isa, m, iopt, aux",
        "cfti": "ngal, wsac1",
        "rft": "c, isa, m, iopt, aux",
        "rfti": "mgal, wsaf1",
    }
    for name, args in want.items():
        m = re.search(r"^subroutine %s\(([^)]*)\)" % name, src, re.M)
        assert m and m.group(1) == args, (name, m.group(1) if m else None)
    assert re.search(r"^subroutine boundary$", src, re.M)


@pytest.mark.gpu
def test_dropin_routines_match_oracle(tmp_path):
    cfg = small_cfg(16, 33)
    N1, N2, ny = cfg["N1"], cfg["N2"], cfg["ny"]
    Ng1, Ng2 = cfg["Ngal1"], cfg["Ngal2"]
    nxh = N1 // 2 + 1
    ncol = nxh * N2
    kstage = 3
    rng = np.random.default_rng(11)
    cplx = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)  # noqa: E731
    ins = dict(rhsu=cplx(ncol, ny + 1), rhsw=cplx(ncol, ny + 1), rhsv=cplx(ncol, ny),
               rhsp=cplx(ncol, ny - 1), rhs0=rng.standard_normal(ny + 1), colsu=cplx(ncol, ny + 1),
               plu=rng.standard_normal((ny + 1, Ng2, Ng1 + 2)))
    for q in (1, 2, 3):
        ins["plane%d" % q] = rng.standard_normal((Ng2, Ng1 + 2))
    mask = textures.planebc(N1, N2, 2, 2, 2.0, 2.0)
    ins["planebc"] = mask.astype(np.int32)
    o = oracle_for(cfg)
    f = init_fields(cfg, o.grid(), seed=91)
    for nm in ("u1", "u2", "u3", "p"):
        ins["f_" + nm] = f[nm]
    for k, a in ins.items():
        a.tofile(str(tmp_path / k))
    nml = str(tmp_path / "dropin.nml")
    _nml(nml, "check", dict(N1=N1, N2=N2, ny=ny, alp=cfg["alp"], bet=cfg["bet"], Reb=cfg["Re"],
                            grid_kind=cfg["grid_kind"], gamma=cfg["gamma"], dtin=cfg["dt"], kstage=kstage,
                            dir=str(tmp_path)))
    _run(os.path.join(BIN, "lh_dropin_check"), nml)
    out = lambda n, s, t=np.float64: _rd(str(tmp_path / n), s, t)  # noqa: E731
    c128 = np.complex128

    # solves: LUsolU_W, LUsolV over the bands, LUsolP, LUsol0
    errs = {
        "LUsolU_W u": relerr(out("sol_u", (ncol, ny + 1), c128), o.lusol_helm(kstage, 0, ins["rhsu"])),
        "LUsolU_W w": relerr(out("sol_w", (ncol, ny + 1), c128), o.lusol_helm(kstage, 0, ins["rhsw"])),
        "LUsolV": relerr(out("sol_v", (ncol, ny), c128), o.lusol_helm(kstage, 1, ins["rhsv"])),
        "LUsolP": relerr(out("sol_p", (ncol, ny - 1), c128), o.lusolp(ins["rhsp"])),
        "LUsol0": relerr(out("sol_0", (ny + 1,)), o.lusol0(kstage, ins["rhs0"].copy())),
    }
    # transposes: modes (i, k) at (2i+1, krow) of plane j, krow = k, or
    # k + Ngal2 - N2 for the upper half (FOU3D.f90:1034-1040)
    kr = np.array([k if k < N2 // 2 else k + Ng2 - N2 for k in range(N2)])
    exp_pl = np.zeros((ny + 1, Ng2, Ng1 + 2))
    cc = ins["colsu"].T.reshape(ny + 1, N2, nxh)
    exp_pl[:, kr, 0:2 * nxh:2] = cc.real
    exp_pl[:, kr, 1:2 * nxh:2] = cc.imag
    assert np.array_equal(out("plu_out", exp_pl.shape), exp_pl)
    plu = ins["plu"]
    back = (plu[:, kr, 0:2 * nxh:2] + 1j * plu[:, kr, 1:2 * nxh:2]).reshape(ny + 1, ncol).T
    assert np.array_equal(out("colsu_out", (ncol, ny + 1), c128), back)
    exp_n = np.zeros((ncol, ny - 1), c128)
    exp_n[:] = back[:, 1:ny]
    assert np.array_equal(out("colsn_out", (ncol, ny - 1), c128), exp_n)
    # plane transforms and FFTPACK's cft / rft in the reference's sequence
    for q in (1, 2, 3):
        errs["four_to_phys_u %d" % q] = relerr(out("phys%d" % q, (Ng2, Ng1 + 2))[:, :Ng1],
                                               o.four_to_phys(ins["plane%d" % q])[:, :Ng1])
    ph1 = o.four_to_phys(ins["plane1"])
    errs["phys_to_four_du"] = relerr(out("four1", (Ng2, Ng1 + 2))[:, :N1 + 2], o.phys_to_four(ph1)[:, :N1 + 2])
    c = ins["plane1"].copy().ravel()
    steps = (("cft_inv", lambda a: po.cft(a, Ng2, Ng1 + 2, 2, nxh, 1)),
             ("rft_inv", lambda a: po.rft(a, Ng1, Ng1 + 2, Ng2, 1)),
             ("rft_fwd", lambda a: po.rft(a, Ng1, Ng1 + 2, Ng2, -1)),
             ("cft_fwd", lambda a: po.cft(a, Ng2, Ng1 + 2, 2, nxh, -1)))
    for name, fn in steps:
        fn(c)
        errs[name] = relerr(out(name, (Ng2 * (Ng1 + 2),)), c)
    # boundary (planeBC of the module state on both walls) + one RK3 step
    o.set_texture(mask, mask)
    for nm in ("u1", "u2", "u3", "p"):
        o.set(nm, f[nm])
    o.init_flowrate()
    o.step(1)
    rows = {"u1": ny + 1, "u2": ny, "u3": ny + 1, "p": ny - 1}
    for nm in ("u1", "u2", "u3", "p"):
        errs["boundary + step " + nm] = relerr(out("step_" + nm, (ncol, rows[nm]), c128), o.get(nm))
    for k, e in errs.items():
        say(k, e)
    bad = {k: e for k, e in errs.items() if not e <= (1e-12 if k.startswith("boundary") else 1e-13)}
    assert not bad, bad
