"""The Fortran host layer (rnl_multiblock_nofft_amd/fortran/): the
ISO_C_BINDING module covers the C-ABI, and the reference-named shims and the
main program agree with the FP64 oracle when run on the GPU.

The Fortran programs run as child processes and exchange data through files
(the record_out field format for the main program, raw FP64 streams for the
shim check)."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from helpers import init_fields, oracle_for, profile_errors, relerr, small_cfg
from rnl_multiblock_nofft_amd import fieldio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "rnl_multiblock_nofft_amd", "fortran")
BIN = os.path.join(FDIR, "bin")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_fortran_module_binds_every_c_entry_point():
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_abi import declared
    src = open(os.path.join(FDIR, "lhgpu.f90")).read()
    bound = sorted(set(re.findall(r"bind\(C,\s*name='(lh_\w+)'\)", src)))
    assert bound == declared()


def _nml(path, group, d):
    with open(path, "w") as f:
        f.write("&%s\n" % group)
        for k, v in d.items():
            if isinstance(v, str):
                f.write("  %s = '%s'\n" % (k, v))
            elif isinstance(v, (list, tuple)):
                f.write("  %s = %s\n" % (k, ", ".join("'%s'" % x for x in v)))
            elif isinstance(v, float):
                f.write("  %s = %s\n" % (k, repr(v).replace("e", "d") if "e" in repr(v) else repr(v) + "d0"))
            else:
                f.write("  %s = %d\n" % (k, v))
        f.write("/\n")


def _run(exe, nml, timeout=120):
    assert os.path.exists(exe), "%s not built (run __graft_entry__.build())" % exe
    say("run", os.path.basename(exe), nml)
    r = subprocess.run([exe, nml], capture_output=True, text=True, timeout=timeout)
    say(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.returncode
    return r.stdout


CFG_KEYS = ("N1", "N2", "ny", "alp", "bet", "Re", "grid_kind", "dyq", "ppp", "gamma",
            "flag_ctpress", "dt")


@pytest.mark.gpu
def test_fortran_main_program_matches_oracle(tmp_path):
    cfg = small_cfg(16, 33)
    nsteps = 2
    o = oracle_for(cfg)
    g = o.grid()
    f = init_fields(cfg, g, seed=77)
    N = np.zeros((4, 3), dtype=np.int32)
    N[0, :] = cfg["N1"]
    N[1, :] = cfg["N2"]
    fin, fout = [], []
    rows = {"u1": 0, "u2": 0, "u3": 0, "p": 1}
    for name in ("u1", "u2", "u3", "p"):
        p = str(tmp_path / ("%s_in.dat" % name))
        y = g["yv"] if name == "u2" else g["yu"]
        d = g["dthdyv"] if name == "u2" else g["dthdyu"]
        fieldio.write_field(p, f[name], rows[name], 2 if name == "u2" else 1, y, d,
                            Re=cfg["Re"], alp=cfg["alp"], bet=cfg["bet"], N=N)
        fin.append(p)
        fout.append(str(tmp_path / ("%s_out.dat" % name)))
        o.set(name, f[name])
    o.init_flowrate()
    o.step(nsteps)
    nml = str(tmp_path / "run.nml")
    d = {k: cfg[k] for k in CFG_KEYS}
    d.update(mpgx=0.0, precision=0, nsteps=nsteps, fin=fin, fout=fout)
    _nml(nml, "run", d)
    _run(os.path.join(BIN, "littleharsh_gpu"), nml)
    sc = o.scalars()
    for q, name in enumerate(("u1", "u2", "u3", "p")):
        r = fieldio.read_field(fout[q])
        assert r["iter"] == nsteps
        assert abs(r["t"] - sc["t"]) <= 1e-12
        assert r["jlo"] == rows[name]
        e = relerr(r["field"], o.get(name))
        say(name, e)
        assert e <= 1e-11, (name, e)


@pytest.mark.gpu
def test_fortran_shims_match_oracle(tmp_path):
    cfg = small_cfg(16, 33)
    o = oracle_for(cfg)
    rng = np.random.default_rng(5)
    Ng1, Ng2 = cfg["Ngal1"], cfg["Ngal2"]
    nplanes = 3
    planes = rng.standard_normal((nplanes, Ng2, Ng1 + 2))
    planes.tofile(str(tmp_path / "planes_in"))
    ncol = (cfg["N1"] // 2 + 1) * cfg["N2"]
    ny = cfg["ny"]
    kstage = 2
    ru = rng.standard_normal((ncol, ny + 1)) + 1j * rng.standard_normal((ncol, ny + 1))
    rw = rng.standard_normal((ncol, ny + 1)) + 1j * rng.standard_normal((ncol, ny + 1))
    rv = rng.standard_normal((ncol, ny)) + 1j * rng.standard_normal((ncol, ny))
    for nm, a in (("rhsu", ru), ("rhsw", rw), ("rhsv", rv)):
        a.astype(np.complex128).tofile(str(tmp_path / nm))
    planesN = rng.standard_normal((nplanes, cfg["N2"], cfg["N1"] + 2))
    planesN.tofile(str(tmp_path / "planesN_in"))
    rp = rng.standard_normal((ncol, ny - 1)) + 1j * rng.standard_normal((ncol, ny - 1))
    rp.tofile(str(tmp_path / "rhsp"))
    r0 = rng.standard_normal(ny + 1)
    r0.tofile(str(tmp_path / "rhs0"))
    cu = rng.standard_normal((ncol, ny + 1)) + 1j * rng.standard_normal((ncol, ny + 1))
    cu.tofile(str(tmp_path / "colsu"))
    plu = rng.standard_normal((ny + 1, Ng2, Ng1 + 2))
    plu.tofile(str(tmp_path / "plu"))
    nml = str(tmp_path / "check.nml")
    d = {k: cfg[k] for k in CFG_KEYS}
    d.update(nplanes=nplanes, kstage=kstage, dir=str(tmp_path))
    _nml(nml, "check", d)
    _run(os.path.join(BIN, "lh_shim_check"), nml)

    phys = np.fromfile(str(tmp_path / "planes_phys")).reshape(planes.shape)
    four = np.fromfile(str(tmp_path / "planes_four")).reshape(planes.shape)
    for q in range(nplanes):
        ref_p = o.four_to_phys(planes[q])
        e1 = relerr(phys[q], ref_p)
        e2 = relerr(four[q], o.phys_to_four(ref_p))
        say("plane", q, e1, e2)
        assert e1 <= 1e-13 and e2 <= 1e-13
    su = np.fromfile(str(tmp_path / "sol_u"), dtype=np.complex128).reshape(ncol, ny + 1)
    sw = np.fromfile(str(tmp_path / "sol_w"), dtype=np.complex128).reshape(ncol, ny + 1)
    sv = np.fromfile(str(tmp_path / "sol_v"), dtype=np.complex128).reshape(ncol, ny)
    eu = relerr(su, o.lusol_helm(kstage, 0, ru))
    ew = relerr(sw, o.lusol_helm(kstage, 0, rw))
    ev = relerr(sv, o.lusol_helm(kstage, 1, rv))
    say("lusol", eu, ew, ev)
    assert max(eu, ew, ev) <= 1e-13

    # N-grid transforms (four_to_phys_N / phys_to_four_N)
    N1, N2 = cfg["N1"], cfg["N2"]
    pN = np.fromfile(str(tmp_path / "planesN_phys")).reshape(planesN.shape)
    fN = np.fromfile(str(tmp_path / "planesN_four")).reshape(planesN.shape)
    for q in range(nplanes):
        ref_p = o.four_to_phys_N(planesN[q])
        e1 = relerr(pN[q][:, :N1], ref_p[:, :N1])
        e2 = relerr(fN[q], o.phys_to_four_N(ref_p))
        say("planeN", q, e1, e2)
        assert e1 <= 1e-13 and e2 <= 1e-13
    # LUsolP, LUsol0
    sp = np.fromfile(str(tmp_path / "sol_p"), dtype=np.complex128).reshape(ncol, ny - 1)
    ep = relerr(sp, o.lusolp(rp))
    s0 = np.fromfile(str(tmp_path / "sol_0"))
    e0 = relerr(s0, o.lusol0(kstage, r0.copy()))
    say("lusolp", ep, "lusol0", e0)
    assert ep <= 1e-13 and e0 <= 1e-14
    # transposes (modes_to_planes_UVP FOU3D.f90:1000-1098: mode (i, k) at
    # (2i+1, krow), krow = k or k + Ngal2 - N2 for the upper half)
    nxh = N1 // 2 + 1
    kr = np.array([k if k < N2 // 2 else k + Ng2 - N2 for k in range(N2)])
    exp_pl = np.zeros((ny + 1, Ng2, Ng1 + 2))
    cc = cu.T.reshape(ny + 1, N2, nxh)
    exp_pl[:, kr, 0:2 * nxh:2] = cc.real
    exp_pl[:, kr, 1:2 * nxh:2] = cc.imag
    plo = np.fromfile(str(tmp_path / "plu_out")).reshape(ny + 1, Ng2, Ng1 + 2)
    assert np.array_equal(plo, exp_pl)
    back = plu[:, kr, 0:2 * nxh:2] + 1j * plu[:, kr, 1:2 * nxh:2]     # (rows, N2, nxh)
    exp_c = back.reshape(ny + 1, ncol).T
    co = np.fromfile(str(tmp_path / "colsu_out"), dtype=np.complex128).reshape(ncol, ny + 1)
    assert np.array_equal(co, exp_c)
    cn = np.fromfile(str(tmp_path / "colsn_out"), dtype=np.complex128).reshape(ncol, ny + 1)
    exp_n = cu.copy()
    exp_n[:, 1:ny] = exp_c[:, 1:ny]
    assert np.array_equal(cn, exp_n)


@pytest.mark.gpu
def test_fortran_main_program_writes_stats(tmp_path):
    """nstat > 0: samples after every nstat steps, written in write_stats'
    layout (stats.f90:254-441), against the oracle's samples."""
    from rnl_multiblock_nofft_amd import statsio
    cfg = small_cfg(16, 33)
    nsteps, nstat = 5, 2
    o = oracle_for(cfg)
    g = o.grid()
    f = init_fields(cfg, g, seed=78)
    N = np.zeros((4, 3), dtype=np.int32)
    N[0, :] = cfg["N1"]
    N[1, :] = cfg["N2"]
    fin, fout = [], []
    rows = {"u1": 0, "u2": 0, "u3": 0, "p": 1}
    for name in ("u1", "u2", "u3", "p"):
        p = str(tmp_path / ("%s_in.dat" % name))
        y = g["yv"] if name == "u2" else g["yu"]
        d = g["dthdyv"] if name == "u2" else g["dthdyu"]
        fieldio.write_field(p, f[name], rows[name], 2 if name == "u2" else 1, y, d,
                            Re=cfg["Re"], alp=cfg["alp"], bet=cfg["bet"], N=N)
        fin.append(p)
        fout.append(str(tmp_path / ("%s_out.dat" % name)))
        o.set(name, f[name])
    o.init_flowrate()
    prof = None
    for _ in range(nsteps // nstat):
        o.step(nstat)
        prof, sp = o.stats() if prof is None else o.stats(prof, sp)
    o.step(nsteps % nstat)
    fst = str(tmp_path / "stats.dat")
    nml = str(tmp_path / "run.nml")
    d = {k: cfg[k] for k in CFG_KEYS}
    fh = str(tmp_path / "hist.dat")
    d.update(mpgx=0.0, precision=0, nsteps=nsteps, fin=fin, fout=fout, nstat=nstat, fstats=fst, fhist=fh)
    _nml(nml, "run", d)
    _run(os.path.join(BIN, "littleharsh_gpu"), nml)
    r = statsio.read_stats(fst)
    assert r["istat"] == nsteps // nstat and r["iter"] == nsteps and r["nband"] == 1
    assert np.allclose(r["yu"], g["yu"], rtol=0, atol=1e-15)
    ny = cfg["ny"]
    got = np.zeros_like(prof)
    for q, name in enumerate(statsio.PROFILES):
        j0, n = {"u": (0, ny + 1), "v": (0, ny), "p": (1, ny - 1)}[statsio.GRID_OF[name]]
        got[q, j0:j0 + n] = r[name]
    errs = profile_errors(got, prof, cfg, nsteps // nstat)
    say(errs)
    for name, e in errs.items():
        assert e <= 1e-11, (name, e)
    # history file: a record at the start and after every nstat steps
    h = fieldio.read_hist(fh)
    assert [r["iter"] for r in h["records"]] == [0, 2, 4]
    assert all(r["flag"] == 0 for r in h["records"])
    last = h["records"][-1]
    assert last["Umax"] > 0.5 and last["err"] < 1e-12 and last["dt"] == cfg["dt"]
