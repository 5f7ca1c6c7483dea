"""On-disk formats (record_out / read_in, FOU3D.f90:2856-3073 and
start.f90:2444-2882; the history file, start.f90:238-253 and
FOU3D.f90:3040-3070): byte-exact round trips of the reference's own files,
and the device's history record against the fields it reduces."""
import os

import numpy as np
import pytest

from helpers import init_fields, relerr, small_cfg
from rnl_multiblock_nofft_amd import fieldio

REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "input")), reason="reference fixtures absent")
@pytest.mark.parametrize("name", ["input/u1_0032x0032x0153_t00000.dat", "input/u2_0032x0032x0153_t00000.dat",
                                  "input/p_0032x0032x0153_t00000.dat", "output/u3_0032x0032x0153_t00052.dat"])
def test_field_file_round_trip_is_byte_exact(tmp_path, name):
    p = os.path.join(REF, name)
    r = fieldio.read_field(p)
    js = sorted(r["rows"])
    yp = np.zeros(max(js) + 2)
    for j in js:
        yp[j] = r["rows"][j][1]
    out = str(tmp_path / "f.dat")
    fieldio.write_field(out, r["field"], r["jlo"], r["rows"][js[0]][0], r["y"], r["dthdy"], t=r["t"], Re=r["Re"],
                        alp=r["alp"], bet=r["bet"], mpgx=r["mpgx"], iter=r["iter"], N=r["N"], yplane=yp,
                        dthetavi=r["dthetavi"])
    assert open(p, "rb").read() == open(out, "rb").read()


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "output")), reason="reference fixtures absent")
def test_history_file_round_trip_is_byte_exact(tmp_path):
    p = os.path.join(REF, "output", "hist_0032x0032x0153.dat")
    h = fieldio.read_hist(p)
    assert h["nband"] == 3 and len(h["records"]) > 2
    flag = h["records"][0]["flag"]
    out = str(tmp_path / "hist.dat")
    fieldio.write_hist(out, h["t"], h["Re"], h["alp"], h["bet"], h["mpgx"], h["N"], h["records"], flag)
    assert open(p, "rb").read() == open(out, "rb").read()
    # records every 5 iterations; the run's step is the viscous cap 2.5 dtv
    r1 = h["records"][1]
    assert r1["iter"] == 5 and r1["dt"] == pytest.approx(2.5 * r1["dtv"], rel=1e-15)


@pytest.mark.gpu
def test_history_record_matches_fields():
    from rnl_multiblock_nofft_amd import Solver
    cfg = small_cfg(16, 33)
    s = Solver([cfg])
    g = s.grid(0)
    for k, v in init_fields(cfg, g, seed=9).items():
        s.upload(0, k, v)
    s.init_flowrate()
    s.step(2)
    s.substep(1)
    s.substep(2)
    m0 = s.scalars(0)["mpgx"]
    s.substep(3)
    h = s.history(0)
    sc = s.scalars(0)
    div = s.download(0, "div")
    u1 = s.download(0, "u1")
    assert h["err"] == pytest.approx(np.abs(div).max(), rel=1e-14)
    yv, yu = g["yv"], g["yu"]
    Q = sum((yv[j] - yv[j - 1]) * u1[0, j].real for j in range(1, cfg["ny"]))
    assert h["Qx"] == pytest.approx(Q, rel=1e-14)
    assert h["Umax"] == u1[0].real.max()
    us = (u1[0, 1].real * (-1 - yu[0]) + u1[0, 0].real * (yu[1] + 1)) / (yu[1] - yu[0])
    assert h["Uslip"] == pytest.approx(us, rel=1e-12, abs=1e-15)
    assert h["dgx"] == pytest.approx(sc["mpgx"] - m0, rel=1e-9, abs=1e-18)
    assert h["t"] == sc["t"] and h["dt"] == sc["dt"] and h["mpgx"] == sc["mpgx"]
    rec = fieldio.hist_record(0, h)
    assert len(rec) == 8 + 8 + 11 * 8


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "input")), reason="reference fixtures absent")
@pytest.mark.parametrize("name", ["input/u1_0032x0032x0153_t00000.dat", "input/u2_0032x0032x0153_t00000.dat",
                                  "input/p_0032x0032x0153_t00000.dat"])
def test_fortran_field_round_trip_is_byte_exact(tmp_path, name):
    """read_in_field + record_out_field of the Fortran host (no device call)."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rnl_multiblock_nofft_amd",
                       "fortran", "bin", "lh_io_check")
    if not os.path.exists(exe):
        pytest.skip("Fortran programs not built")
    out = str(tmp_path / "f.dat")
    r = subprocess.run([exe, os.path.join(REF, name), out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert open(os.path.join(REF, name), "rb").read() == open(out, "rb").read()
