"""The oracle's right-hand side, projection, interpolation and y operators
against what the mathematics fixes (tests/oracle_pins.py), and proof that
each check has teeth: copies of the oracle with one planted mistake each
(a flipped sign, a dropped term, swapped weights) are compiled into a
temporary directory and must fail the check aimed at them.  CPU only."""
import json
import os
import shutil
import subprocess
import sys

import pytest

import oracle_pins

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "oracle", "lh_oracle.c")


@pytest.mark.parametrize("name", sorted(oracle_pins.PINS))
def test_pin_holds(name):
    err = oracle_pins.PINS[name]()
    print(name, err)
    assert err <= oracle_pins.BOUNDS[name], (name, err)


# (label, text in lh_oracle.c, its replacement, the pin that must fail)
MISTAKES = [
    ("ikx p sign", "cplx C3x = -s->dRK[q] * (I * s->kx[i]);",
     "cplx C3x = s->dRK[q] * (I * s->kx[i]);", "projection"),
    ("ikz p sign", "cplx C3z = -s->dRK[q] * (I * s->kz[k]);",
     "cplx C3z = s->dRK[q] * (I * s->kz[k]);", "projection"),
    ("D_y p sign", "double C3 = -s->dRK[q] * s->ddthetavi * s->dthdyv[j];",
     "double C3 = s->dRK[q] * s->ddthetavi * s->dthdyv[j];", "projection"),
    ("mean-mode pin row", "a[3 * ns + 2] = 0.0;", "a[3 * ns + 2] = a[3 * ns + 1];", "projection"),
    ("laplacian k^2 dropped (u)", "cplx L1 = a[0] * u1[j - 1] + (k2x + k2z + a[1]) * u1[j] + a[2] * u1[j + 1];",
     "cplx L1 = a[0] * u1[j - 1] + (a[1]) * u1[j] + a[2] * u1[j + 1];", "stokes_u1"),
    ("laplacian k^2 sign (w)", "cplx L3 = a[0] * u3[j - 1] + (k2x + k2z + a[1]) * u3[j] + a[2] * u3[j + 1];",
     "cplx L3 = a[0] * u3[j - 1] + (-k2x - k2z + a[1]) * u3[j] + a[2] * u3[j + 1];", "stokes_u3"),
    ("dyu2i lower/upper weights swapped",
     "a[3 * j + 0] = 2.0 / ((yu[j] - yu[j - 1]) * (yu[j + 1] - yu[j - 1]));",
     "a[3 * j + 0] = 2.0 / ((yu[j + 1] - yu[j]) * (yu[j + 1] - yu[j - 1]));", "d2_quadratic"),
    ("dyv2i centre coefficient", "b[3 * j + 1] = -2.0 / ((yv[j + 1] - yv[j]) * (yv[j] - yv[j - 1]));",
     "b[3 * j + 1] = -2.0 / ((yv[j + 1] - yv[j]) * (yv[j + 1] - yv[j - 1]));", "d2_quadratic"),
    ("interp_u weights swapped",
     "o1[j] = ((yv[j] - yu[j]) * u1[j + 1] + (yu[j + 1] - yv[j]) * u1[j]) / (yu[j + 1] - yu[j]);",
     "o1[j] = ((yu[j + 1] - yv[j]) * u1[j + 1] + (yv[j] - yu[j]) * u1[j]) / (yu[j + 1] - yu[j]);",
     "interp_linear"),
    ("interp_v as plain average",
     "o2[j] = ((yu[j] - yv[j - 1]) * u2[j] + (yv[j] - yu[j]) * u2[j - 1]) / (yv[j] - yv[j - 1]);",
     "o2[j] = 0.5 * (u2[j] + u2[j - 1]);", "interp_linear"),
    ("old nonlinear term sign", "double C1 = s->aRK[q] / s->c.Re, C2 = -s->cRK[q];",
     "double C1 = s->aRK[q] / s->c.Re, C2 = s->cRK[q];", "rk3_advection"),
    ("RK3 gamma_2", "const double g[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};",
     "const double g[3] = {8.0 / 15.0, 7.0 / 12.0, 3.0 / 4.0};", "rk3_advection"),
]


def _build(tmp, label, old, new):
    src = open(SRC).read()
    assert src.count(old) >= 1, "planted-mistake site not found (oracle changed?): " + label
    d = os.path.join(tmp, "m%d" % abs(hash(label)))
    os.makedirs(d)
    with open(os.path.join(d, "lh_oracle.c"), "w") as f:
        f.write(src.replace(old, new, 1))
    shutil.copy(os.path.join(ROOT, "oracle", "lh_oracle.h"), d)
    so = os.path.join(d, "liblhoracle.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-fcx-limited-range", "-std=gnu11", "-shared",
                    "-o", so, os.path.join(d, "lh_oracle.c"), "-lm"], check=True)
    return so


def test_planted_mistakes_fail_their_pin(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc absent")
    failed = {}
    for label, old, new, pin in MISTAKES:
        so = _build(str(tmp_path), label, old, new)
        env = dict(os.environ, LHO_LIB=so, OMP_NUM_THREADS="2")
        out = subprocess.run([sys.executable, os.path.join(HERE, "oracle_pins.py"), pin],
                             env=env, capture_output=True, text=True, check=True).stdout
        val = json.loads(out.strip().splitlines()[-1])[pin]
        print(label, "->", pin, val)
        caught = isinstance(val, str) or not (val <= oracle_pins.BOUNDS[pin])
        failed[label] = caught
    assert all(failed.values()), [k for k, v in failed.items() if not v]
