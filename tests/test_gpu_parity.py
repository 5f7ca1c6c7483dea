"""GPU parity: liblhgpu (through the C-ABI) against the FP64 CPU oracle on the
same seeded fields.  Bar: relative L2 <= 1e-12 per field after one RK3 step
in FP64 (and per phase); mixed precision <= 1e-5."""
import sys

import numpy as np
import pytest

from helpers import init_fields, oracle_for, relerr, small_cfg
from rnl_multiblock_nofft_amd import Solver, _lib, configs, textures

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(cfg, seed=1234, precision="fp64"):
    s = Solver([cfg], precision=precision)
    o = oracle_for(cfg)
    g = s.grid(0)
    og = o.grid()
    for k in ("yu", "yv", "dthdyu", "dthdyv"):
        assert np.array_equal(g[k], og[k]), k
    f = init_fields(cfg, g, seed)
    for name in ("u1", "u2", "u3", "p"):
        s.upload(0, name, f[name])
        o.set(name, f[name])
    s.init_flowrate()
    o.init_flowrate()
    assert abs(s.scalars(0)["QxT"] - o.scalars()["QxT"]) <= 1e-15 * abs(o.scalars()["QxT"])
    return s, o, f


CASES = {
    "N8_ny17": small_cfg(8, 17),
    "N16_ny33_poly_ctP": small_cfg(16, 33, grid_kind=0, flag_ctpress=1, mpgx=-2.0 / 4200.0),
    "N32_ny153": configs.smooth_32()[0],
}


TALL = {"N16_ny257": small_cfg(16, 257)}
# the 384-point transforms of a 256x256 channel on a short y grid
SHORT = {"N256_ny9": small_cfg(256, 9)}


@pytest.mark.parametrize("name", list(CASES))
def test_phases_kRK(name):
    cfg = CASES[name]
    say("case", name)
    s, o, f = make_pair(cfg)
    dt = cfg["dt"]
    for kRK in (1, 2, 3):
        s.phase(_lib.PHASE_NONLINEAR, kRK)
        s.phase(_lib.PHASE_RHS, kRK)
        o.phase("rhs0", kRK)
        o.phase("nonlinear", kRK)
        for nm in ("Nu1", "Nu2", "Nu3"):
            e = relerr(s.download(0, nm), o.get(nm))
            say(name, kRK, nm, e)
            assert e <= TOL, (nm, e)
        for nm, un, dn in (("r1", "u1", "du1"), ("r2", "u2", "du2"), ("r3", "u3", "du3")):
            ref = o.get(un) + dt * o.get(dn)
            ref[:, 0] = 0
            ref[:, -1] = 0
            e = relerr(s.download(0, nm), ref)
            say(name, kRK, nm, e)
            assert e <= TOL, (nm, e)
        s.phase(_lib.PHASE_SOLVE, kRK)
        o.phase("solve_uw", kRK)
        o.phase("solve_v", kRK)
        if cfg["flag_ctpress"] == 0:
            o.phase("meanflow_ctU", kRK)
        o.phase("solve_p", kRK)
        o.phase("v_corr", kRK)
        for nm in ("u1", "u2", "u3", "p"):
            e = relerr(s.download(0, nm), o.get(nm))
            say(name, kRK, nm, e)
            # p accumulates psi = div/(dt(a+b)); the divergence of the
            # predicted velocity is a near-cancellation, so mid-step p carries
            # ~10x the velocity round-off.  The one-step bar below stays 1e-12.
            assert e <= (5 * TOL if nm == "p" else TOL), (nm, e)
        # psi is the small pressure correction (a near-cancellation): its
        # error is measured against the size of p
        e = np.linalg.norm(s.download(0, "psi") - o.get("psi")) / np.linalg.norm(o.get("p"))
        say(name, kRK, "psi/|p|", e)
        assert e <= 5 * TOL, ("psi", e)
        d = s.download(0, "div")
        assert np.abs(d).max() < 1e-10 * max(1.0, np.abs(s.download(0, "u1")).max())
        # mpgx accumulates g = (QxT - Qx0)/Qx1, a difference of two flow rates
        # that agree to O(dt): field round-off (1e-16) is amplified ~1e6 here
        assert abs(s.scalars(0)["mpgx"] - o.scalars()["mpgx"]) <= 1e-8 * abs(o.scalars()["mpgx"]) + 1e-18


@pytest.mark.parametrize("name", list(CASES))
def test_one_step_graph(name):
    cfg = CASES[name]
    say("case", name)
    s, o, f = make_pair(cfg, seed=99)
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say(name, nm, e)
        assert e <= TOL, (nm, e)
    # a second and third step replay the captured graph
    s.step(2)
    o.step(2)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say(name, "3 steps", nm, e)
        assert e <= 1e-11, (nm, e)


@pytest.mark.parametrize("cols", [-1, 3])
def test_step_psi_div_are_the_last_substeps(cols):
    """A captured step stores psi and div at its last substep only (the wave
    form skips them at substeps 1 and 2): after two steps they are the
    oracle's last-substep psi, a divergence at round-off, and the history's
    error is that divergence."""
    cfg = CASES["N32_ny153"]
    s, o, f = make_pair(cfg, seed=41)
    s.set_tuning(_lib.TUNE_SOLVE_COLS, cols)
    s.step(2)
    o.step(2)
    e = np.linalg.norm(s.download(0, "psi") - o.get("psi")) / np.linalg.norm(o.get("p"))
    say("step psi/|p|", cols, e)
    assert e <= 5e-11, e
    d = s.download(0, "div")
    assert np.abs(d).max() < 1e-10 * max(1.0, np.abs(s.download(0, "u1")).max())
    assert s.history(0)["err"] == pytest.approx(np.abs(d).max(), rel=1e-14)


@pytest.mark.parametrize("cols,scan,waves,case", [(0, 0, 1, "N32_ny153"), (1, 0, 1, "N32_ny153"),
                                                  (1, 1, 1, "N32_ny153"), (2, 1, 1, "N32_ny153"),
                                                  (3, 0, 1, "N32_ny153"), (3, 1, 1, "N32_ny153"),
                                                  (3, 1, 4, "N32_ny153"), (4, 0, 2, "N32_ny153"),
                                                  (1, 1, 1, "N8_ny17"), (2, 1, 4, "N8_ny17"),
                                                  (10, 0, 1, "N16_ny33_poly_ctP"), (10, 1, 4, "N16_ny33_poly_ctP"),
                                                  (-1, 0, 1, "N32_ny153"), (-1, 0, 1, "N8_ny17"),
                                                  (-1, 0, 4, "N32_ny153"), (-1, 0, 4, "N8_ny17"),
                                                  (-1, 0, 1, "N16_ny33_poly_ctP"), (-1, 0, 1, "N16_ny257"),
                                                  (0, 0, 1, "N16_ny257")])
def test_column_solve_layouts(cols, scan, waves, case):
    """Every column-solve form (one lane per column from HBM; C columns per
    workgroup staged in LDS, with each recurrence on one lane or split over a
    lane group; one column per wave, its rows in registers) gives the
    oracle's step.  At 8 x 8 x 17 the Poisson sweeps have fewer rows than
    lanes (empty chunks); 16 x 16 x 257 puts five rows on each lane of the
    wave form."""
    cfg = CASES[case] if case in CASES else TALL[case]
    say("solve layout", cols, scan, waves, case)
    s, o, f = make_pair(cfg, seed=7)
    s.set_tuning(_lib.TUNE_SOLVE_COLS, cols)
    s.set_tuning(_lib.TUNE_SOLVE_SCAN, scan)
    s.set_tuning(_lib.TUNE_SOLVE_WAVES, waves)
    assert s.get_tuning(_lib.TUNE_SOLVE_COLS) == cols
    assert s.get_tuning(_lib.TUNE_SOLVE_SCAN) == scan
    assert s.get_tuning(_lib.TUNE_SOLVE_WAVES) == waves
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say("cols", cols, "scan", scan, "waves", waves, nm, e)
        assert e <= TOL, (nm, e)


@pytest.mark.parametrize("cols,scan,waves,case", [(0, 0, 1, "N32_ny153"), (3, 1, 1, "N32_ny153"),
                                                  (2, 1, 4, "N8_ny17"), (10, 1, 4, "N16_ny33_poly_ctP"),
                                                  (-1, 0, 1, "N32_ny153"), (-1, 0, 1, "N16_ny33_poly_ctP"),
                                                  (-1, 0, 4, "N32_ny153"), (-1, 0, 4, "N16_ny33_poly_ctP")])
def test_rhs_fused_into_solve(cols, scan, waves, case):
    """Smooth-walled steps with the right-hand sides computed inside the
    column solve give the oracle's steps and, to round-off, the steps of the
    separate right-hand-side kernel (the same arithmetic per point)."""
    cfg = CASES[case]
    say("rhs fused", cols, scan, waves, case)
    out = {}
    for fused in (1, 0):
        s, o, f = make_pair(cfg, seed=11)
        for key, val in ((_lib.TUNE_SOLVE_COLS, cols), (_lib.TUNE_SOLVE_SCAN, scan),
                         (_lib.TUNE_SOLVE_WAVES, waves), (_lib.TUNE_RHS_FUSED, fused)):
            s.set_tuning(key, val)
        assert s.get_tuning(_lib.TUNE_RHS_FUSED) == fused
        s.step(2)
        out[fused] = {nm: s.download(0, nm) for nm in ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3")}
        if fused:
            o.step(2)
            for nm in ("u1", "u2", "u3", "p"):
                e = relerr(out[1][nm], o.get(nm))
                say("fused", nm, e)
                assert e <= 1e-11, (nm, e)
    for nm in out[1]:
        e = relerr(out[1][nm], out[0][nm])
        say("fused vs separate", nm, e)
        assert e <= 1e-13, (nm, e)


@pytest.mark.parametrize("walls", ["posts_stripes", "ridges", "posts_bottom"])
def test_rhs_fused_textured_wave_solve(walls):
    """Textured walls with the right-hand sides computed inside the
    one-column-per-wave solve: the near-wall rows the wall sources read are
    computed first, the solve reads the wall rows' sources and computes every
    other row itself.  Steps equal the oracle's and, to round-off, the
    separate right-hand-side kernel's (x-varying masks on both walls,
    x-independent masks, one textured wall)."""
    cfg = small_cfg(16, 33)
    masks = {"posts_stripes": (textures.posts(16, 16, 8, 4), textures.stripes(16, 16, 4, 0.5)),
             "ridges": (textures.stripes(16, 16, 4, 0.5), textures.stripes(16, 16, 4, 0.5)),
             "posts_bottom": (textures.posts(16, 16, 4, 2), None)}[walls]
    out = {}
    for fused in (1, 0):
        s, o = tex_pair(cfg, masks, seed=29)
        s.set_tuning(_lib.TUNE_SOLVE_COLS, -1)
        s.set_tuning(_lib.TUNE_RHS_FUSED, fused)
        assert s.get_tuning(_lib.TUNE_RHS_FUSED) == fused
        s.step(2)
        out[fused] = {nm: s.download(0, nm) for nm in ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3")}
        if fused:
            o.step(2)
            for nm in ("u1", "u2", "u3", "p"):
                e = relerr(out[1][nm], o.get(nm))
                say("fused textured", walls, nm, e)
                assert e <= 1e-11, (nm, e)
    for nm in out[1]:
        e = relerr(out[1][nm], out[0][nm])
        say("fused textured vs separate", walls, nm, e)
        assert e <= 1e-13, (nm, e)


@pytest.mark.parametrize("key,points", [(_lib.TUNE_ZINV_POINTS, 256), (_lib.TUNE_ZINV_POINTS, 4096),
                                        (_lib.TUNE_XPROD_POINTS, 128), (_lib.TUNE_XPROD_POINTS, 2048),
                                        (_lib.TUNE_ZFWD_POINTS, 100), (_lib.TUNE_ZFWD_POINTS, 3000)])
def test_stage_tiles(key, points):
    """Every tile size of the nonlinear stages (ragged last tiles included)
    gives the oracle's step."""
    cfg = CASES["N32_ny153"]
    s, o, f = make_pair(cfg, seed=9)
    s.set_tuning(key, points)
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say("tiles", key, points, nm, e)
        assert e <= TOL, (nm, e)


@pytest.mark.parametrize("stage,points,case", [("xprod", 384, "N32_ny153"), ("xprod", 768, "N32_ny153"),
                                               ("xprod", 2048, "N32_ny153"), ("xprod", 512, "N16_ny33_poly_ctP"),
                                               ("xprod", 256, "N8_ny17"), ("zfwd", 384, "N32_ny153"),
                                               ("zfwd", 100, "N32_ny153"), ("zfwd", 3000, "N32_ny153"),
                                               ("zfwd", 200, "N16_ny33_poly_ctP"), ("zinv", 384, "N32_ny153"),
                                               ("zinv", 4096, "N32_ny153"), ("zinv", 768, "N16_ny33_poly_ctP"),
                                               ("zinv", 256, "N8_ny17")])
def test_stage_inplace(stage, points, case):
    """The z-inverse, x-product and z-forward stages with their transforms
    in place in one LDS buffer (where every stage has at most one item per
    thread; otherwise the two-buffer path) give the oracle's step."""
    cfg = CASES[case]
    s, o, f = make_pair(cfg, seed=11)
    pk, ik = {"xprod": (_lib.TUNE_XPROD_POINTS, _lib.TUNE_XPROD_INPLACE),
              "zfwd": (_lib.TUNE_ZFWD_POINTS, _lib.TUNE_ZFWD_INPLACE),
              "zinv": (_lib.TUNE_ZINV_POINTS, _lib.TUNE_ZINV_INPLACE)}[stage]
    s.set_tuning(pk, points)
    s.set_tuning(ik, 1)
    assert s.get_tuning(ik) == 1
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say(stage, "in place", points, case, nm, e)
        assert e <= TOL, (nm, e)


@pytest.mark.parametrize("mode,points,case", [(1, 48, "N32_ny153"), (1, 240, "N32_ny153"), (1, 768, "N32_ny153"),
                                              (2, 240, "N32_ny153"), (2, 1536, "N32_ny153"),
                                              (1, 72, "N16_ny33_poly_ctP"), (1, 768, "N16_ny33_poly_ctP"),
                                              (2, 768, "N16_ny33_poly_ctP"), (1, 384, "N8_ny17")])
def test_zfwd_fieldwise(mode, points, case):
    """The z-forward stage transforming a plane's fields one after another,
    in workgroups of 256 (mode 1) or 512 (mode 2) threads, in tiles of 1, 3,
    5, 9 and 16 x-modes (ragged last tiles of 2 and 1 modes; at mode 2 and
    1536 points a 17-mode tile), gives the oracle's step; at Ngal2 = 12 (no
    compiled plan) the switch leaves the stage as it was."""
    cfg = CASES[case]
    say("zfwd fieldwise", mode, points, case)
    s, o, f = make_pair(cfg, seed=17)
    s.set_tuning(_lib.TUNE_ZFWD_FIELDWISE, mode)
    s.set_tuning(_lib.TUNE_ZFWD_POINTS, points)
    assert s.get_tuning(_lib.TUNE_ZFWD_FIELDWISE) == mode
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say("fieldwise", mode, points, case, nm, e)
        assert e <= TOL, (nm, e)


@pytest.mark.parametrize("mode", [1, 2])
def test_zfwd_fieldwise_mixed(mode):
    """The same in mixed precision, against the FP64 oracle at the mixed
    bar."""
    cfg = CASES["N32_ny153"]
    s, o, f = make_pair(cfg, seed=18, precision="mixed")
    s.set_tuning(_lib.TUNE_ZFWD_FIELDWISE, mode)
    s.set_tuning(_lib.TUNE_ZFWD_POINTS, 240)
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3"):
        e = relerr(s.download(0, nm), o.get(nm))
        say("fieldwise mixed", nm, e)
        assert e <= 1e-5, (nm, e)


@pytest.mark.parametrize("threads,ip,points,case", [(128, 0, 768, "N32_ny153"), (128, 1, 384, "N32_ny153"),
                                                    (128, 1, 768, "N32_ny153"), (128, 0, 2048, "N32_ny153"),
                                                    (192, 1, 576, "N32_ny153"), (192, 0, 768, "N32_ny153"),
                                                    (192, 1, 768, "N256_ny9"), (192, 0, 1536, "N256_ny9"),
                                                    (256, 1, 768, "N256_ny9")])
def test_xprod_threads(threads, ip, points, case):
    """The x-product stage in workgroups of 128, 192 or 256 threads
    (two-buffer and in-place forms; 48- and 384-point lengths) gives the
    oracle's step."""
    cfg = CASES[case] if case in CASES else SHORT[case]
    s, o, f = make_pair(cfg, seed=12)
    s.set_tuning(_lib.TUNE_XPROD_POINTS, points)
    s.set_tuning(_lib.TUNE_XPROD_THREADS, threads)
    s.set_tuning(_lib.TUNE_XPROD_INPLACE, ip)
    assert s.get_tuning(_lib.TUNE_XPROD_THREADS) == threads
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say("xprod threads", threads, ip, points, nm, e)
        assert e <= TOL, (nm, e)


def test_autotune_keeps_state():
    cfg = CASES["N32_ny153"]
    s, o, f = make_pair(cfg, seed=8)
    before = {nm: s.download(0, nm) for nm in ("u1", "u2", "u3", "p")}
    sc = s.scalars(0)
    tuned = s.autotune(3)
    say("autotune", tuned)
    for nm, a in before.items():
        assert np.array_equal(s.download(0, nm), a), nm
    assert s.scalars(0)["mpgx"] == sc["mpgx"]
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        assert relerr(s.download(0, nm), o.get(nm)) <= TOL, nm


def profiles(f, nxh):
    """Mean (mode (0,0), real part) and rms (all other modes; x-modes i > 0
    counted twice for their conjugates) of a field, per y row.  Column
    c = i + nxh k with nxh = N1/2+1."""
    ncol = f.shape[0]
    w = np.where(np.arange(ncol) % nxh == 0, 1.0, 2.0)
    w[0] = 0.0
    return f[0].real.copy(), np.sqrt((w[:, None] * np.abs(f) ** 2).sum(axis=0))


def test_100_steps_mean_and_rms_profiles():
    """The plain FP64 step against the oracle over 100 RK3 steps of the
    32x32x153 smooth channel: mean-velocity and rms profiles within 1e-6."""
    cfg = configs.smooth_32()[0]
    say("100 steps")
    s, o, f = make_pair(cfg, seed=2024)
    s.step(100)
    o.step(100)
    for nm in ("u1", "u2", "u3"):
        a, b = s.download(0, nm), o.get(nm)
        nxh = cfg["N1"] // 2 + 1
        ma, ra = profiles(a, nxh)
        mb, rb = profiles(b, nxh)
        # the mean of v vanishes by continuity: errors are measured against
        # the field's own size (largest mean or rms value)
        scale = max(np.abs(mb).max(), rb.max(), 1e-300)
        em = np.abs(ma - mb).max() / scale
        er = np.abs(ra - rb).max() / scale
        say("100 steps", nm, "mean", em, "rms", er, "field", relerr(a, b))
        assert em <= 1e-6 and er <= 1e-6, (nm, em, er)
    assert abs(s.scalars(0)["t"] - o.scalars()["t"]) <= 1e-12


def tex_pair(cfg, masks, seed=5, precision="fp64"):
    s, o, f = make_pair(cfg, seed=seed, precision=precision)
    s.set_texture(0, *masks)
    o.set_texture(*masks)
    return s, o


TEX_CASES = {
    "N16_ny33_stripes_posts": (small_cfg(16, 33), lambda: (textures.stripes(16, 16, 4, 0.5),
                                                           textures.posts(16, 16, 8, 4))),
    "N32_ny153_ridges_both": (configs.smooth_32()[0], lambda: (textures.stripes(32, 32, 8, 0.5),
                                                               textures.stripes(32, 32, 8, 0.5))),
    "N16_ny33_bottom_only_ctP": (small_cfg(16, 33, grid_kind=0, flag_ctpress=1, mpgx=-2.0 / 4200.0),
                                 lambda: (textures.posts(16, 16, 4, 2), None)),
}


@pytest.mark.parametrize("name", list(TEX_CASES))
def test_textured_step(name):
    cfg, mk = TEX_CASES[name]
    say("texture", name)
    s, o = tex_pair(cfg, mk())
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say(name, nm, e)
        assert e <= TOL, (nm, e)
    s.step(2)
    o.step(2)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say(name, "3 steps", nm, e)
        assert e <= 1e-11, (nm, e)


@pytest.mark.parametrize("fused", [1, 0])
def test_posts_wall_launch_forms(fused):
    """x-varying masks (posts on both walls, one wall stripes): the wall
    iteration with the planes resident in one workgroup's LDS (fused = 1)
    and as separate launches (z transforms, then the x transforms and the
    mask in one launch, the form of channels too large for LDS) both match
    the oracle, one and three steps."""
    cfg = small_cfg(16, 33)
    s, o = tex_pair(cfg, (textures.posts(16, 16, 8, 4), textures.stripes(16, 16, 4, 0.5)), seed=23)
    s.set_tuning(_lib.TUNE_TEX_FUSED, fused)
    assert s.get_tuning(_lib.TUNE_TEX_FUSED) == fused
    for n, tol in ((1, TOL), (2, 1e-11)):
        s.step(n)
        o.step(n)
        for nm in ("u1", "u2", "u3", "p"):
            e = relerr(s.download(0, nm), o.get(nm))
            say("posts wall launches", fused, n, nm, e)
            assert e <= tol, (nm, e)


@pytest.mark.parametrize("sep", [1, 0])
def test_ridges_wall_paths(sep):
    """x-independent masks: the per-x-mode wall iteration (sep = 1) and the
    two-dimensional wall transforms (sep = 0) both match the oracle."""
    cfg = small_cfg(32, 33)
    m = textures.stripes(32, 32, 4, 0.5)
    s, o = tex_pair(cfg, (m, m), seed=17)
    s.set_tuning(_lib.TUNE_TEX_SEPARABLE, sep)
    assert s.get_tuning(_lib.TUNE_TEX_SEPARABLE) == sep
    s.step(2)
    o.step(2)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say("ridges sep", sep, nm, e)
        assert e <= 1e-11, (nm, e)


@pytest.mark.parametrize("points,case", [(768, "N32_ny153"), (4096, "N32_ny153"), (768, "N16_ny33_poly_ctP"),
                                         (256, "N8_ny17")])
def test_zinv_512_threads(points, case):
    """The z-inverse stage in place in workgroups of 512 threads (tiles of
    16 and 17 x-modes at 48 points, which 256 threads cannot run in place;
    9 modes at 24; the two-buffer path at 12) gives the oracle's step."""
    cfg = CASES[case]
    say("zinv 512", points, case)
    s, o, f = make_pair(cfg, seed=21)
    s.set_tuning(_lib.TUNE_ZINV_INPLACE, 2)
    s.set_tuning(_lib.TUNE_ZINV_POINTS, points)
    assert s.get_tuning(_lib.TUNE_ZINV_INPLACE) == 2
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say("zinv 512", points, case, nm, e)
        assert e <= TOL, (nm, e)


@pytest.mark.parametrize("n,precision,zinv_ip", [(32, "fp64", 0), (64, "fp64", 0), (32, "mixed", 0),
                                                 (64, "fp64", 2)])
def test_ghost_pass_in_zinv(n, precision, zinv_ip):
    """Captured steps that leave the ridges' ghost-row pass to extra
    workgroups of the next z-inverse launch (within a step and across the
    steps of one call) give the oracle's steps and the separate launch's
    steps (the same arithmetic), also with the z-inverse stage in 512-thread
    workgroups.  In mixed precision the z-inverse stage is FP32 and the ghost
    pass FP64; the oracle bound there is the mixed mode's 1e-5."""
    cfg = small_cfg(n, 33)
    m = textures.stripes(n, n, 4, 0.5)
    tol = 1e-11 if precision == "fp64" else 1e-5
    out = {}
    for g in (1, 0):
        s, o = tex_pair(cfg, (m, m), seed=19, precision=precision)
        s.set_tuning(_lib.TUNE_ZINV_INPLACE, zinv_ip)
        s.set_tuning(_lib.TUNE_GHOST_IN_ZINV, g)
        assert s.get_tuning(_lib.TUNE_GHOST_IN_ZINV) == g
        s.step(3)
        out[g] = {nm: s.download(0, nm) for nm in ("u1", "u2", "u3", "p")}
        if g:
            o.step(3)
            for nm in out[g]:
                e = relerr(out[g][nm], o.get(nm))
                say("ghost in zinv", n, precision, nm, e)
                assert e <= tol, (nm, e)
    for nm in out[1]:
        e = relerr(out[1][nm], out[0][nm])
        say("ghost in zinv vs separate", n, nm, e)
        assert e <= 1e-14, (nm, e)


def test_texture_all_no_slip_is_smooth():
    cfg = CASES["N16_ny33_poly_ctP"]
    s, o, f = make_pair(cfg, seed=12)
    ns = np.ones((16, 16), np.int32)
    s.set_texture(0, ns, ns)
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        assert e <= TOL, (nm, e)


def test_ensemble_textured_and_smooth():
    cfgs = [small_cfg(16, 33), small_cfg(8, 17), small_cfg(16, 33), small_cfg(16, 25)]
    masks = [(textures.stripes(16, 16, 2, 0.5), None), None,
             (textures.posts(16, 16, 8, 4), textures.stripes(16, 16, 4, 0.25)), (None, textures.posts(16, 16, 4, 2))]
    s = Solver(cfgs)
    oracles = []
    for q, c in enumerate(cfgs):
        o = oracle_for(c)
        f = init_fields(c, s.grid(q), seed=300 + q)
        for name in ("u1", "u2", "u3", "p"):
            s.upload(q, name, f[name])
            o.set(name, f[name])
        if masks[q] is not None:
            s.set_texture(q, *masks[q])
            o.set_texture(*masks[q])
        o.init_flowrate()
        oracles.append(o)
    s.init_flowrate()
    s.step(2)
    for q, o in enumerate(oracles):
        o.step(2)
        for nm in ("u1", "u2", "u3", "p"):
            e = relerr(s.download(q, nm), o.get(nm))
            say("tex chan", q, nm, e)
            assert e <= 1e-11, (q, nm, e)


def test_texture_rejects_unsupported_wall_length():
    """Textured walls transform at N1, N2 (the N grid); a length the LDS
    transforms cannot do (20 = 2^2 5) is an error, not a partial transform."""
    cfg = dict(small_cfg(8, 17), N1=20, N2=20, Ngal1=32, Ngal2=32)
    s = Solver([cfg])
    m = np.ones((20, 20), np.int32)
    m[:, ::2] = 2
    with pytest.raises(_lib.LhError):
        s.set_texture(0, m, None)
    s.close()


def test_bad_texture_code_raises():
    s = Solver([small_cfg(8, 17)])
    m = np.full((8, 8), 3, np.int32)
    with pytest.raises(_lib.LhError):
        s.set_texture(0, m, None)


def test_dtc_matches_oracle():
    """dtc_calc on the device (CFL maxima of the first substep's physical
    planes) equals the oracle's, with dt fixed."""
    cfg = CASES["N32_ny153"]
    s, o, f = make_pair(cfg, seed=31)
    s.step(1)
    o.step(1)
    a, b = s.scalars(0), o.scalars()
    say("dtc", a["dtc"], b["dtc"], "t", a["t"], b["t"])
    assert abs(a["dtc"] - b["dtc"]) <= 1e-13 * b["dtc"]
    assert abs(a["t"] - b["t"]) <= 1e-15


@pytest.mark.parametrize("name", ["N16_ny33_poly_ctP", "N32_ny153"])
def test_adaptive_time_step(name):
    """Adaptive dt = min(2.5 dtv, CFL dtc) at every first substep, with the
    pivots and mean-mode solutions rebuilt on the device (FOU3D.f90:288-297)."""
    cfg = dict(CASES[name], adaptive_dt=1, CFL=0.5)
    say("adaptive", name)
    s = Solver([cfg])
    o = oracle_for(dict(cfg, fixed_dt=0))
    f = init_fields(cfg, s.grid(0), 41)
    for nm in ("u1", "u2", "u3", "p"):
        s.upload(0, nm, f[nm])
        o.set(nm, f[nm])
    s.init_flowrate()
    o.init_flowrate()
    for n in range(3):
        s.step(1)
        o.step(1)
        a, b = s.scalars(0), o.scalars()
        say("adaptive step", n, "dt", a["dt"], b["dt"], "t", a["t"], b["t"])
        assert abs(a["dt"] - b["dt"]) <= 1e-12 * b["dt"]
        assert abs(a["t"] - b["t"]) <= 1e-12 * b["t"]
        assert a["dt"] != cfg["dt"]
        for nm in ("u1", "u2", "u3", "p"):
            e = relerr(s.download(0, nm), o.get(nm))
            say(name, n, nm, e)
            assert e <= 1e-11, (nm, e)


@pytest.mark.parametrize("groups", [0, 1])
def test_ensemble_uneven(groups):
    """Channels of different sizes in one launch per stage (groups = 0) or
    one launch per plane shape (LH_TUNE_SHAPE_GROUPS = 1) each equal their
    own oracle over two steps."""
    cfgs = [small_cfg(8, 17), small_cfg(16, 33), small_cfg(8, 17, grid_kind=0), small_cfg(16, 25)]
    say("ensemble", len(cfgs), "groups", groups)
    s = Solver(cfgs)
    s.set_tuning(_lib.TUNE_SHAPE_GROUPS, groups)
    assert s.get_tuning(_lib.TUNE_SHAPE_GROUPS) == groups
    oracles = []
    for q, c in enumerate(cfgs):
        o = oracle_for(c)
        f = init_fields(c, s.grid(q), seed=1234 + q)
        for name in ("u1", "u2", "u3", "p"):
            s.upload(q, name, f[name])
            o.set(name, f[name])
        o.init_flowrate()
        oracles.append(o)
    s.init_flowrate()
    s.step(2)
    for q, o in enumerate(oracles):
        o.step(2)
        for nm in ("u1", "u2", "u3", "p"):
            e = relerr(s.download(q, nm), o.get(nm))
            say("chan", q, nm, e)
            assert e <= 1e-11, (q, nm, e)


@pytest.mark.parametrize("mode,points,zinv_ip", [(1, 768, 0), (2, 768, 1), (1, 96, 2)])
def test_zfwd_fieldwise_mixed_lengths(mode, points, zinv_ip):
    """One launch with channels of z lengths 24, 48 and 24 (each channel
    takes its own length's field-wise z-forward path, and its own length's
    z-inverse path, in place or not; 768 points: tiles of 9 and 16 x-modes,
    96: 4 and 2) gives each channel the oracle's step."""
    cfgs = [small_cfg(16, 33), configs.smooth_32()[0], small_cfg(16, 25, grid_kind=0)]
    say("fieldwise mixed lengths", mode, points, zinv_ip)
    s = Solver(cfgs)
    s.set_tuning(_lib.TUNE_ZFWD_FIELDWISE, mode)
    s.set_tuning(_lib.TUNE_ZFWD_POINTS, points)
    s.set_tuning(_lib.TUNE_ZINV_INPLACE, zinv_ip)
    oracles = []
    for q, c in enumerate(cfgs):
        o = oracle_for(c)
        f = init_fields(c, s.grid(q), seed=300 + q)
        for name in ("u1", "u2", "u3", "p"):
            s.upload(q, name, f[name])
            o.set(name, f[name])
        o.init_flowrate()
        oracles.append(o)
    s.init_flowrate()
    s.step(1)
    for q, o in enumerate(oracles):
        o.step(1)
        for nm in ("u1", "u2", "u3", "p"):
            e = relerr(s.download(q, nm), o.get(nm))
            say("fieldwise mixed lengths chan", q, nm, e)
            assert e <= TOL, (q, nm, e)


def test_mixed_precision_error():
    """Mixed precision on the smooth 32x32x153 channel: <= 1e-5 per field
    against the FP64 oracle after one step, and the bars the FP64 plane-mean
    terms make possible (velocities <= 1e-8, p <= 1e-6; with the means in the
    FP32 transforms the step gives u3 4.5e-9, p 6.0e-6)."""
    cfg = configs.smooth_32()[0]
    say("mixed")
    s, o, f = make_pair(cfg, precision="mixed")
    s.step(1)
    o.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        say("mixed", nm, e)
        assert e <= 1e-5, (nm, e)
        assert e <= (1e-6 if nm == "p" else 1e-8), (nm, e)


@pytest.mark.parametrize("knobs", [dict(xprod_ip=1, xprod_threads=256, zfwd_ip=0, xpts=768, zpts=384),
                                   dict(xprod_ip=1, xprod_threads=128, zfwd_ip=1, xpts=384, zpts=256),
                                   dict(xprod_ip=0, xprod_threads=128, zfwd_ip=1, xpts=1024, zpts=384)])
def test_mixed_precision_launch_forms(knobs):
    """Mixed precision with the in-place and 128-thread stage forms the
    autotuner can pick: same bound as the default forms against the FP64
    oracle, and the default forms' result to FP32 rounding (velocities 1e-7;
    p, whose mixed-precision error is itself ~2e-7, 1e-6)."""
    cfg = configs.smooth_32()[0]
    s, o, f = make_pair(cfg, precision="mixed")
    ref = Solver([cfg], precision="mixed")
    for name in ("u1", "u2", "u3", "p"):
        ref.upload(0, name, f[name])
    ref.init_flowrate()
    s.set_tuning(_lib.TUNE_XPROD_POINTS, knobs["xpts"])
    s.set_tuning(_lib.TUNE_ZFWD_POINTS, knobs["zpts"])
    s.set_tuning(_lib.TUNE_XPROD_THREADS, knobs["xprod_threads"])
    s.set_tuning(_lib.TUNE_XPROD_INPLACE, knobs["xprod_ip"])
    s.set_tuning(_lib.TUNE_ZFWD_INPLACE, knobs["zfwd_ip"])
    s.step(1)
    o.step(1)
    ref.step(1)
    for nm in ("u1", "u2", "u3", "p"):
        e = relerr(s.download(0, nm), o.get(nm))
        d = relerr(s.download(0, nm), ref.download(0, nm))
        say("mixed forms", knobs, nm, e, d)
        assert e <= 1e-5, (nm, e)
        assert d <= (1e-6 if nm == "p" else 1e-7), (nm, d)
    ref.close()


def test_bad_config_raises():
    cfg = small_cfg(8, 17)
    cfg["Ngal1"] = 10  # below the 3/2 rule
    with pytest.raises(_lib.LhError):
        Solver([cfg])
    cfg = small_cfg(8, 17, dt=0.0)
    with pytest.raises(_lib.LhError):
        Solver([cfg])
