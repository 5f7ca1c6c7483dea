"""Learned closure: the numpy restatement (oracle/closure_oracle.py) pinned
on the CPU, and the device kernels (through the C-ABI) against it on the GPU.

Bars: the oracle's exact (float64) gradient equals central differences to
1e-6 and torch autograd to 1e-10; the device network (bf16 weights and
activations, fp32 accumulation) equals the oracle's bf16 emulation to 2e-2
relative L2 per tensor (accumulation order flips an occasional bf16
rounding); the sample pipeline (fp32 planes) equals the FP64 oracle to 1e-5;
a coarse step with the closure on equals the oracle's to 1e-2 of the
closure's own effect.
"""
import sys

import numpy as np
import pytest

from helpers import init_fields, oracle_for, relerr, small_cfg
from oracle import closure_oracle as co

gpu = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_params(seed, gain=1.0, out_gain=0.5, bias=0.1):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in co.SHAPES:
        if name.startswith("W"):
            d[name] = rng.standard_normal(shp) * gain / np.sqrt(shp[1]) * (out_gain if name == "W4" else 1.0)
        else:
            d[name] = bias * rng.standard_normal(shp)
    return co.pack(d)


# ---------------------------------------------------------------------------
# CPU: the oracle

def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 5000),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(co.bf16(x), ref)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert co.bf16(np.float32([1 + 2 ** -8]))[0] == 1.0
    assert co.bf16(np.float32([1 + 3 * 2 ** -8]))[0] == np.float32(1 + 2 ** -6)


def test_gelu_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.linspace(-8, 8, 2001)
    t = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(t, approximate="tanh")
    y.sum().backward()
    assert np.max(np.abs(co.gelu(x) - y.detach().numpy())) < 1e-14
    assert np.max(np.abs(co.dgelu(x) - t.grad.numpy())) < 1e-13
    assert co.gelu(0.0) == 0.0 and co.dgelu(0.0) == 0.5


def test_gradient_finite_differences():
    rng = np.random.default_rng(3)
    P = rand_params(1).astype(np.float64)
    X = rng.standard_normal((16, co.NIN))
    T = rng.standard_normal((16, co.NOUT))
    _, g = co.loss_grad(P, X, T, emulate=False)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in co.SHAPES])
    idx = np.concatenate([rng.integers(offs[q], offs[q + 1], 6) for q in range(len(co.SHAPES))])
    h = 1e-5
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (co.loss_grad(Pp, X, T, emulate=False)[0] - co.loss_grad(Pm, X, T, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()), (i, fd, g[i])


def test_gradient_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    P = rand_params(2).astype(np.float64)
    X = rng.standard_normal((32, co.NIN))
    T = rng.standard_normal((32, co.NOUT))
    loss, g = co.loss_grad(P, X, T, emulate=False)
    d = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in co.unpack(P).items()}
    h = torch.tensor(X)
    for L in (1, 2, 3):
        h = torch.nn.functional.gelu(h @ d["W%d" % L].T + d["b%d" % L], approximate="tanh")
    y = h @ d["W4"].T + d["b4"]
    lt = ((y - torch.tensor(T)) ** 2).mean()
    lt.backward()
    gt = co.pack({k: v.grad.numpy() for k, v in d.items()})
    assert abs(loss - lt.item()) <= 1e-14 * abs(loss)
    assert relerr(g, gt) <= 1e-10


def test_emulated_network_near_exact():
    """The bf16 emulation stays within bf16-sized distance of the exact
    network (a planted mistake in it -- say a dropped bias -- does not)."""
    rng = np.random.default_rng(5)
    P = rand_params(3)
    X = rng.standard_normal((256, co.NIN))
    T = rng.standard_normal((256, co.NOUT))
    ye, _ = co.forward(P, X, emulate=True)
    yx, _ = co.forward(P, X, emulate=False)
    assert relerr(ye, yx) < 3e-2
    _, ge = co.loss_grad(P, X, T, emulate=True)
    _, gx = co.loss_grad(P, X, T, emulate=False)
    assert relerr(ge, gx) < 5e-2
    bad = co.unpack(P.copy())
    bad["b2"][:] = 0
    yb, _ = co.forward(co.pack(bad), X, emulate=True)
    assert relerr(yb, yx) > 0.1


def test_n_grid_transforms_match_fftpack_oracle():
    """modes_to_phys / phys_to_modes against the C oracle's four_to_phys_N /
    phys_to_four_N (FFTPACK conventions, pinned by the DFT tests)."""
    cfg = small_cfg(16, 17)
    o = oracle_for(cfg)
    N1, N2 = 16, 16
    nxh = N1 // 2 + 1
    rng = np.random.default_rng(6)
    cols = rng.standard_normal(nxh * N2) + 1j * rng.standard_normal(nxh * N2)
    plane = np.zeros((N2, N1 + 2))
    plane[:, 0::2] = cols.reshape(N2, nxh).real
    plane[:, 1::2] = cols.reshape(N2, nxh).imag
    ref = o.four_to_phys_N(plane.ravel()).reshape(N2, N1 + 2)[:, :N1]
    assert np.max(np.abs(co.modes_to_phys(cols, N1, N2) - ref)) < 1e-12 * np.abs(ref).max()
    u = rng.standard_normal((N2, N1))
    p2 = np.zeros((N2, N1 + 2))
    p2[:, :N1] = u
    ref2 = o.phys_to_four_N(p2.ravel()).reshape(N2, N1 + 2)
    got = co.phys_to_modes(u, N1, N2).reshape(N2, nxh)
    assert np.max(np.abs(got.real - ref2[:, 0::2])) < 1e-14
    assert np.max(np.abs(got.imag - ref2[:, 1::2])) < 1e-14


def test_features_closed_forms():
    N1, N2, ny = 8, 8, 9
    nyu = ny + 1
    phys = np.zeros((3, nyu, N2, N1))
    j, z, x = np.meshgrid(np.arange(nyu), np.arange(N2), np.arange(N1), indexing="ij")
    phys[0] = 100 * j + 10 * z + x      # u1 encodes its own position
    phys[1] = 7.0                       # u2 constant
    phys[2] = -j                        # u3 = -row
    norm = np.array([0, 1, 0, 1, 0.5, 2, 1, 1, 1], np.float32)
    pts = np.array([0, 5 * N1 * N2 + 3 * N1 + 7, (ny - 2) * N1 * N2 + N1 * N2 - 1])
    X = co.features(phys, pts, norm)
    for n, p in enumerate(pts):
        jj, r = 1 + p // (N1 * N2), p % (N1 * N2)
        zz, xx = divmod(r, N1)
        for dy in range(-2, 3):
            for dz in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    f = (dy + 2) * 9 + (dz + 1) * 3 + dx + 1
                    jr = min(max(jj + dy, 0), nyu - 1)
                    assert X[n, f] == 100 * jr + 10 * ((zz + dz) % N2) + (xx + dx) % N1
                    assert X[n, 45 + f] == 0.5 * (7.0 - 1.0)
                    assert X[n, 90 + f] == -2.0 * min(max(jj + dy, 0), nyu - 1)


def test_box_filter_closed_form():
    cf, cc = small_cfg(16, 9), small_cfg(8, 9)
    nxf, nxc = 9, 5
    cols = np.zeros((nxf * 16, 3), np.complex128)
    cols[2 + 3 * nxf] = 1.0 + 2.0j            # (i=2, m=3)
    cols[1 + 13 * nxf] = -0.5j                # (i=1, m=-3)
    cols[6 + 1 * nxf] = 5.0                   # i=6 is outside the coarse band
    cols[0 + 4 * nxf] = 3.0                   # m=4: the coarse z-Nyquist (k = 4), cleared
    out = co.box_filter_restrict(cols, cf, cc)
    expect = np.zeros((nxc * 8, 3), np.complex128)
    s = lambda a: np.sin(np.pi * a) / (np.pi * a) if a else 1.0
    expect[2 + 3 * nxc] = s(2 / 8) * s(3 / 8) * (1.0 + 2.0j)
    expect[1 + 5 * nxc] = s(1 / 8) * s(-3 / 8) * (-0.5j)
    assert np.max(np.abs(out - expect)) < 1e-15


def _single_mode_fields(cfg, grid):
    """u1 = A cos(alp x) f(y) + U(y), v = w = 0 on the oracle's grid."""
    nxh = cfg["N1"] // 2 + 1
    ncol = nxh * cfg["N2"]
    yu = grid["yu"]
    u1 = np.zeros((ncol, cfg["ny"] + 1), np.complex128)
    u1[0] = 1 - yu ** 2
    u1[1] = 0.5 * 0.3 * (1 - yu ** 2) ** 2
    return dict(u1=u1, u2=np.zeros((ncol, cfg["ny"]), np.complex128),
                u3=np.zeros((ncol, cfg["ny"] + 1), np.complex128),
                p=np.zeros((ncol, cfg["ny"] - 1), np.complex128))


def test_sample_closed_form():
    """u1 = U(y) + A cos(x) f(y), v = w = 0: N1 = d/dx(u1^2) lives on x-modes
    1 and 2, and tau on mode 2 is (G2 - G1^2) N1(2) (mode 1 is exact:
    G1 N(1) = N_c(G u)(1) since U has G = 1), tau2 = tau3 = 0."""
    cf, cc = small_cfg(16, 17), small_cfg(8, 17)
    of, oc = oracle_for(cf), oracle_for(cc)
    f = _single_mode_fields(cf, of.grid())
    for k, v in f.items():
        of.set(k, v)
    phys, tau = co.sample(of, cf, oc, cc)
    of.phase("nonlinear", 2)
    N1f = of.get("Nu1")
    G1, G2 = np.sinc(1 / 8), np.sinc(2 / 8)
    nxc = 5
    tmodes = np.zeros((nxc * 8, 16), np.complex128)
    tmodes[2] = (G2 - G1 ** 2) * N1f[2]
    expect = np.zeros_like(tau[0])
    for jj in range(16):
        expect[jj + 1] = co.modes_to_phys(tmodes[:, jj], 8, 8)
    scale = np.abs(expect).max()
    assert scale > 1e-6
    assert np.max(np.abs(tau[0] - expect)) < 1e-13 * max(scale, np.abs(N1f).max())
    assert np.max(np.abs(tau[1])) < 1e-15 and np.max(np.abs(tau[2])) < 1e-15
    # the coarse state holds the filtered field
    assert np.max(np.abs(phys[0, :, 0, :] - (1 - oc.grid()["yu"][:, None] ** 2)
                         - 0.3 * G1 * (1 - oc.grid()["yu"][:, None] ** 2) ** 2
                         * np.cos(2 * np.pi * np.arange(8) / 8)[None, :])) < 1e-13


# ---------------------------------------------------------------------------
# GPU: the kernels through the C-ABI

def _trainer(batch, P, norm):
    from rnl_multiblock_nofft_amd import closure
    return closure.Trainer(batch, P, norm)


NORM1 = np.array([0, 0, 0, 1, 1, 1, 1, 1, 1], np.float32)


@gpu
def test_gpu_forward_and_gradient_match_oracle():
    rng = np.random.default_rng(11)
    B = 512
    P = rand_params(7).astype(np.float32)
    X = rng.standard_normal((B, co.NIN)).astype(np.float32)
    T = rng.standard_normal((B, co.NOUT)).astype(np.float32)
    tr = _trainer(B, P, NORM1)
    assert np.array_equal(tr.params(), P)
    tr.set_batch(X, T)
    y = tr.forward()
    ye, _ = co.forward(P, X, emulate=True)
    e = relerr(y, ye)
    say("forward", e)
    assert e < 2e-2
    loss = tr.loss_grad()
    lo, go = co.loss_grad(P, X, T, emulate=True)
    say("loss", loss, lo)
    assert abs(loss - lo) <= 2e-3 * lo
    g = tr.grad()
    for name, part in co.unpack(g).items():
        ref = co.unpack(go)[name]
        e = relerr(part, ref)
        say("grad", name, e)
        assert e < 2e-2, name


@gpu
def test_gpu_adam_step():
    """One AdamW step from the device gradient (first step: w - lr sign(g)
    for |g| >> eps, weight decay decoupled)."""
    rng = np.random.default_rng(12)
    B = 256
    P = rand_params(8).astype(np.float32)
    tr = _trainer(B, P, NORM1)
    tr.set_batch(rng.standard_normal((B, co.NIN)), rng.standard_normal((B, co.NOUT)))
    tr.loss_grad()
    g = tr.grad()
    lr, wd = 1e-3, 0.01
    tr.adam(lr, wd=wd, eps=1e-30)
    new = tr.params()
    expect = P - lr * (np.sign(g) + wd * P)
    sel = np.abs(g) > 0
    assert np.max(np.abs(new[sel] - expect[sel])) < 1e-6


def _coarse_pair(N=16, ny=17, seed=1234):
    from rnl_multiblock_nofft_amd import Solver
    cfg = small_cfg(N, ny)
    s = Solver([cfg])
    o = oracle_for(cfg)
    f = init_fields(cfg, s.grid(0), seed)
    for name in ("u1", "u2", "u3", "p"):
        s.upload(0, name, f[name])
        o.set(name, f[name])
    s.init_flowrate()
    o.init_flowrate()
    return cfg, s, o


@gpu
def test_gpu_planes_and_network_match_oracle():
    from rnl_multiblock_nofft_amd import closure
    cfg, s, o = _coarse_pair()
    N1, N2 = cfg["N1"], cfg["N2"]
    P = rand_params(9).astype(np.float32)
    norm = np.array([0.3, 0.0, 0.0, 2.0, 20.0, 20.0, 0.05, 0.02, 0.03], np.float32)
    closure.set_network(s, P, norm)
    phys, corr = closure.evaluate(s, 0)
    ref = co.phys_planes(o.get("u1"), o.get("u2"), o.get("u3"), N1, N2)
    e = np.max(np.abs(phys - ref)) / np.abs(ref).max()
    say("planes", e)
    assert e < 1e-6
    cref = co.network_correction(P, norm, phys.astype(np.float64), emulate=True)
    e = relerr(corr, cref)
    say("network on the planes", e)
    assert e < 2e-2
    # the fused network equals the GEMM path on the same device inputs
    B = 1024
    tr = _trainer(B, P, norm)
    ph = closure.DeviceArray(phys.shape)
    ph.set(phys)
    pts = np.random.default_rng(1).integers(0, (cfg["ny"] - 1) * N1 * N2, B)
    tr.set_points(cfg, ph.ptr, pts)
    y = tr.forward() * norm[6:9]
    j, r = 1 + pts // (N1 * N2), pts % (N1 * N2)
    got = corr[:, j, r // N1, r % N1].T
    e = relerr(got, y)
    say("fused vs GEMM", e)
    assert e < 1e-2


@gpu
def test_gpu_zero_network_is_the_plain_step():
    from rnl_multiblock_nofft_amd import closure
    cfg, s, o = _coarse_pair(seed=5)
    cfg2, s2, _ = _coarse_pair(seed=5)
    P = rand_params(10).astype(np.float32)
    d = co.unpack(P)
    d["W4"][:] = 0
    d["b4"][:] = 0
    closure.set_network(s, co.pack(d).astype(np.float32), NORM1)
    closure.enable(s, 0)
    s.step(2)
    s2.step(2)
    for name in ("u1", "u2", "u3", "p"):
        assert np.array_equal(s.download(0, name), s2.download(0, name)), name


@gpu
def test_gpu_closure_step_matches_oracle():
    from oracle import pyoracle as po
    from rnl_multiblock_nofft_amd import closure
    cfg, s, o = _coarse_pair(seed=21)
    o_plain = oracle_for(cfg)
    for name in ("u1", "u2", "u3", "p"):
        o_plain.set(name, o.get(name))
    o_plain.init_flowrate()
    P = rand_params(13).astype(np.float32)
    norm = np.array([0.3, 0.0, 0.0, 2.0, 20.0, 20.0, 0.5, 0.2, 0.3], np.float32)
    closure.set_network(s, P, norm)
    closure.enable(s, 0)
    s.step(1)
    for kRK in (1, 2, 3):
        co.substep_with_closure(o, kRK, P, norm, emulate=True)
    o_plain.step(1)
    for name in ("u1", "u2", "u3", "p"):
        g, r, r0 = s.download(0, name), o.get(name), o_plain.get(name)
        effect = np.linalg.norm(r - r0)
        e = np.linalg.norm(g - r) / effect
        say(name, "closure effect", effect / np.linalg.norm(r0), "gpu - oracle / effect", e)
        assert effect > 1e-8 * np.linalg.norm(r0)
        assert e < 1e-2, name


@gpu
def test_gpu_closure_once_per_step_matches_oracle():
    """Mode 2 (network at the first substep, correction held for the step)
    equals the oracle's held-correction substeps to 1e-2 of the closure's
    effect, over two steps (a graph replay)."""
    from rnl_multiblock_nofft_amd import closure
    cfg, s, o = _coarse_pair(seed=22)
    o_plain = oracle_for(cfg)
    for name in ("u1", "u2", "u3", "p"):
        o_plain.set(name, o.get(name))
    o_plain.init_flowrate()
    P = rand_params(14).astype(np.float32)
    norm = np.array([0.3, 0.0, 0.0, 2.0, 20.0, 20.0, 0.5, 0.2, 0.3], np.float32)
    closure.set_network(s, P, norm)
    closure.enable(s, 0, once_per_step=True)
    s.step(2)
    for step in range(2):
        held = {}
        for kRK in (1, 2, 3):
            co.substep_with_closure(o, kRK, P, norm, emulate=True, held=held)
    o_plain.step(2)
    for name in ("u1", "u2", "u3", "p"):
        g, r, r0 = s.download(0, name), o.get(name), o_plain.get(name)
        effect = np.linalg.norm(r - r0)
        e = np.linalg.norm(g - r) / effect
        say(name, "held closure effect", effect / np.linalg.norm(r0), "gpu - oracle / effect", e)
        assert effect > 1e-8 * np.linalg.norm(r0)
        assert e < 1e-2, name


@gpu
def test_gpu_sample_matches_oracle():
    from rnl_multiblock_nofft_amd import Solver, closure
    cf, cc = small_cfg(16, 17), small_cfg(8, 17)
    fine, coarse = Solver([cf]), Solver([cc])
    of, oc = oracle_for(cf), oracle_for(cc)
    f = init_fields(cf, fine.grid(0), 77)
    for name in ("u1", "u2", "u3", "p"):
        fine.upload(0, name, f[name])
        of.set(name, f[name])
    snaps = closure.Snapshots(cc, 2)
    snaps.add(fine, 0, coarse, 0)
    phys, tau = co.sample(of, cf, oc, cc)
    gp, gt = snaps.phys.get()[0], snaps.tau.get()[0]
    e1 = np.max(np.abs(gp - phys)) / np.abs(phys).max()
    e2 = np.max(np.abs(gt - tau)) / np.abs(tau).max()
    say("sample planes", e1, "tau", e2, "tau scale", np.abs(tau).max())
    assert e1 < 1e-6 and e2 < 1e-5
    # the coarse state now holds the filtered fields
    for name in ("u1", "u2", "u3", "p"):
        assert relerr(coarse.download(0, name), oc.get(name)) < 1e-14, name
    # batches drawn by the device sampler: loss against the host's reading of
    # the same points
    snaps.add(fine, 0, coarse, 0)
    norm = snaps.norm()
    B = 256
    P = rand_params(14).astype(np.float32)
    tr = closure.Trainer(B, P, norm)
    tr.sample(snaps, seed=99)
    loss = tr.loss_grad()
    allp = np.stack([snaps.phys.get()[q] for q in range(2)])
    allt = np.stack([snaps.tau.get()[q] for q in range(2)])
    plane = cc["N1"] * cc["N2"]
    npts = (cc["ny"] - 1) * plane
    idx = closure.batch_points(99, B, 2 * npts)
    X = np.zeros((B, co.NIN))
    T = np.zeros((B, co.NOUT))
    for n, h in enumerate(idx):
        q, p = divmod(int(h), npts)
        X[n] = co.features(allp[q].astype(np.float64), [p], norm)[0]
        j, r = 1 + p // plane, p % plane
        T[n] = allt[q][:, j, r // cc["N1"], r % cc["N1"]] / norm[6:9]
    lo, _ = co.loss_grad(P, X, T, emulate=True)
    say("sampled-batch loss", loss, lo)
    assert abs(loss - lo) <= 5e-3 * lo


@gpu
def test_gpu_training_gains_over_no_closure():
    """Samples from a 32 x 32 x 33 run for a 16 x 16 x 33 channel; a short
    training must explain a clear part of tau on snapshots it never saw
    (the zero correction scores 1)."""
    from rnl_multiblock_nofft_amd import Solver, closure
    cf, cc = small_cfg(32, 33), small_cfg(16, 33)
    fine, coarse = Solver([cf]), Solver([cc])
    f = init_fields(cf, fine.grid(0), 3, amp=0.3)
    for name in ("u1", "u2", "u3", "p"):
        fine.upload(0, name, f[name])
    fine.init_flowrate()
    nsnap = 24
    snaps = closure.Snapshots(cc, nsnap)
    for q in range(nsnap):
        fine.step(5)
        snaps.add(fine, 0, coarse, 0)
    norm = snaps.norm(max_snap=18)
    tr = closure.Trainer(2048, closure.init_params(0), norm)
    before = closure.apriori(tr, snaps, 18, 6)
    hist = closure.train(tr, snaps, 600, lr=2e-3, first=0, count=18, seed=5)
    after = closure.apriori(tr, snaps, 18, 6)
    say("loss first/last", hist[:5].mean(), hist[-20:].mean())
    say("held-out error per output before", before, "after", after)
    assert hist[-20:].mean() < 0.8 * hist[:5].mean()
    assert np.all(after < 0.8)
