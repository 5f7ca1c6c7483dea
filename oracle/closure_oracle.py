"""CPU restatement of the learned closure (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product never
imports it.

The closure has no counterpart in the reference.  What it reuses of the
reference's conventions is restated here:
  * the N1 x N2 physical grid of four_to_phys_N / phys_to_four_N
    (FOU3D.f90:786-804, 2274-2293): backward transforms unscaled, forward
    ones scaled by 1/n, the real x transform ignoring the imaginary parts of
    modes 0 and N1/2 (FFTPACK rfftb) and producing none (rfftf);
  * the full nonlinear term, whose y-derivative parts follow der_yu_h /
    der_yv_h (FOU3D.f90:636-676) -- taken from the FP64 oracle (lho_nonlinear).

Everything else (inputs, network, box filter, sample definition) follows the
closure's own definition, stated in the docstrings.  The network functions
take `emulate=True` to round where the device rounds to bf16 (inputs,
weights, hidden activations, backward deltas), else they compute in float64
with no rounding, the form the finite-difference and autograd checks pin.
"""
import numpy as np

NIN, NH, NOUT = 135, 128, 3
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)), ("b3", (NH,)),
          ("W4", (NOUT, NH)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(vec):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(vec[o:o + n]).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name]).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as
    float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


C0, C1 = 0.7978845608028654, 0.044715


def gelu(x):
    """GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi)(x + 0.044715 x^3)))."""
    return 0.5 * x * (1.0 + np.tanh(C0 * (x + C1 * x ** 3)))


def dgelu(x):
    t = np.tanh(C0 * (x + C1 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C0 * (1.0 + 3.0 * C1 * x * x)


def forward(params, X, emulate=True):
    """Network outputs y[B][3] for inputs X[B][135] and the intermediates
    (pre-activations Z, activations H) for the backward pass."""
    P = unpack(params)
    if emulate:
        r = bf16
        H = r(np.asarray(X, np.float32))
        dt = np.float32
    else:
        r = lambda a: a
        H = np.asarray(X, np.float64)
        dt = np.float64
    Zs, Hs = [], [H]
    for L in (1, 2, 3):
        W = r(P["W%d" % L].astype(dt))
        Z = H @ W.T + P["b%d" % L].astype(dt)
        H = r(gelu(Z).astype(dt))
        Zs.append(Z)
        Hs.append(H)
    y = H @ r(P["W4"].astype(dt)).T + P["b4"].astype(dt)
    return y, (Zs, Hs)


def loss_grad(params, X, T, emulate=True):
    """loss = mean over points and outputs of (y - t)^2 and its gradient
    (parameter-vector layout)."""
    P = unpack(params)
    y, (Zs, Hs) = forward(params, X, emulate)
    B = X.shape[0]
    dt = np.float32 if emulate else np.float64
    r = bf16 if emulate else (lambda a: a)
    d = y - np.asarray(T, dt)
    loss = float(np.mean(d.astype(np.float64) ** 2))
    dY = r((2.0 * d / (3.0 * B)).astype(dt))
    G = {"W4": dY.T @ Hs[3], "b4": dY.sum(axis=0)}
    dH = dY @ r(P["W4"].astype(dt))
    for L in (3, 2, 1):
        dZ = r((dH * dgelu(Zs[L - 1])).astype(dt))
        G["W%d" % L] = dZ.T @ Hs[L - 1]
        G["b%d" % L] = dZ.sum(axis=0)
        if L > 1:
            dH = dZ @ r(P["W%d" % L].astype(dt))
    return loss, pack({k: np.asarray(v, np.float64) for k, v in G.items()})


# ---------------------------------------------------------------------------
# inputs on the N1 x N2 grid

def modes_to_phys(cols, N1, N2):
    """Columns (ncol,) with c = i + k*(N1/2+1) -> physical N2 x N1 values
    (four_to_phys_N conventions)."""
    nxh = N1 // 2 + 1
    c = np.asarray(cols).reshape(N2, nxh)  # [k][i]
    g = np.fft.ifft(c, axis=0) * N2        # backward z, unscaled
    return np.fft.irfft(g, n=N1, axis=1) * N1


def phys_to_modes(u, N1, N2):
    """Physical N2 x N1 values -> columns (phys_to_four_N conventions)."""
    x = np.fft.rfft(np.asarray(u, np.float64), axis=1) / N1
    x[:, 0] = x[:, 0].real
    x[:, N1 // 2] = x[:, N1 // 2].real
    c = np.fft.fft(x, axis=0) / N2
    return c.reshape(-1)


def phys_planes(u1, u2, u3, N1, N2):
    """Planes [3][ny+1][N2][N1] of fields in the oracle's column layout
    (ncol, rows): u1, u3 rows 0..ny, u2 rows 0..ny-1 (row ny zero)."""
    nyu = u1.shape[1]
    out = np.zeros((3, nyu, N2, N1))
    for f, a in enumerate((u1, u2, u3)):
        for j in range(a.shape[1]):
            out[f, j] = modes_to_phys(a[:, j], N1, N2)
    return out


def features(phys, pts, norm9):
    """Inputs X[len(pts)][135] of points pts (index (j-1)*N1*N2 + z*N1 + x,
    j = 1..ny-1): input f = field*45 + (dy+2)*9 + (dz+1)*3 + dx+1, the value of
    field (u1, u2, u3) at row j+dy clamped to the field's rows (0..ny for u1,
    u3, 0..ny-1 for u2), z+dz and x+dx periodic, normalised by norm9."""
    _, nyu, N2, N1 = phys.shape
    pts = np.asarray(pts)
    j = 1 + pts // (N1 * N2)
    r = pts % (N1 * N2)
    z, x = r // N1, r % N1
    X = np.zeros((len(pts), NIN))
    for f in range(NIN):
        fld, rr = divmod(f, 45)
        dy, r9 = rr // 9 - 2, rr % 9
        dz, dx = r9 // 3 - 1, r9 % 3 - 1
        jmax = nyu - 1 if fld != 1 else nyu - 2
        jj = np.clip(j + dy, 0, jmax)
        X[:, f] = (phys[fld, jj, (z + dz) % N2, (x + dx) % N1] - norm9[fld]) * norm9[3 + fld]
    return X


def correction_modes(corr, N1, N2, ny):
    """Correction planes [3][ny+1][N2][N1] -> the mode-space addition to
    Nu1, Nu2, Nu3 in the oracle's layout ((ncol, ny-1), (ncol, ny-2),
    (ncol, ny-1)); x-Nyquist and z-Nyquist modes untouched (zero)."""
    nxh = N1 // 2 + 1
    ncol = nxh * N2
    i = np.arange(ncol) % nxh
    k = np.arange(ncol) // nxh
    keep = (i != N1 // 2) & (k != N2 // 2)
    out = []
    for f, rows in ((0, ny - 1), (1, ny - 2), (2, ny - 1)):
        a = np.zeros((ncol, rows), np.complex128)
        for jj in range(rows):
            a[:, jj] = phys_to_modes(corr[f, jj + 1], N1, N2) * keep
        out.append(a)
    return out


def network_correction(params, norm9, phys, emulate=True):
    """The network's correction planes [3][ny+1][N2][N1] (u rows 1..ny-1;
    rows 0 and ny zero)."""
    _, nyu, N2, N1 = phys.shape
    npts = (nyu - 2) * N1 * N2
    X = features(phys, np.arange(npts), norm9)
    y, _ = forward(params, X, emulate)
    corr = np.zeros((3, nyu, N2, N1))
    for o in range(3):
        corr[o, 1:nyu - 1] = (y[:, o] * norm9[6 + o]).reshape(nyu - 2, N2, N1)
    return corr


# ---------------------------------------------------------------------------
# training samples

def box_filter_restrict(cols_f, cfg_f, cfg_c):
    """Fine columns (ncol_f, rows) -> coarse (ncol_c, rows): coarse column
    (i, k) takes fine mode (i, m), m the z wavenumber of k (k >= N2c/2 ->
    k - N2c), times sinc(pi i/N1c) sinc(pi m/N2c); the coarse z-Nyquist row
    k = N2c/2 is zero."""
    N1c, N2c, N1f, N2f = cfg_c["N1"], cfg_c["N2"], cfg_f["N1"], cfg_f["N2"]
    nxc, nxf = N1c // 2 + 1, N1f // 2 + 1
    out = np.zeros((nxc * N2c, cols_f.shape[1]), np.complex128)
    for c in range(nxc * N2c):
        i, k = c % nxc, c // nxc
        if k == N2c // 2:
            continue
        m = k if k < N2c // 2 else k - N2c
        G = np.sinc(i / N1c) * np.sinc(m / N2c)  # numpy sinc(x) = sin(pi x)/(pi x)
        out[c] = G * cols_f[i + (m % N2f) * nxf]
    return out


def clear_nyquist(a, N1, N2):
    nxh = N1 // 2 + 1
    c = np.arange(a.shape[0])
    a = a.copy()
    a[(c % nxh == N1 // 2) | (c // nxh == N2 // 2)] = 0
    return a


def sample(fine, cfg_f, coarse, cfg_c):
    """(phys, tau) planes of one training sample from oracle states `fine`
    and `coarse` (pyoracle.Oracle): the coarse state gets the filtered fine
    fields; tau = filtered full N of the fine state - full N of the coarse
    state, Nyquist modes cleared."""
    ny = cfg_c["ny"]
    for name in ("u1", "u2", "u3", "p"):
        coarse.set(name, box_filter_restrict(fine.get(name), cfg_f, cfg_c))
    fine.phase("nonlinear", 2)
    coarse.phase("nonlinear", 2)
    N1, N2 = cfg_c["N1"], cfg_c["N2"]
    taus = []
    for name in ("Nu1", "Nu2", "Nu3"):
        t = box_filter_restrict(fine.get(name), cfg_f, cfg_c) - coarse.get(name)
        taus.append(clear_nyquist(t, N1, N2))
    phys = phys_planes(coarse.get("u1"), coarse.get("u2"), coarse.get("u3"), N1, N2)
    tau = np.zeros_like(phys)
    for f, t in enumerate(taus):
        for jj in range(t.shape[1]):
            tau[f, jj + 1] = modes_to_phys(t[:, jj], N1, N2)
    return phys, tau


def substep_with_closure(o, kRK, params, norm9, emulate=True, held=None):
    """One oracle substep (lho_substep order) with the network's correction
    added to the full nonlinear term (which the next substep also sees as
    its old N) and to du.  With `held` (a dict), the correction is computed
    at kRK = 1 only, kept in held["add"] and reused at kRK = 2, 3 (the
    once-per-step mode)."""
    cfg = o.cfg
    N1, N2, ny = cfg["N1"], cfg["N2"], cfg["ny"]
    o.phase("rhs0", kRK)
    phys = phys_planes(o.get("u1"), o.get("u2"), o.get("u3"), N1, N2)
    o.phase("nonlinear", kRK)
    if held is not None and kRK != 1:
        add = held["add"]
    else:
        add = correction_modes(network_correction(params, norm9, phys, emulate), N1, N2, ny)
        if held is not None:
            held["add"] = add
    g = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)[kRK - 1]
    for f, a in enumerate(add):
        nu = "Nu%d" % (f + 1)
        o.set(nu, o.get(nu) + a)
        du = "du%d" % (f + 1)
        d = o.get(du)
        d[:, 1:1 + a.shape[1]] -= g * a
        o.set(du, d)
    o.phase("solve_uw", kRK)
    o.phase("solve_v", kRK)
    if cfg.get("flag_ctpress", 0) == 0:
        o.phase("meanflow_ctU", kRK)
    o.phase("solve_p", kRK)
    o.phase("v_corr", kRK)
