"""Learned closure of the nonlinear term: host side.

No counterpart in the reference.  A coarse channel (e.g. 32 x 32 x 153)
adds, at every substep, a small network's estimate of the part of the
nonlinear term that its grid misses.  The network is trained on samples that
the GPU solver itself makes from a fine run (e.g. 128 x 128 x 153) of the
same box and y grid:

    tau = G N(u_fine) - N_coarse(G u_fine)

G the box filter of one coarse cell restricted to the coarse modes
(lh_closure_sample).  Everything that touches the fields runs through the
C-ABI of liblhgpu.so (lh_closure_*, lh_mlp_*); this module holds parameters,
device buffers and the training loop.

Parameter vector (LH_CL_NPARAM floats): W1[128][135] b1 W2[128][128] b2
W3[128][128] b3 W4[3][128] b4, row-major, y = W x + b; normalisation
norm9 = mean[3], inv_scale[3], out_scale[3].
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check

NIN, NH, NOUT = _lib.CL_NIN, _lib.CL_NHID, _lib.CL_NOUT
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)), ("b3", (NH,)),
          ("W4", (NOUT, NH)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
assert NPARAM == _lib.CL_NPARAM


def unpack(vec):
    """Parameter vector -> dict of arrays (views)."""
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = vec[o:o + n].reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float32).ravel() for name, _ in SHAPES]).astype(np.float32)


def init_params(seed=0, out_gain=0.1):
    """Random weights: normal with variance 1/fan_in (hidden layers), the
    output layer scaled by out_gain; zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            w = rng.standard_normal(shp) / np.sqrt(shp[1])
            d[name] = w * (out_gain if name == "W4" else 1.0)
        else:
            d[name] = np.zeros(shp)
    return pack(d)


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class DeviceArray:
    """A float32 device buffer (lh_dev_alloc), for snapshots and planes."""

    def __init__(self, shape):
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = 4 * int(np.prod(self.shape))
        self.ptr = ctypes.c_void_p()
        L = _lib.load()
        check(L.lh_dev_alloc(self.nbytes, ctypes.byref(self.ptr)))
        self.L = L

    def at(self, index):
        """Device address of element `index` of the first axis."""
        stride = self.nbytes // self.shape[0]
        return ctypes.c_void_p(self.ptr.value + index * stride)

    def get(self):
        out = np.empty(self.shape, np.float32)
        check(self.L.lh_copy_to_host(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes))
        return out

    def get_rows(self, first, count):
        """Elements first .. first+count-1 of the first axis."""
        out = np.empty((count,) + self.shape[1:], np.float32)
        check(self.L.lh_copy_to_host(out.ctypes.data_as(ctypes.c_void_p), self.at(first), out.nbytes))
        return out

    def set(self, arr):
        a = np.ascontiguousarray(arr, np.float32)
        assert a.shape == self.shape
        check(self.L.lh_copy_to_device(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes))

    def free(self):
        if self.ptr:
            self.L.lh_dev_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def plane_shape(cfg):
    """Shape of one set of physical planes of a channel: (3, ny+1, N2, N1)."""
    return (3, cfg["ny"] + 1, cfg["N2"], cfg["N1"])


def set_network(solver, params, norm9):
    check(solver.L.lh_closure_set(solver.h, _fp(np.ascontiguousarray(params, np.float32)),
                                  _fp(np.ascontiguousarray(norm9, np.float32))))


def enable(solver, ch, on=True, once_per_step=False):
    """Switch the correction on (every substep evaluates the network) or, with
    once_per_step, evaluate it at the first substep of each step and reuse
    it at the other two."""
    check(solver.L.lh_closure_enable(solver.h, ch, (2 if once_per_step else 1) if on else 0))


def evaluate(solver, ch, with_correction=True):
    """(phys, corr) of the current fields: the N1 x N2 planes of u1, u2, u3
    and, if asked, the network's correction (else None)."""
    shp = plane_shape(solver.cfgs[ch])
    ph = DeviceArray(shp)
    co = DeviceArray(shp) if with_correction else None
    check(solver.L.lh_closure_eval(solver.h, ch, ph.ptr, co.ptr if co else None))
    return ph.get(), (co.get() if co else None)


def closure_time(solver, nrep=20):
    """ms of [physical planes, network, correction to modes] per substep."""
    ms = np.zeros(3)
    check(solver.L.lh_closure_time(solver.h, int(nrep), ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return ms


class Snapshots:
    """nsnap training samples of a coarse channel on the device."""

    def __init__(self, cfg, nsnap):
        self.cfg = dict(cfg)
        self.nsnap = int(nsnap)
        shp = (self.nsnap,) + plane_shape(cfg)
        self.phys = DeviceArray(shp)
        self.tau = DeviceArray(shp)
        self.count = 0

    def add(self, fine, chf, coarse, chc):
        """One sample (lh_closure_sample): replaces the coarse channel's
        fields with the filtered fine ones."""
        if self.count >= self.nsnap:
            raise ValueError("snapshot buffer full")
        check(fine.L.lh_closure_sample(fine.h, chf, coarse.h, chc, self.phys.at(self.count),
                                       self.tau.at(self.count)))
        self.count += 1

    def norm(self, max_snap=None):
        """Normalisation from the stored samples: per field mean and
        1/std of the inputs (interior rows), rms of each target."""
        n = self.count if max_snap is None else min(self.count, max_snap)
        pick = np.unique(np.linspace(0, n - 1, min(n, 64)).astype(int))
        ph = np.stack([self.phys.get_rows(int(q), 1)[0] for q in pick])[:, :, 1:-1]
        ta = np.stack([self.tau.get_rows(int(q), 1)[0] for q in pick])[:, :, 1:-1]
        mean = ph.mean(axis=(0, 2, 3, 4))
        std = ph.std(axis=(0, 2, 3, 4))
        rms = np.sqrt((ta.astype(np.float64) ** 2).mean(axis=(0, 2, 3, 4)))
        return np.concatenate([mean, 1.0 / np.maximum(std, 1e-12), np.maximum(rms, 1e-30)]).astype(np.float32)


class Trainer:
    """lh_mlp: training batches of `batch` points on the device."""

    def __init__(self, batch, params, norm9):
        self.L = _lib.load()
        self.batch = int(batch)
        h = ctypes.c_void_p()
        check(self.L.lh_mlp_create(self.batch, ctypes.byref(h)))
        self.h = h
        self.norm9 = np.ascontiguousarray(norm9, np.float32)
        self.set_params(params)

    def close(self):
        if getattr(self, "h", None):
            self.L.lh_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        check(self.L.lh_mlp_set_params(self.h, _fp(np.ascontiguousarray(params, np.float32)), _fp(self.norm9)))

    def params(self):
        out = np.zeros(NPARAM, np.float32)
        check(self.L.lh_mlp_get_params(self.h, _fp(out)))
        return out

    def grad(self):
        out = np.zeros(NPARAM, np.float32)
        check(self.L.lh_mlp_get_grad(self.h, _fp(out)))
        return out

    def set_batch(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        assert x.shape == (self.batch, NIN) and t.shape == (self.batch, NOUT)
        check(self.L.lh_mlp_set_batch(self.h, _fp(x), _fp(t)))

    def sample(self, snaps, seed, first=0, count=None):
        """A random batch from snapshots first .. first+count-1."""
        count = snaps.count - first if count is None else count
        c = snaps.cfg
        check(self.L.lh_mlp_sample(self.h, c["N1"], c["N2"], c["ny"], snaps.phys.at(first), snaps.tau.at(first),
                                   int(count), ctypes.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF)))

    def set_points(self, cfg, phys_ptr, pts):
        p = np.ascontiguousarray(pts, np.int32)
        assert p.shape == (self.batch,)
        check(self.L.lh_mlp_set_points(self.h, cfg["N1"], cfg["N2"], cfg["ny"], phys_ptr,
                                       p.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def forward(self):
        y = np.zeros((self.batch, NOUT), np.float32)
        check(self.L.lh_mlp_forward(self.h, _fp(y)))
        return y

    def loss_grad(self):
        loss = ctypes.c_double()
        check(self.L.lh_mlp_grad(self.h, ctypes.byref(loss)))
        return loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, wd=0.0):
        check(self.L.lh_mlp_adam(self.h, float(lr), float(beta1), float(beta2), float(eps), float(wd)))


def splitmix64(x):
    """The batch sampler's hash (numpy uint64 arithmetic, wrapping)."""
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def batch_points(seed, batch, total):
    """Point indices lh_mlp_sample draws (snapshot-major)."""
    n = np.arange(batch, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = splitmix64(np.uint64(seed) ^ (n * np.uint64(0xD1B54A32D192ED03)))
    return (h % np.uint64(total)).astype(np.int64)


def train(trainer, snaps, steps, lr=1e-3, first=0, count=None, seed=1, log_every=0, wd=0.0, lr_final=None):
    """AdamW steps on random batches of snapshots first .. first+count-1;
    cosine decay of the rate to lr_final (default lr/20).  Returns the loss
    history."""
    hist = []
    lr_final = lr / 20 if lr_final is None else lr_final
    for it in range(steps):
        trainer.sample(snaps, seed * 1000003 + it, first, count)
        loss = trainer.loss_grad()
        a = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * it / max(steps - 1, 1)))
        trainer.adam(a, wd=wd)
        hist.append(loss)
        if log_every and (it % log_every == 0 or it == steps - 1):
            print("closure train %5d  loss %.5f  lr %.2e" % (it, loss, a), flush=True)
    return np.array(hist)


def apriori(trainer, snaps, first, count, max_points=None):
    """A-priori error of the network on snapshots first .. first+count-1:
    per output, sum (y - t)^2 / sum t^2 over every point of u rows
    1..ny-1 (t = tau / out_scale; the zero correction scores 1), or over
    max_points points spread evenly through each snapshot."""
    c = snaps.cfg
    plane = c["N1"] * c["N2"]
    npts = (c["ny"] - 1) * plane
    B = trainer.batch
    scale = trainer.norm9[6:9].astype(np.float64)
    num, den = np.zeros(NOUT), np.zeros(NOUT)
    use = np.arange(npts) if max_points is None or max_points >= npts else \
        np.linspace(0, npts - 1, max_points).astype(np.int64)
    for s in range(first, first + count):
        tau = snaps.tau.get_rows(s, 1)[0]
        t_all = tau[:, 1:c["ny"]].reshape(NOUT, -1).T / scale
        for p0 in range(0, len(use), B):
            k = min(B, len(use) - p0)
            pts = use[np.arange(p0, p0 + B) % len(use)]
            trainer.set_points(c, snaps.phys.at(s), pts)
            y = trainer.forward().astype(np.float64)
            t = t_all[pts[:k]]
            num += ((y[:k] - t) ** 2).sum(axis=0)
            den += (t ** 2).sum(axis=0)
    return num / np.maximum(den, 1e-300)
