"""Benchmark: grid-point RK3 substeps per second of the plain-FFT RK3 step
on one MI355X (FP64 plain by default; --precision mixed for FP32 FFTs).
The default workload is the largest single-GPU channel: 256x256x257,
Re_b = 10000, 8x8 posts at spacing 16 on both walls, dt = 1e-3.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                  [--precision fp64|mixed] [--no-cpu]

Prints ONE JSON line.  value = grid points (N1*N2*ny summed over channels)
x RK3 substeps / wall second, with the fields resident in HBM, timed between
two device synchronisations around K graph-replayed steps.  `roofline` is for
the dominant kernel, from its HIP-event duration measured here and the
algorithmic bytes that kernel must move.  `cpu_baseline` times the CPU
oracle (a from-scratch FP64 port of the reference algorithm) on a bounded
sample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pt RK3 substeps/s, FP64 plain vs mixed/batched/closure; HBM % of peak"
HBM_PEAK_GBS = 8000.0
KERNELS = ["k_zinv", "k_xprod", "k_zfwd", "k_rhs", "k_solve"]
# the textured-wall kernels (k_tex_adj + k_tex_fused x 2 launches), timed as
# one bucket; not a single kernel, so not a roofline candidate
TEX = "k_tex"


def grid_points(c, cfgs):
    """Fourier modes x ny a channel advances; a wall band of a banded
    channel counts its own pencils only (the middle band's modes are the
    middle band's)"""
    role = c.get("band", (None,))[0]
    if role in (0, 2):
        mid = next(m for m in cfgs if m.get("band") == (1, c["band"][1]))
        return (c["N1"] * c["N2"] - mid["N1"] * mid["N2"]) * c["ny"]
    return c["N1"] * c["N2"] * c["ny"]


def algorithmic_bytes(cfg, precision):
    """Bytes each kernel of one substep must read and write at least once
    (intermediates of the nonlinear pipeline at their stored precision)."""
    N1, N2, ny = cfg["N1"], cfg["N2"], cfg["ny"]
    Ng2 = cfg["Ngal2"]
    nxh = N1 // 2 + 1
    ncol = nxh * N2
    nyu, nyv = ny + 1, ny
    nplu, nplv = nyu - 2, nyv - 2
    npu, npv = nplu, nplv  # planes the nonlinear stages transform
    if cfg.get("band", (None,))[0] == 1:
        # the middle band of a banded channel transforms its own planes only
        # (u rows nb+2 .. ny-nb-2, v rows nb+1 .. ny-nb-2)
        nb = cfg["band"][1]
        npu, npv = ny - 3 - 2 * nb, ny - 2 - 2 * nb
    row = ncol * 16
    es = 8 if precision == "mixed" else 16
    plane = Ng2 * nxh * es
    hz = (npu + npv) * 3 * plane
    gx = (npu * 4 + npv * 2) * plane
    zinv = (2 * nyu + nyv) * row + hz
    xprod = hz + gx
    zfwd = gx + (3 * npu + 3 * npv) * row
    if precision == "mixed":
        # the FP64 modes of the plane-mean terms: u1, u3 on u-planes, u2 on v-planes
        zfwd += (2 * npu + npv) * row
    rhs = ((2 * nyu + nyv + (nyu - 2)) + (2 * nplu + nplv) * 2 + (2 * nyv + nplu)) * row \
        + (2 * nyu + nyv + 2 * nplu + nplv) * row
    # reads r1, r3, r2, p and the pivots of the U/W, V and Poisson operators
    # (nyu doubles per column each); writes u1, u3, u2, p, psi, div
    ne = nyu - 2
    solve = (2 * nyu + nyv + ne) * row + 3 * nyu * ncol * 8 + (2 * nyu + nyv + 3 * ne) * row
    return dict(k_zinv=zinv, k_xprod=xprod, k_zfwd=zfwd, k_rhs=rhs, k_solve=solve)


def baseline_model_bytes(cfg, precision):
    """B_sub = 78 S + 57 F of the roofline model the metric is quoted with."""
    N1, N2, ny = cfg["N1"], cfg["N2"], cfg["ny"]
    S = (N1 // 2 + 1) * N2 * ny * 16
    F = (cfg["Ngal1"] + 2) * cfg["Ngal2"] * ny * (4 if precision == "mixed" else 8)
    return 78 * S + 57 * F


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by what
    the counter reads for a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "%s_%s_hbm_traffic.md" % (workload, precision))
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes" and f[1] == kernel:
            return int(f[2])
    return None


def triad_setup(cfg, mode):
    """Lists of the triad-list path for one channel: 'full' = the threshold-0
    lists on every plane, 'thinned' = 65 % of their rows (the density of the
    reference's own 32x32x153 files, 791 627 of 1 216 945 rows for plane 1).
    Returns (files, map_u, map_v)."""
    from rnl_multiblock_nofft_amd import triads
    full = triads.full_lists(cfg["N1"], cfg["N2"])
    lists = full if mode == "full" else triads.thinned(full, 0.65, seed=1)
    ny = cfg["ny"]
    return [lists], [0] * (ny - 1), [0] * (ny - 2)


def triad_bytes(cfg, files):
    """Algorithmic bytes of the list stage per substep (k_triad + k_triad_fin):
    every table entry (4 B per row and z mirror) and target offset read once,
    the operand rows of each plane read once, the product rows written and
    read back once, the derivative and flux rows written."""
    N1, N2, ny = cfg["N1"], cfg["N2"], cfg["ny"]
    ncol = (N1 // 2 + 1) * N2
    row = ncol * 16
    nplu, nplv = ny - 1, ny - 2
    lists = files[0]

    def ent(types):
        n = 0
        for t in types:
            a = lists[t].astype(int)
            n += 2 * len(a) - int(((a[:, 1] + a[:, 3]) == 0).sum())
        return n * 4 + len(types) * (ncol + 1) * 4

    return (nplu * ent((1, 3, 5, 7, 9)) + nplv * ent((2, 4, 6, 8)) + (4 * nplu + 5 * nplv) * row
            + 2 * (5 * nplu + 4 * nplv) * row + 3 * (nplu + nplv) * row)


def host_cpu():
    """(nproc, CPU model) of this host"""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def cpu_baseline(cfg, seconds=20.0, max_sub=100000, triads_mode=None):
    """The FP64 CPU oracle (OpenMP) on one channel of the workload for about
    `seconds` of CPU work: a reported baseline.  Threads: this process's CPU
    share (OMP_NUM_THREADS, which the GPU hosts set to their per-GPU share),
    else every core the OS reports."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import init_fields, oracle_for
    from oracle import pyoracle as po
    from rnl_multiblock_nofft_amd import Solver  # noqa: F401

    nproc, model = host_cpu()
    cores = int(os.environ.get("OMP_NUM_THREADS") or nproc)
    cores = po.set_threads(cores)
    o = oracle_for(cfg)
    g = o.grid()
    f = init_fields(cfg, g, seed=1234)
    for name in ("u1", "u2", "u3", "p"):
        o.set(name, f[name])
    o.init_flowrate()
    if triads_mode:
        o.set_triads(*triad_setup(cfg, triads_mode))
    gp = cfg["N1"] * cfg["N2"] * cfg["ny"]
    if gp < 2000000:
        o.substep(1)  # warm plans / threads (a large channel's first substep is its sample)
    n = 0
    t0 = time.perf_counter()
    while n < max_sub:
        o.substep(n % 3 + 1)
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return dict(value=gp * n / dt, unit="grid-pt RK3 substeps/s", cores=cores, kind="port",
                nproc=nproc, cpu_model=model,
                sample="%d RK3 substeps of one %dx%dx%d channel (FP64 CPU oracle%s, OpenMP, %d threads "
                       "of %d cores, %s), %.1f s"
                       % (n, cfg["N1"], cfg["N2"], cfg["ny"],
                          ", triad-list path" if triads_mode else "", cores, nproc, model, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="large_256x256x257")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", default="",
                    help="fixed launch shapes instead of autotuning, e.g. "
                         "zinv_points=768,xprod_points=384,zfwd_points=384,solve_cols=3")
    ap.add_argument("--smooth", action="store_true",
                    help="drop the workload's wall textures (smooth no-slip walls)")
    ap.add_argument("--kernel-substeps", type=int, default=30)
    ap.add_argument("--triads", default="", choices=["", "full", "thinned"],
                    help="run the triad-list nonlinear path (the reference's restricted "
                         "nonlinear term) with generated lists on every plane")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB, 5 times (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters against known traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one channel ensemble on one GPU; --gpus %d is not "
              "supported (multi-GPU sharding is described in DESIGN.md, not implemented)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch

    from rnl_multiblock_nofft_amd import Solver, _lib, configs, initial

    torch.cuda.set_device(0)
    if args.copy_probe_mib:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_probe_bytes=n * 8, launches=5)), flush=True)
        return
    cfgs = configs.WORKLOADS[args.workload]()
    if args.smooth:
        cfgs = configs.smooth(cfgs)

    def start(precision):
        """a solver of the workload holding the seeded initial fields"""
        s = Solver(cfgs, precision=precision)
        for q, c in enumerate(cfgs):
            if c.get("band", (1,))[0] != 1:
                continue  # a wall band of a banded channel starts from zero short pencils
            kmax = min(8, c["N1"] // 2 - 1, c["N2"] // 2 - 1)
            f = initial.perturbed_poiseuille(c, s.grid(q), 1234 + q, kmax=kmax)
            for name in ("u1", "u2", "u3", "p"):
                s.upload(q, name, f[name])
        s.init_flowrate()
        return s

    s = start(args.precision)
    tri_files = {}
    if args.triads:
        for q, c in enumerate(cfgs):
            tri_files[q] = triad_setup(c, args.triads)
            s.set_triads(q, *tri_files[q])
    if args.tune:
        keys = {"zinv_points": _lib.TUNE_ZINV_POINTS, "xprod_points": _lib.TUNE_XPROD_POINTS,
                "zfwd_points": _lib.TUNE_ZFWD_POINTS, "solve_cols": _lib.TUNE_SOLVE_COLS,
                "solve_scan": _lib.TUNE_SOLVE_SCAN, "solve_waves": _lib.TUNE_SOLVE_WAVES,
                "xprod_inplace": _lib.TUNE_XPROD_INPLACE, "zfwd_inplace": _lib.TUNE_ZFWD_INPLACE,
                "zinv_inplace": _lib.TUNE_ZINV_INPLACE,
                "zfwd_fieldwise": _lib.TUNE_ZFWD_FIELDWISE,
                "xprod_threads": _lib.TUNE_XPROD_THREADS, "rhs_fused": _lib.TUNE_RHS_FUSED,
                "ghost_in_zinv": _lib.TUNE_GHOST_IN_ZINV, "tex_fused": _lib.TUNE_TEX_FUSED,
                "tex_separable": _lib.TUNE_TEX_SEPARABLE, "shape_groups": _lib.TUNE_SHAPE_GROUPS}
        for kv in args.tune.split(","):
            k, v = kv.split("=")
            s.set_tuning(keys[k], int(v))
        tuning = s.tuning()
    else:
        tuning = s.autotune(30)
    s.step(args.warmup)
    s.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    gp = sum(grid_points(c, cfgs) for c in cfgs)
    value = gp * 3 * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps

    # live per-kernel durations (HIP events on the library's stream)
    nsub = max(3, args.kernel_substeps)
    ms = s.time_kernels(nsub) / nsub
    kernels = list(KERNELS)
    alg = {k: 0 for k in KERNELS}
    for c in cfgs:
        for k, v in algorithmic_bytes(c, args.precision).items():
            alg[k] += v
    if args.triads:
        # the list kernels are timed in the x-product slot; the FFT stages
        # do not run
        kernels[1] = "k_triad"
        alg["k_triad"] = sum(triad_bytes(c, tri_files[q][0]) for q, c in enumerate(cfgs))
        ms[0] = ms[2] = 0.0
    dom = kernels[int(max(range(5), key=lambda q: ms[q]))]  # the largest single kernel
    dom_ms = float(ms[kernels.index(dom)])
    achieved = alg[dom] / (dom_ms * 1e-3) / 1e9
    kname = dom
    if dom == "k_solve" and tuning.get("solve_cols", 0) > 0:
        kname = "k_solve_lds"  # the LDS-staged form the autotuner chose
    if dom == "k_solve" and tuning.get("solve_cols", 0) < 0:
        kname = "k_solve_wave"  # one column per wave
    if dom == "k_zfwd" and tuning.get("zfwd_fieldwise", 0) > 0:
        kname = "k_zfwd_fs"  # the field-by-field form
    traffic = measured_traffic(args.workload, args.precision, kname)
    roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic, kernel=kname,
                    kernel_ms=round(dom_ms, 5), alg_bytes=alg[dom],
                    kernels_ms=dict({k: round(float(ms[q]), 5) for q, k in enumerate(kernels)},
                                    **{TEX: round(float(ms[5]), 5)}))
    bmodel = sum(baseline_model_bytes(c, args.precision) for c in cfgs)
    sub_per_s = 3 * args.steps / elapsed
    hbm_pct = 100.0 * bmodel * sub_per_s / (HBM_PEAK_GBS * 1e9)
    # what the fused kernels must move per substep (their algorithmic bytes,
    # texture passes left out): the bandwidth the step actually uses
    step_alg = sum(alg[k] for k in kernels)
    hbm_pct_alg = 100.0 * step_alg * sub_per_s / (HBM_PEAK_GBS * 1e9)

    mixed_err = None
    if args.precision == "mixed":
        # one step of the same seeded fields in mixed precision and in FP64
        # (default launch shapes): relative L2 per field, the largest over
        # the channels
        import numpy as np
        pair = [start("mixed"), start("fp64")]
        for x in pair:
            x.step(1)
            x.sync()
        mixed_err = {}
        for name in ("u1", "u2", "u3", "p"):
            e = 0.0
            for q in range(len(cfgs)):
                a, b = pair[0].download(q, name), pair[1].download(q, name)
                nb = float(np.linalg.norm(b))
                if nb > 0:
                    e = max(e, float(np.linalg.norm(a - b)) / nb)
            mixed_err[name] = float("%.3g" % e)
        for x in pair:
            x.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(max(cfgs, key=lambda c: c["N1"] * c["N2"] * c["ny"]), triads_mode=args.triads or None)
    rec = dict(metric=METRIC, value=round(value, 1), unit="grid-pt RK3 substeps/s", n_gpus=1,
               steps=args.steps, warmup=args.warmup, ms_per_step=round(ms_per_step, 5),
               higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64" if args.precision == "fp64" else "f32-fft/f64-solve", data="synthetic",
               config=dict(workload=args.workload, channels=len(cfgs),
                           grids=sorted({"%dx%dx%d" % (c["N1"], c["N2"], c["ny"]) for c in cfgs}),
                           precision=args.precision,
                           walls="textured" if any(c.get("texture") is not None for c in cfgs) else "smooth",
                           dt=cfgs[0]["dt"],
                           nonlinear=("triad lists (%s)" % args.triads) if args.triads else "fft",
                           grid_points=gp, hbm_pct_of_peak_model=round(hbm_pct, 3),
                           alg_bytes_per_substep=step_alg,
                           hbm_pct_of_peak_alg_bytes=round(hbm_pct_alg, 3),
                           tuning=tuning),
               roofline=roofline, cpu_baseline=cpu)
    if mixed_err is not None:
        rec["config"]["mixed_error_vs_fp64_one_step"] = mixed_err
    print(json.dumps(rec), flush=True)
    s.close()


if __name__ == "__main__":
    main()
