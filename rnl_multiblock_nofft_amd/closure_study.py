"""The learned-closure workload end to end on one GPU.

  python -m rnl_multiblock_nofft_amd.closure_study [--fine 128] [--coarse 32]
      [--ny 153] [--snapshots 2000] [--every 5] [--spinup 200]
      [--train-steps 4000] [--batch 8192] [--post-steps 2000] [--amp 0.1]
      [--save params.npz]

1. A fine channel (fine x fine x ny, Re_b = 2800, constant flow rate,
   dt = 1e-3) starts from the seeded perturbed Poiseuille field and runs
   `spinup` steps; then every `every` steps a training sample is taken for
   the coarse channel (coarse x coarse x ny, same box and y grid): the fine
   fields box-filtered over a coarse cell, and tau = filtered fine nonlinear
   term - coarse nonlinear term of the filtered fields (lh_closure_sample).
2. The network (135 -> 128 -> 128 -> 128 -> 3) trains on the first 90 % of
   the samples with AdamW; its a-priori error is measured on the last 10 %.
3. From the filtered fine state at the end of sampling, a coarse run with
   the closure (network at every substep), one with the network at the
   first substep of each step only (correction held for the step), the
   held form in mixed precision (FP32 transforms, FP64 solves) and one
   without advance `post_steps` steps beside the fine run, all accumulating mean and rms profiles every 10 steps (the fine run's
   filtered fields are accumulated too); the coarse profiles are compared
   with the fine run's.
4. GPU time per step of the fine run and of the coarse runs.

Prints progress lines and one JSON line with the results.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import Solver, closure, configs, initial


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def profiles(solver, ch=0):
    """Mean u and rms u', v', w' per interior row from the device statistics
    (sums over the channel's physical grid and the samples)."""
    prof, istat = solver.stats(ch)
    c = solver.cfgs[ch]
    n = c["Ngal1"] * c["Ngal2"] * max(istat, 1)
    ny = c["ny"]
    U = prof[0, 1:ny] / n
    urms = np.sqrt(np.maximum(prof[1, 1:ny] / n - U ** 2, 0))
    V = prof[2, 1:ny - 1] / n
    vrms = np.sqrt(np.maximum(prof[3, 1:ny - 1] / n - V ** 2, 0))
    W = prof[4, 1:ny] / n
    wrms = np.sqrt(np.maximum(prof[5, 1:ny] / n - W ** 2, 0))
    return dict(U=U, urms=urms, vrms=vrms, wrms=wrms)


def profile_errors(p, ref):
    """max |p - ref| / max |ref| per profile."""
    return {k: float(np.max(np.abs(p[k] - ref[k])) / max(np.max(np.abs(ref[k])), 1e-300)) for k in ref}


def time_steps(solver, n):
    solver.step(2)
    solver.sync()
    t0 = time.perf_counter()
    solver.step(n)
    solver.sync()
    return (time.perf_counter() - t0) / n


def run(args):
    Re = configs.re_code(2800.0)
    cf = configs.channel(args.fine, args.fine, args.ny, Re)
    cc = configs.channel(args.coarse, args.coarse, args.ny, Re)
    fine = Solver([cf])
    sampler = Solver([cc])
    f0 = initial.perturbed_poiseuille(cf, fine.grid(0), args.seed, amp=args.amp,
                                      kmax=min(8, args.coarse // 2 - 1))
    for name in ("u1", "u2", "u3", "p"):
        fine.upload(0, name, f0[name])
    fine.init_flowrate()
    fine.autotune(5)
    say("fine %dx%dx%d spin-up %d steps" % (args.fine, args.fine, args.ny, args.spinup))
    fine.step(args.spinup)
    fine.sync()
    snaps = closure.Snapshots(cc, args.snapshots)
    t0 = time.perf_counter()
    for q in range(args.snapshots):
        fine.step(args.every)
        snaps.add(fine, 0, sampler, 0)
        if q % max(1, args.snapshots // 10) == 0:
            say("sample %d / %d  t = %.3f" % (q, args.snapshots, fine.scalars(0)["t"]))
    t_data = time.perf_counter() - t0
    ntrain = max(1, int(0.9 * args.snapshots))
    nheld = args.snapshots - ntrain
    norm = snaps.norm(max_snap=ntrain)
    say("normalisation", np.array2string(norm, precision=4))
    tr = closure.Trainer(args.batch, closure.init_params(args.seed), norm)
    t0 = time.perf_counter()
    hist = closure.train(tr, snaps, args.train_steps, lr=args.lr, first=0, count=ntrain, seed=args.seed,
                         log_every=max(1, args.train_steps // 10))
    t_train = time.perf_counter() - t0
    held = closure.apriori(tr, snaps, ntrain, nheld, max_points=args.apriori_points) if nheld else None
    train_err = closure.apriori(tr, snaps, 0, min(ntrain, nheld or 1), max_points=args.apriori_points)
    say("a-priori error (held-out)", held, "(training)", train_err)
    params = tr.params()
    if args.save:
        np.savez(args.save, params=params, norm=norm)

    # a-posteriori: coarse runs from the filtered fine state, beside the fine run
    runs = {}
    tmp = closure.Snapshots(cc, 1)
    mpgx = fine.scalars(0)["mpgx"]
    for name, on, prec in (("closure", 1, "fp64"), ("closure_held", 2, "fp64"), ("closure_held_mixed", 2, "mixed"),
                           ("plain", 0, "fp64")):
        s = Solver([cc], precision=prec)
        tmp.count = 0
        tmp.add(fine, 0, s, 0)
        s.init_flowrate()
        s.set_scalars(0, t=fine.scalars(0)["t"], mpgx=mpgx)
        if on:
            closure.set_network(s, params, norm)
            closure.enable(s, 0, once_per_step=on == 2)
        s.autotune(5)
        runs[name] = s
    fine.stats_reset()
    sampler.stats_reset()
    for s in runs.values():
        s.stats_reset()
    done = 0
    while done < args.post_steps:
        n = min(10, args.post_steps - done)
        fine.step(n)
        for s in runs.values():
            s.step(n)
        done += n
        fine.stats_accumulate()
        tmp.count = 0
        tmp.add(fine, 0, sampler, 0)   # filtered fine fields into the sampler
        sampler.stats_accumulate()
        for s in runs.values():
            s.stats_accumulate()
        if done % max(10, args.post_steps // 5 // 10 * 10) == 0:
            say("a-posteriori step %d / %d" % (done, args.post_steps))
    pf = profiles(fine)
    pff = profiles(sampler)
    res = {}
    for name, s in runs.items():
        p = profiles(s)
        res[name] = dict(vs_fine=profile_errors(p, pf), vs_filtered_fine=profile_errors(p, pff))
        say(name, "vs fine", res[name]["vs_fine"])
        say(name, "vs filtered fine", res[name]["vs_filtered_fine"])

    ms_fine = 1e3 * time_steps(fine, 50)
    ms_cl = 1e3 * time_steps(runs["closure"], 200)
    ms_held = 1e3 * time_steps(runs["closure_held"], 200)
    ms_held_mixed = 1e3 * time_steps(runs["closure_held_mixed"], 200)
    ms_plain = 1e3 * time_steps(runs["plain"], 200)
    ms_parts = closure.closure_time(runs["closure"], 50)
    out = dict(fine="%dx%dx%d" % (args.fine, args.fine, args.ny), coarse="%dx%dx%d" % (args.coarse, args.coarse, args.ny),
               snapshots=args.snapshots, every=args.every, spinup=args.spinup, amp=args.amp,
               train_steps=args.train_steps, batch=args.batch,
               loss_first=float(np.mean(hist[:10])), loss_last=float(np.mean(hist[-50:])),
               apriori_heldout=None if held is None else [float(v) for v in held],
               apriori_train=[float(v) for v in train_err],
               post_steps=args.post_steps, profile_errors=res,
               ms_per_step=dict(fine=ms_fine, coarse_closure=ms_cl, coarse_closure_held=ms_held,
                                coarse_closure_held_mixed=ms_held_mixed,
                                coarse_plain=ms_plain),
               closure_ms_per_substep=dict(planes=ms_parts[0], network=ms_parts[1], to_modes=ms_parts[2]),
               gpu_time_ratio_fine_over_closure=ms_fine / ms_cl,
               gpu_time_ratio_fine_over_closure_held=ms_fine / ms_held,
               gpu_time_ratio_fine_over_closure_held_mixed=ms_fine / ms_held_mixed,
               seconds=dict(data=t_data, train=t_train))
    print(json.dumps(out), flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--fine", type=int, default=128)
    ap.add_argument("--coarse", type=int, default=32)
    ap.add_argument("--ny", type=int, default=153)
    ap.add_argument("--snapshots", type=int, default=2000)
    ap.add_argument("--every", type=int, default=5)
    ap.add_argument("--spinup", type=int, default=200)
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--post-steps", type=int, default=2000)
    ap.add_argument("--amp", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--apriori-points", type=int, default=65536)
    ap.add_argument("--save", default="")
    run(ap.parse_args(argv))


if __name__ == "__main__":
    main()
