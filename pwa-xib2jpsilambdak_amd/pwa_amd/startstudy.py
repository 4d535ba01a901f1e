"""Wall-time to the global minimum: plain random-start fits (finite-difference
gradient of the forward FCN) against the accelerated procedure (analytic
gradient + learned start proposer), on one GPU.

  python -m pwa_amd.startstudy [--config 1] [--ndata 20000] [--nmc 100000]

Procedure (all fits within the same parameter limits):
  1. n_random random starts fitted with the analytic gradient: the best FCN
     of all fits made here defines the global minimum F*; a fit reaches it
     when its FCN < F* + 0.1.  p_random = fraction of random starts that do.
     Both gradient modes run the same minimiser, so they sample the same
     basins of attraction; single trajectories from one start may still end
     in different minima (the finite-difference gradient's truncation error
     changes the path), so p_random is a property of the start distribution
     and the landscape, estimated with the cheap analytic fits.
  2. n_plain random starts fitted in the plain mode: mean time t_plain.
     Expected wall-time of the plain procedure = t_plain / p_random.
  3. Accelerated: n_train analytic fits from random starts record every
     (parameters, FCN, gradient) with the fit's final FCN; the MLP is trained
     on them; 4096 random starts are ranked by their predicted fit outcome;
     the top n_top are refit (analytic gradient) in rank order until one
     reaches F*.
     Wall-time = data fits + training + ranking + refits up to that success.
Prints one JSON line.
"""
import argparse
import json
import time

import numpy as np

from . import abi, models, toys
from ._lib import Context, generate_phsp
from .fit import default_limits, fit, random_starts
from .proposer import Proposer


def _log(msg):
    print(f"[startstudy] {msg}", flush=True)


def setup(model, nd, nm, seed):
    mcols = generate_phsp(nm, seed)
    dcols = toys.accept_reject(model, nd, seed + 1)
    ctx = Context(model.desc)
    sd = models.window_schemes(dcols[abi.COL["mjpsiL"]]) if model.desc.n_schemes else None
    sm = models.window_schemes(mcols[abi.COL["mjpsiL"]]) if model.desc.n_schemes else None
    ctx.upload(abi.PWA_DATA, dcols, np.ones(nd), sd)
    ctx.upload(abi.PWA_MC, mcols, np.ones(nm), sm)
    return ctx


def run(config=1, ndata=20000, nmc=100000, n_random=1000, n_plain=12, n_train=64, n_top=32, n_rank=4096,
        epochs=300, seed=11, fit_seconds=20.0, plain_seconds=120.0):
    model = {0: models.config_tiny, 1: models.config_small, 2: models.config_nominal,
             3: models.config_kmatrix}[config]()
    ctx = setup(model, ndata, nmc, seed)
    lim = default_limits(model)
    rng = np.random.default_rng(seed + 5)
    res = {"config": config, "n_data": ndata, "n_mc": nmc, "n_float": model.n_float}

    # 1. random starts, analytic gradient: the global minimum and p_random
    starts = random_starts(model, n_random, rng)
    t0 = time.perf_counter()
    rand = []
    for j, x0 in enumerate(starts):
        rand.append(fit(ctx, model, x0, "analytic", lim, max_iter=1000, max_seconds=fit_seconds))
        if (j + 1) % 100 == 0:
            _log(f"random-start analytic fits: {j + 1}/{n_random}, best {min(r.fval for r in rand):.4f}")
    t_rand = time.perf_counter() - t0
    f_rand = np.array([r.fval for r in rand])
    t_an = np.array([r.seconds for r in rand])

    # 2. plain fits (finite-difference gradient of the forward FCN)
    plain = []
    for j, x0 in enumerate(starts[:n_plain]):
        plain.append(fit(ctx, model, x0, "fd", lim, max_iter=1000, max_seconds=plain_seconds))
        _log(f"plain fit {j + 1}/{n_plain}: fcn {plain[-1].fval:.4f} ({plain[-1].seconds:.1f} s)")
    t_plain = np.array([r.seconds for r in plain])
    agree = [abs(p.fval - a.fval) < 1e-3 for p, a in zip(plain, rand)]

    # 3. accelerated procedure
    t_acc0 = time.perf_counter()
    tr_starts = random_starts(model, n_train, rng)
    X, F, G, E = [], [], [], []
    train_fits = []
    for x0 in tr_starts:
        r = fit(ctx, model, x0, "analytic", lim, max_iter=1000, max_seconds=fit_seconds, keep_trajectory=True)
        train_fits.append(r)
        for x, f, g in r.trajectory:
            if np.isfinite(f) and np.all(np.isfinite(g)):
                X.append(x)
                F.append(f)
                G.append(g)
                E.append(r.fval)
    t_data = time.perf_counter() - t_acc0
    _log(f"training data: {len(X)} samples from {n_train} fits in {t_data:.1f} s")
    t1 = time.perf_counter()
    prop = Proposer(model.n_float, *lim, seed=seed)
    loss = prop.train(np.array(X), np.array(F), np.array(G), np.array(E), epochs=epochs)
    t_train = time.perf_counter() - t1
    _log(f"MLP trained: loss {loss:.4f} in {t_train:.1f} s")
    t2 = time.perf_counter()
    cand, score = prop.rank(random_starts(model, n_rank, rng), n_top)
    t_rank = time.perf_counter() - t2
    refits = []
    for x0 in cand:
        refits.append(fit(ctx, model, x0, "analytic", lim, max_iter=1000, max_seconds=fit_seconds))
    prop.close()
    f_star = min(f_rand.min(), min(r.fval for r in refits), min(r.fval for r in train_fits),
                 min(r.fval for r in plain))
    ok_rand = f_rand < f_star + 0.1
    ok_ref = np.array([r.fval < f_star + 0.1 for r in refits])
    p_rand = float(ok_rand.mean())
    f_ref = np.array([r.fval for r in refits])
    # a looser yardstick: the best 5 % of the random-start minima
    f_q05 = float(np.percentile(f_rand, 5))
    good = {"threshold_fcn": f_q05, "random": float(np.mean(f_rand <= f_q05)),
            "proposer": float(np.mean(f_ref <= f_q05)),
            "median_fcn_random": float(np.median(f_rand)), "median_fcn_proposer": float(np.median(f_ref))}
    # accelerated wall-time: data + training + ranking + refits up to the first success
    if ok_ref.any():
        k = int(np.argmax(ok_ref))
        t_acc = t_data + t_train + t_rank + sum(r.seconds for r in refits[:k + 1])
    else:
        t_acc = float("nan")
    t_plain_proc = float(t_plain.mean() / p_rand) if p_rand > 0 else float("inf")
    t_an_proc = float(t_an.mean() / p_rand) if p_rand > 0 else float("inf")
    res.update({
        "global_min_fcn": f_star,
        "random_starts": n_random, "p_random": p_rand,
        "mean_fit_s": {"plain": float(t_plain.mean()), "analytic": float(t_an.mean())},
        "plain_analytic_same_minimum": f"{sum(agree)}/{len(agree)}",
        "expected_time_to_global_min_s": {"plain_random_starts": t_plain_proc,
                                          "analytic_random_starts": t_an_proc,
                                          "analytic_plus_proposer": t_acc},
        "proposer": {"train_fits": n_train, "samples": len(X), "train_s": t_train, "data_s": t_data,
                     "rank_s": t_rank, "top": n_top, "refit_success_fraction": float(ok_ref.mean()),
                     "first_success_rank": int(np.argmax(ok_ref)) if ok_ref.any() else None},
        "best_5pct_of_minima": good,
        "gain": {"plain_over_accelerated": t_plain_proc / t_acc if np.isfinite(t_acc) else None,
                 "plain_over_analytic_random": t_plain_proc / t_an_proc if np.isfinite(t_an_proc) else None,
                 "proposer_success_vs_random": float(ok_ref.mean()) / p_rand if p_rand > 0 else None},
        "seconds_random_fits": t_rand,
    })
    ctx.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--ndata", type=int, default=20000)
    ap.add_argument("--nmc", type=int, default=100000)
    ap.add_argument("--random", type=int, default=1000)
    ap.add_argument("--plain", type=int, default=12)
    ap.add_argument("--train-fits", type=int, default=64)
    ap.add_argument("--top", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--seed", type=int, default=11)
    a = ap.parse_args()
    r = run(a.config, a.ndata, a.nmc, a.random, a.plain, a.train_fits, a.top, epochs=a.epochs, seed=a.seed)
    print(json.dumps(r))


if __name__ == "__main__":
    main()
