#!/usr/bin/env python3
"""Benchmark: NLL + analytic-gradient evaluations per second of the
Lambda_b -> J/psi p K- amplitude likelihood on one MI355X.

Workload (nominal size): 300k data events drawn by accept-reject from the
model, 3M uniform phase-space MC events; 14 Lambda* Breit-Wigners (J <= 7/2),
3 narrow Pc Breit-Wigners, 1 Zcs; Pc lineshapes smeared over 40 Gaussian
sub-bins (sigma = 2.7 MeV) for events with 4.2 < m(J/psi p) < 4.6 GeV.
All arithmetic FP64.

A step is one NLL + gradient evaluation (pwa_eval_grad: the analytic
adjoint kernel).  The plain method it is compared with is the same FP64
forward NLL kernel used for a central finite-difference gradient
(2P + 1 forward evaluations per gradient, P = floating parameters).

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "pwa-xib2jpsilambdak_amd"))

import numpy as np  # noqa: E402

from pwa_amd import abi, models, toys  # noqa: E402
from pwa_amd._lib import Context, generate_phsp  # noqa: E402

PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 (vector and matrix)
PEAK_HBM = 8000.0  # GB/s


def algorithmic_flops(model, n_events, nbins, smeared, grad=True):
    """Work model of the hot path per event: coherent contraction
    F_fwd = 8 H K (H = 8 helicity rows, K = complex coupling columns),
    adjoint F_adj = 2 F_fwd, smearing 8 S R^2 + 8 H R^2 for smeared events,
    K-matrix row-0 solve 299 S n_cls per smeared event."""
    d = model.desc
    from pwa_amd._lib import helicity_couplings
    K = len(helicity_couplings(d, model.params)) / 2.0
    R = d.n_pc
    ncls = len({(d.spin2[r], d.parity[r]) for r in range(d.n_lambda, d.n_lambda + d.n_pc)
                if d.type[r] == abi.PWA_KMATRIX})
    f = 64.0 * K * (3.0 if grad else 1.0)
    S = nbins if smeared else 1
    f += 8.0 * S * R * R + 8.0 * 8 * R * R
    f += 299.0 * S * ncls
    return f * n_events


def build_workload(nd, nm, seed=2024):
    model = models.config_nominal(2)
    t0 = time.time()
    mcols = generate_phsp(nm, seed)
    dcols = toys.accept_reject(model, nd, seed + 1)
    # start point of the evaluations: a random point near the generating one
    rng = np.random.default_rng(seed + 2)
    p = model.params.copy()
    fl = np.flatnonzero(model.float_mask)
    p[fl] *= 1.0 + 0.05 * rng.standard_normal(len(fl))
    return model, p, dcols, mcols, time.time() - t0


def cpu_baseline(model, p, dcols, mcols, nd, nm, budget_s=20.0):
    """The CPU FP64 oracle on tables prepared once, single thread, on a
    bounded event sample; scaled to the full workload and to NLL + central
    finite-difference gradient (2P + 1 NLL evaluations)."""
    import oracle
    frac = 0.01
    sd, sm = max(int(nd * frac), 1), max(int(nm * frac), 1)
    dsub, msub = dcols[:, :sd], mcols[:, :sm]
    P = oracle.Prepared(dsub, np.ones(sd), models.window_schemes(dsub[1]), msub, np.ones(sm),
                        models.window_schemes(msub[1]))
    t0 = time.time()
    k = 0
    while True:
        P.nll(model.desc, p)
        k += 1
        if time.time() - t0 > budget_s or k >= 50:
            break
    t_sample = (time.time() - t0) / k
    t_full = t_sample * (nd + nm) / (sd + sm)
    nfl = int(model.float_mask.sum())
    return {"value": 1.0 / (t_full * (2 * nfl + 1)), "unit": "NLL+grad evals/s", "cores": 1, "kind": "port",
            "sample": f"oracle NLL on {sd} data + {sm} MC events (1% of the workload), {k} evaluations, "
                      f"{t_sample:.3f} s each; scaled x{(nd + nm) / (sd + sm):.0f} to the full sample and "
                      f"x{2 * nfl + 1} (central finite differences over {nfl} parameters)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ndata", type=int, default=300_000)
    ap.add_argument("--nmc", type=int, default=3_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fd-evals", type=int, default=20, help="forward evaluations timed for the FD baseline")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "check": float(b[-1].item())}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(the likelihood would shard by events with one all-reduce per evaluation; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    model, p, dcols, mcols, t_setup = build_workload(args.ndata, args.nmc)
    nd, nm = dcols.shape[1], mcols.shape[1]
    ctx = Context(model.desc)
    ds, ms = models.window_schemes(dcols[1]), models.window_schemes(mcols[1])
    ctx.upload(abi.PWA_DATA, dcols, np.ones(nd), ds)
    ctx.upload(abi.PWA_MC, mcols, np.ones(nm), ms)
    nfl = int(model.float_mask.sum())
    print(f"[bench] setup {t_setup:.1f}s, {nd} data + {nm} MC events, {nfl} floating parameters", file=sys.stderr,
          flush=True)

    # --- analytic NLL + gradient (the metric)
    for _ in range(args.warmup):
        ctx.eval_grad(p)
    launches = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nll, g = ctx.eval_grad(p)
        launches.append(ctx.last_launches())
    t_grad = (time.perf_counter() - t0) / args.steps
    print(f"[bench] grad: {t_grad * 1e3:.3f} ms/step", file=sys.stderr, flush=True)

    # --- plain method: forward FP64 NLL, central FD gradient
    for _ in range(2):
        ctx.eval(p)
    t0 = time.perf_counter()
    for _ in range(args.fd_evals):
        ctx.eval(p)
    t_fwd = (time.perf_counter() - t0) / args.fd_evals
    fd_rate = 1.0 / (t_fwd * (2 * nfl + 1))
    print(f"[bench] forward: {t_fwd * 1e3:.3f} ms/eval -> FD gradient {1 / fd_rate:.3f} s", file=sys.stderr,
          flush=True)

    # --- roofline of the dominant kernel (the MC-sample gradient kernel,
    #     all its launches of one evaluation), device times from HIP events
    frac_smeared = float(np.mean(ms > 0))
    mc_launch_ms, mc_flops, mc_events = 0.0, 0.0, 0
    nb = model.desc.scheme_nbins[0]
    for step in launches:
        for L in step:
            if L["grad"] and L["kind"] == abi.PWA_MC:
                mc_launch_ms += L["ms"]
                mc_events += L["n"]
                mc_flops += algorithmic_flops(model, L["n"], nb, L["scheme"] > 0)
    n_launch = sum(1 for step in launches for L in step if L["grad"] and L["kind"] == abi.PWA_MC)
    avg_ms = mc_launch_ms / n_launch
    avg_flops = mc_flops / n_launch
    achieved = avg_flops / (avg_ms * 1e-3) / 1e12
    hbm_bytes = 368.0  # bytes read per event (46 FP64 planes)
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64, 4), "traffic": None,
            "kernel": "k_grad<3,MC> (per launch; MC segments smeared + unsmeared)",
            "avg_launch_ms": round(avg_ms, 4), "flops_per_launch": avg_flops,
            "hbm_input_GBps": round(hbm_bytes * (mc_events / n_launch) / (avg_ms * 1e-3) / 1e9, 1)}

    out = {
        "metric": "NLL+grad evals/s (300k ev, 3M MC); wall-time to global min vs plain GPU FCN",
        "value": round(1.0 / t_grad, 3),
        "unit": "NLL+grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_grad * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded phase-space MC; data by accept-reject from the model)",
        "config": {"workload": "nominal: 300k data / 3M MC, 14 Lambda* + 3 Pc BW + 1 Zcs, 40-sub-bin smearing",
                   "n_data": nd, "n_mc": nm, "n_float_params": nfl, "smeared_fraction_mc": round(frac_smeared, 4)},
        "plain_fd": {"forward_ms": round(t_fwd * 1e3, 4), "evals_per_s": fd_rate,
                     "speedup": round((1.0 / t_grad) / fd_rate, 2)},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(model, p, dcols, mcols, nd, nm)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
