"""Train and evaluate the learned flux closure on the GPU.

    python aeromhd-cuda_amd/train_closure.py --fields 8 --train 200 --test 200:220 \
        --iters 3000 --out aeromhd-cuda_amd/closure_weights.npz --report closure_report.json

(the recipe of the shipped weights: the mixing flux alone, no aggregation
rounds, the correction gain and the coefficient of the limited
anti-diffusion of C calibrated together on training seeds 0-7), and to
evaluate given weights (calibrating them first unless --calibrate ''):

    python aeromhd-cuda_amd/train_closure.py --evaluate-only W.npz --out W_calibrated.npz

Data: fine 2048x1024 magnetic-sieve runs (FP32, 2000 steps) for seeds
0..train-1 restricted 8x to 256x128; held-out seeds for the validation
loss and the fine-grid evaluation.  After the first fit, each aggregation round
adds pairs taken along the corrected coarse solver's own trajectories
(rollout_pairs) and trains further on the union.  The evaluation also times the plain FP64
fine-grid solve of one seed for the speed comparison."""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def log(*a):
    print(*a, flush=True)


def final_evaluation(rep, model, eval_seeds, variants, args):
    """Evaluate on the held-out seeds (with ablations: the network without
    the sharpening of C, the sharpening without the network, no limiters),
    time the FP64 fine solve, write the report."""
    import amhd_closure as cl

    extra = tuple(v for v in ("network", "sharpen", "nolimit") if v not in variants)
    rows = cl.evaluate(eval_seeds, model, steps=args.fine_steps, log=log, variants=variants + extra)
    rep["evaluation"] = rows
    rep["fp64_fine"] = cl.fp64_fine_throughput(eval_seeds[0], args.fine_steps)
    fine_cells = 2048 * 1024 * args.fine_steps
    tags = ("plain",) + variants + extra
    summ = cl.summarize(rows, tags)
    fp64 = rep["fp64_fine"]["mcups"]
    for t in tags:
        summ[f"speedup_{t}_vs_fp64_fine"] = fine_cells / summ[f"wall_{t}_s_mean"] / 1e6 / fp64
    bt = cl.evaluate_batched(eval_seeds, model, steps=args.fine_steps, log=log)
    fine_cells = 2048 * 1024 * args.fine_steps * len(eval_seeds)
    for t in ("plain", "closure", "sharpen"):
        if f"wall_{t}_s" in bt:
            bt[f"speedup_{t}_vs_fp64_fine"] = fine_cells / bt[f"wall_{t}_s"] / 1e6 / fp64
    rep["batched"] = bt
    rep["summary"] = summ
    log(f"[closure] summary {summ}")
    json.dump(rep, open(args.report, "w"), indent=1, default=float)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200)
    ap.add_argument("--test", default="200:220")
    ap.add_argument("--pairs", type=int, default=16)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--fine-steps", type=int, default=2000)
    ap.add_argument("--eval-seeds", default="", help="seed range a:b evaluated at the end (default: the test seeds)")
    ap.add_argument("--dagger-rounds", type=int, default=0)
    ap.add_argument("--dagger-seeds", default="0:40")
    ap.add_argument("--dagger-iters", type=int, default=2000)
    ap.add_argument("--dagger-lr", type=float, default=3e-4)
    ap.add_argument("--pullback", type=float, default=0.1)
    ap.add_argument("--quick-eval", type=int, default=4, help="held-out seeds evaluated after every phase")
    ap.add_argument("--variants", default="closure")
    ap.add_argument("--fields", default="8", help="network fields the closure corrects (8 = the mixing flux)")
    ap.add_argument("--evaluate-only", default="", help="evaluate these weights on the test seeds; no training")
    ap.add_argument("--calibrate", default="0:8", help="training seeds a:b for the gain calibration ('' to skip)")
    ap.add_argument("--gains", default="0.25,0.5,0.75,1.0,1.25,1.5,1.75")
    ap.add_argument("--kappas", default="0.008,0.009,0.01,0.011,0.012,0.014",
                    help="anti-diffusion coefficients, calibrated with the gain on the grid gains x kappas "
                         "('' : the gain alone, no sharpening)")
    ap.add_argument("--kappas-alone", default="0.012,0.014,0.016,0.018,0.02,0.022,0.025",
                    help="coefficients tried for the sharpening-alone ablation ('' to skip)")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "closure_weights.npz"))
    ap.add_argument("--report", default="closure_report.json")
    args = ap.parse_args()
    import amhd_closure as cl

    a, b = (int(v) for v in args.test.split(":"))
    test_seeds = list(range(a, b))
    eval_seeds = test_seeds if not args.eval_seeds else list(range(*(int(v) for v in args.eval_seeds.split(":"))))
    variants = tuple(args.variants.split(","))
    gains = [float(g) for g in args.gains.split(",")]

    def calibrate(model, rep):
        if not args.calibrate:
            return
        ca, cb = (int(v) for v in args.calibrate.split(":"))
        # the network alone (no sharpening): its own gain, kept for the ablation
        ga, gerr = cl.calibrate_gain(model, range(ca, cb), gains, steps=args.fine_steps, log=log, kappa=0.0)
        model.stats["gain_alone"] = np.asarray(ga, dtype=np.float32)
        rep["calibration"] = {"seeds": args.calibrate, "mw_err_by_gain_alone": gerr, "gain_alone": ga}
        if args.kappas:
            g, k, err = cl.calibrate_gain_kappa(model, range(ca, cb), gains,
                                                [float(v) for v in args.kappas.split(",")],
                                                steps=args.fine_steps, log=log)
            rep["calibration"].update({"mw_err_by_gain_kappa": err, "gain": g, "kappa": k})
        else:
            model.stats["kappa"] = np.asarray(0.0)
            rep["calibration"].update({"gain": ga, "kappa": 0.0})
        if args.kappas_alone:
            k, kerr = cl.calibrate_kappa(model, range(ca, cb), [float(v) for v in args.kappas_alone.split(",")],
                                         steps=args.fine_steps, log=log, alone=True)
            rep["calibration"].update({"mw_err_by_kappa_alone": kerr, "kappa_alone": k})
        model.save(args.out)

    if args.evaluate_only:
        model = cl.FluxCNN.load(args.evaluate_only)
        rep = {"weights": args.evaluate_only, "fields": cl.applied_fields(model.stats)}
        calibrate(model, rep)
        final_evaluation(rep, model, eval_seeds, variants, args)
        return
    t0 = time.time()
    data = cl.generate_pairs(range(args.train), args.pairs, steps=args.fine_steps, log=log)
    test = cl.generate_pairs(test_seeds, 4, steps=args.fine_steps, log=log)
    t_data = time.time() - t0
    dx, dy = 6.0 / 256, 2.0 / 128
    log(f"[closure] data: train {tuple(data[0].shape)}, test {tuple(test[0].shape)}, {t_data:.0f} s")
    t1 = time.time()
    fields = None if args.fields == "all" else [int(v) for v in args.fields.split(",")]
    model, hist = cl.train(data, dx, dy, iters=args.iters, batch=args.batch, lr=args.lr, log=log, holdout=test,
                           fields=fields)
    model.save(args.out)
    phases = []

    def quick(tag):
        model.save(args.out.replace(".npz", f".{tag.replace(' ', '')}.npz"))
        if args.quick_eval <= 0:
            return
        rows = cl.evaluate(test_seeds[:args.quick_eval], model, steps=args.fine_steps, variants=variants)
        ph = {"phase": tag, **cl.summarize(rows, ("plain",) + variants)}
        phases.append(ph)
        log(f"[closure] phase {tag}: " + ", ".join(f"{k}={v:.4g}" if isinstance(v, float) else f"{k}={v}"
                                                   for k, v in ph.items()))

    quick("fit")
    da, db = (int(v) for v in args.dagger_seeds.split(":"))
    for rnd in range(args.dagger_rounds):
        roll = cl.rollout_pairs(model, range(da, db), steps=args.fine_steps, pullback=args.pullback, log=log)
        rt = cl.rollout_pairs(model, test_seeds[:2], starts=4, steps=args.fine_steps, pullback=args.pullback)
        log(f"[closure] aggregation round {rnd}: {roll[0].shape[0]} rollout pairs; "
            f"holdout rollout loss {cl.eval_loss(model, rt, dx, dy):.4f}")
        data = cl.join(data, roll)
        del roll
        model, h = cl.train(data, dx, dy, iters=args.dagger_iters, batch=args.batch, lr=args.dagger_lr, log=log,
                            holdout=rt, model=model)
        hist += [dict(r, round=rnd + 1) for r in h]
        model.save(args.out)
        quick(f"aggregation {rnd + 1}")
    t_train = time.time() - t1
    rep = {"train_seeds": args.train, "pairs_per_run": args.pairs, "test_seeds": test_seeds,
           "iters": args.iters, "fields": args.fields, "dagger_rounds": args.dagger_rounds, "dagger_seeds": args.dagger_seeds,
           "dagger_iters": args.dagger_iters, "dagger_lr": args.dagger_lr, "batch": args.batch, "lr": args.lr, "data_s": t_data, "train_s": t_train,
           "history": hist, "phases": phases, "pullback": args.pullback, "holdout_loss": cl.eval_loss(model, test, dx, dy),
           "holdout_loss_plain": cl.baseline_loss(test[1], model.stats)}
    log(f"[closure] holdout loss {rep['holdout_loss']:.4f} vs plain {rep['holdout_loss_plain']:.4f}")
    calibrate(model, rep)
    final_evaluation(rep, model, eval_seeds, variants, args)
    log(f"[closure] wrote {args.out} and {args.report}")


if __name__ == "__main__":
    main()
