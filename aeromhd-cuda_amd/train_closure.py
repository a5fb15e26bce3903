"""Train and judge the learned flux closure on the GPU.

    python aeromhd-cuda_amd/train_closure.py --train 200 --test 200:220 \
        --iters 3000 --out aeromhd-cuda_amd/closure_weights.npz --report closure_report.json

Data: fine 2048x1024 magnetic-sieve runs (FP32, 2000 steps) for seeds
0..train-1 restricted 8x to 256x128; held-out seeds for the validation
loss and the fine-grid judge.  The judge also times the plain FP64
fine-grid solve of one seed for the speed comparison."""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def log(*a):
    print(*a, flush=True)


def fp64_fine_throughput(seed: int, steps: int):
    import torch
    from amhd_ics import magnetic_sieve
    from physics import prim_to_cons
    from solver import MHDSolver

    cfg, W = magnetic_sieve(2048, 1024, precision="fp64", seed=seed, max_steps=steps)
    cfg.t_end = 1e30
    s = MHDSolver(cfg)
    s.U = prim_to_cons(W, cfg.gamma)
    s.advance(5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = s.advance(steps - 5)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"mcups": 2048 * 1024 * n / wall / 1e6, "wall_s_for_run": wall * steps / n, "steps": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200)
    ap.add_argument("--test", default="200:220")
    ap.add_argument("--pairs", type=int, default=16)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--fine-steps", type=int, default=2000)
    ap.add_argument("--judge", default="")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "closure_weights.npz"))
    ap.add_argument("--report", default="closure_report.json")
    args = ap.parse_args()
    import amhd_closure as cl

    a, b = (int(v) for v in args.test.split(":"))
    test_seeds = list(range(a, b))
    judge_seeds = test_seeds if not args.judge else list(range(*(int(v) for v in args.judge.split(":"))))
    t0 = time.time()
    X, R = cl.generate_pairs(range(args.train), args.pairs, steps=args.fine_steps, log=log)
    Xt, Rt = cl.generate_pairs(test_seeds, 4, steps=args.fine_steps, log=log)
    t_data = time.time() - t0
    dx, dy = 6.0 / 256, 2.0 / 128
    log(f"[closure] data: train {tuple(X.shape)}, test {tuple(Xt.shape)}, {t_data:.0f} s")
    t1 = time.time()
    model, hist = cl.train(X, R, dx, dy, iters=args.iters, batch=args.batch, lr=args.lr, log=log, holdout=(Xt, Rt))
    t_train = time.time() - t1
    model.save(args.out)
    rep = {"train_seeds": args.train, "pairs_per_run": args.pairs, "test_seeds": test_seeds,
           "iters": args.iters, "batch": args.batch, "lr": args.lr, "data_s": t_data, "train_s": t_train,
           "history": hist, "holdout_loss": cl.eval_loss(model, Xt, Rt, dx, dy),
           "holdout_loss_plain": cl.baseline_loss(Rt, model.stats)}
    log(f"[closure] holdout loss {rep['holdout_loss']:.4f} vs plain {rep['holdout_loss_plain']:.4f}")
    rows = cl.judge(judge_seeds, model, steps=args.fine_steps, log=log)
    rep["judge"] = rows
    rep["fp64_fine"] = fp64_fine_throughput(judge_seeds[0], args.fine_steps)
    fine_cells = 2048 * 1024 * args.fine_steps
    rep["closure_fine_equiv_mcups"] = [fine_cells / r["wall_closure_s"] / 1e6 for r in rows]
    json.dump(rep, open(args.report, "w"), indent=1, default=float)
    log(f"[closure] wrote {args.out} and {args.report}")


if __name__ == "__main__":
    main()
