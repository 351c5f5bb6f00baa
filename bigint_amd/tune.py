"""Autotune the plan table and train the MLP cost model on this GPU, then measure regret.

  python -m bigint_amd.tune --train 7000 --heldout 3500 --out bigint_amd/plans/planner.json \
         --report gpurun_out/planner_report.json

1. Draws training shapes (operand sizes log-uniform over [2^10, 2^28] bits, batch sizes
   log-uniform in [1, 64]) and times every allowed plan of each (bigint_amd.planner.Timer:
   planned device batches, HIP events on a dedicated stream).
2. Builds the autotuned table from the training timings (per cell the best fixed plan, or the
   built-in heuristic where it does better).
3. Trains the MLP on the GPU (bi_hip_mlp_grad / bi_hip_mlp_adam; full-batch Adam, 40,000 epochs
   with a cosine learning-rate decay by default).
4. Times every allowed plan on held-out shapes and reports, for the plain plan (0), the table and
   the MLP: mean regret (t_chosen / t_best - 1), geometric-mean speedup over plain, and the
   total-time ratio plain / chosen.  Exactness is independent of the plan (tests/test_planner.py).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from bigint_amd import engine  # noqa: E402
from bigint_amd import planner as P  # noqa: E402


def log(msg: str) -> None:
    print(f"[tune {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def mlp_alone(keys, params, norm):
    """The model's own argmin over the allowed plans (the learned mode without its table prior):
    forward pass on the GPU (bi_hip_mlp_forward), plus the log2 transform-count cost per plan."""
    from bigint_amd import hip
    X = np.array([P.features(28 * la, 28 * lb, k) for la, lb, k in keys], np.float32)
    Xn = ((X - norm[:3]) / norm[3:]).astype(np.float32)
    dx, dp = hip.Buffer.from_array(Xn), hip.Buffer.from_array(np.asarray(params, np.float32))
    dy = hip.Buffer(4 * len(keys) * P.K)
    st = hip.Stream()
    rc = engine.lib().bi_hip_mlp_forward(dx.ptr, len(keys), dp.ptr, P.K, dy.ptr, st.handle)
    if rc:
        raise engine.EngineError(rc, "bi_hip_mlp_forward")
    st.synchronize()
    Y = dy.download(np.float32, len(keys) * P.K).reshape(len(keys), P.K)
    out = []
    for i, (la, lb, k) in enumerate(keys):
        cand = [(Y[i][p] + np.log2(P.plan_cost(la, lb, p)), p) for p in P.allowed(la, lb)]
        out.append(min(cand)[1])
    return out


def evaluate(recs_h, table, params, norm):
    """Regret of plain / heuristic / table / learned choices on held-out records; `learned` is the
    engine's learned mode (the model with the table's plan as its prior), `mlp_alone` the model's
    own argmin."""
    shapes = P.group_best(recs_h)
    keys = list(shapes)
    P.install(table, params, norm)
    ch_table = P.choose([(la, lb, k) for la, lb, k in keys], engine.PLAN_TABLE)
    ch_mlp = P.choose([(la, lb, k) for la, lb, k in keys], engine.PLAN_LEARNED)
    ch_heur = P.choose([(la, lb, k) for la, lb, k in keys], engine.PLAN_FIXED)  # built-in heuristic
    ch_alone = mlp_alone(keys, params, norm)
    out = {}
    for name, ch in [("plain", [0] * len(keys)), ("heuristic", ch_heur), ("table", ch_table),
                     ("mlp_alone", ch_alone), ("mlp", ch_mlp)]:
        reg, sp, tot_c, tot_p, hit = [], [], 0.0, 0.0, 0
        for key, c in zip(keys, ch):
            plans = shapes[key]
            best = min(plans.values())
            t = plans.get(c, plans[0])
            reg.append(t / best - 1)
            sp.append(plans[0] / t)
            tot_c += t * key[2]
            tot_p += plans[0] * key[2]
            hit += int(t == best)
        out[name] = {"mean_regret": float(np.mean(reg)), "p95_regret": float(np.percentile(reg, 95)),
                     "gmean_speedup_vs_plain": float(np.exp(np.mean(np.log(sp)))),
                     "total_time_speedup_vs_plain": tot_p / tot_c, "best_plan_rate": hit / len(keys)}
    oracle_tot = sum(min(v.values()) * k[2] for k, v in shapes.items())
    out["best_possible_total_speedup_vs_plain"] = sum(v[0] * k[2] for k, v in shapes.items()) / oracle_tot
    out["shapes"] = len(keys)
    out["runs"] = len(recs_h)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=7000, help="training shapes")
    ap.add_argument("--heldout", type=int, default=3500, help="held-out shapes")
    ap.add_argument("--epochs", type=int, default=40000)
    ap.add_argument("--lr", type=float, default=3e-3, help="Adam learning rate (start)")
    ap.add_argument("--lr-end", type=float, default=3e-5,
                    help="final learning rate of the cosine decay (negative: constant rate)")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--out", default=P.DEFAULT_PATH)
    ap.add_argument("--report", default=None)
    ap.add_argument("--records", default=None, help="save raw timing records (JSON lines)")
    ap.add_argument("--from-records", default=None,
                    help="train and evaluate on saved records (JSON lines with a 'split' field) "
                         "instead of timing new shapes")
    ap.add_argument("--evaluate", default=None,
                    help="with --from-records: evaluate this planner file (table and model as "
                         "committed) on the held-out records, without training")
    args = ap.parse_args()

    engine.init(0)
    timer = P.Timer()
    max_out = timer.out_words - 64
    def shapes(n, seed):
        sh = P.sample_shapes(n, seed)
        res = []
        for la, lb, k in sh:
            while k > 1 and k * (la + lb) > max_out:
                k //= 2
            if k * (la + lb) <= max_out and la <= timer.max_limbs and lb <= timer.max_limbs:
                res.append((la, lb, k))
        return res

    if args.evaluate:
        if not args.from_records:
            ap.error("--evaluate needs --from-records")
        with open(args.from_records) as f:
            recs_h = [r for r in (json.loads(line) for line in f if line.strip()) if r.get("split") == "heldout"]
        table, params, norm, _ = P.load(args.evaluate, do_install=False)
        ev = evaluate(recs_h, table, params, norm)
        log(json.dumps(ev))
        if args.report:
            with open(args.report, "w") as f:
                json.dump({"evaluation": ev, "planner": os.path.basename(args.evaluate)}, f, indent=1)
        return 0

    t0 = time.time()
    if args.from_records:
        with open(args.from_records) as f:
            allrecs = [json.loads(line) for line in f if line.strip()]
        recs = [r for r in allrecs if r.get("split") == "train"]
        saved_heldout = [r for r in allrecs if r.get("split") == "heldout"]
        train_shapes = sorted({(r["la"], r["lb"], r["k"]) for r in recs})
    else:
        train_shapes = shapes(args.train, args.seed)
        recs = P.collect(train_shapes, timer, log)
    log(f"training timings: {len(recs)} runs over {len(train_shapes)} shapes in {time.time() - t0:.0f}s")
    keys = sorted(P.group_best(recs))
    heur = dict(zip(keys, P.choose(keys, engine.PLAN_FIXED)))  # the built-in heuristic's choices
    table = P.build_table(recs, heur)
    X, T, M = P.dataset(recs)
    t1 = time.time()
    params, norm, hist = P.train_gpu(X, T, M, epochs=args.epochs, lr=args.lr, seed=args.seed, log=log,
                                     lr_end=args.lr_end if args.lr_end >= 0 else None)
    log(f"MLP trained on the GPU in {time.time() - t1:.1f}s, final loss {hist[-1][1]:.5f}")

    t2 = time.time()
    if args.from_records:
        recs_h = saved_heldout
    else:
        held = shapes(args.heldout, args.seed + 1)
        recs_h = P.collect(held, timer, log)
    log(f"held-out timings: {len(recs_h)} runs in {time.time() - t2:.0f}s")
    ev = evaluate(recs_h, table, params, norm)
    log(json.dumps(ev))
    meta = {"train_runs": len(recs), "train_shapes": len(train_shapes), "heldout": ev,
            "loss_history": hist, "epochs": args.epochs, "lr": args.lr, "lr_end": args.lr_end}
    P.save(args.out, table, params, norm, meta)
    if args.report:
        with open(args.report, "w") as f:
            json.dump({"evaluation": ev, "train_runs": len(recs), "loss_history": hist}, f, indent=1)
    if args.records:
        with open(args.records, "w") as f:
            for r in recs:
                f.write(json.dumps(dict(r, split="train")) + "\n")
            for r in recs_h:
                f.write(json.dumps(dict(r, split="heldout")) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
