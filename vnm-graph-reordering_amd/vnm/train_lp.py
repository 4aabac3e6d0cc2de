"""Train the learned pair proposer on exact swap traces, on the GPU.

    python -m vnm.train_lp --out DIR [--graphs 200] [--epochs 8] ...

Data (all seeded, nothing downloaded):
  graphs   symmetric R-MAT, (a, b, c) = (0.57, 0.19, 0.19), average degree 16,
           2^14 .. 2^16 nodes (the scale drawn uniformly, so node counts are
           log-uniform on [16384, 65536]), graph i uses seed 1000 + i.
  states   for each graph, `--states` points along a reorder trajectory: the
           input, then the state after each further `--advance` iterations of
           the exhaustive-window proposer reorder (mode 2).
  labels   at each state the exact reorderer (GetSegSwapScore greedy,
           schedule 0) is traced for `--trace-rounds` rounds on a copy; its
           (primary, target) segment pairs, plus `--rand-pairs` pairs of an
           invalid segment with a random other segment, are scored exactly
           (vnm_pair_score); every column pair (u, v) of each segment pair is
           a training pair with target sign(g) * log1p(|g|), g = lgain + rgain.
Training: Adam through vnm_lp_train_step (BF16 forward on MFMA, FP32
backward, the library's own kernels), cosine learning-rate decay, one graph
state per step with up to `--batch` sampled pairs.  The last `--val` graphs
are held out: reported are their loss and how much more often the model's
top-5% pairs have a positive exact gain than pairs on average.

Writes DIR/lp_rmat_deg16.f32 (22784 little-endian FP32 parameters) and
DIR/lp_rmat_deg16.json (settings, per-epoch losses, validation metrics).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import vnm  # noqa: E402

N, M = 2, 8


def target_of(gain):
    g = gain.astype(np.float64)
    return (np.sign(g) * np.log1p(np.abs(g))).astype(np.float32)


def label_state(g, rng, trace_rounds, rand_pairs):
    """Training pairs (current ids) and targets at the graph's current state."""
    ns = g.nseg
    c = g.clone()
    _, tr = c.nm_trace(N, 0, 10, round_cap=trace_rounds, cap=4096)
    c.close()
    sp = set()
    for rec in tr:
        if rec[1] != rec[2]:
            sp.add((int(rec[1]), int(rec[2])))
    cnt, _ = g.locate(N)
    inval = np.nonzero(cnt > 0)[0]
    if inval.size:
        for p in rng.choice(inval, size=rand_pairs):
            q = int(rng.integers(0, ns - 1))
            q = q + 1 if q >= p else q
            sp.add((int(p), q))
    if not sp:
        return None
    segp = np.array([a for a, _ in sorted(sp)], np.int32)
    segq = np.array([b for _, b in sorted(sp)], np.int32)
    sc = g.pair_score(N, segp, segq).reshape(len(segp), M * M, 2)
    gain = sc[:, :, 0] + sc[:, :, 1]
    b1, b2 = np.divmod(np.arange(M * M), M)
    u = (segp[:, None] * M + b1[None, :]).ravel()
    v = (segq[:, None] * M + b2[None, :]).ravel()
    pairs = np.stack([u, v], 1).astype(np.int32)
    return pairs, target_of(gain.ravel()), gain.ravel().astype(np.int32), len(tr)


def build_dataset(ctx, a, log):
    rng = np.random.default_rng(a.seed)
    examples = []  # (graph index, host COO, pairs, targets, gains)
    t0 = time.time()
    for i in range(a.graphs):
        scale = int(rng.integers(14, 17))
        n = 1 << scale
        coo = vnm.gen_rmat(scale, n * 16 // 2, 1000 + i, M=M)
        g = vnm.Graph(ctx, coo, M)
        for s in range(a.states):
            lab = label_state(g, rng, a.trace_rounds, a.rand_pairs)
            if lab is not None:
                examples.append((i, g.download(), lab[0], lab[1], lab[2]))
            if s + 1 < a.states:
                g.lp_reorder(None, N, mode=vnm.LP_EXHAUSTIVE, topk=1024, window=8, maxiter=a.advance,
                             seed=7 + i)
        g.close()
        if (i + 1) % 20 == 0:
            log(f"[train_lp] data: {i + 1}/{a.graphs} graphs, {len(examples)} states, "
                f"{time.time() - t0:.1f} s")
    return examples


def evaluate(ctx, model, examples):
    losses, lifts = [], []
    for _, coo, pairs, tg, gain in examples:
        g = vnm.Graph(ctx, coo, M)
        loss, _ = model.loss_grad(g, N, pairs, tg, bf16=True, want_grad=False)
        Y = model.forward(g, N, bf16=True)
        g.close()
        u, v = pairs[:, 0], pairs[:, 1]
        s = Y[u, 0] + Y[v, 1] + np.einsum("ij,ij->i", Y[u, 2:18], Y[v, 18:34])
        k = max(1, len(s) // 20)
        top = np.argsort(-s, kind="stable")[:k]
        base = (gain > 0).mean()
        lifts.append(((gain[top] > 0).mean() / base) if base > 0 else 1.0)
        losses.append(loss)
    return float(np.mean(losses)), float(np.mean(lifts))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--graphs", type=int, default=200)
    ap.add_argument("--val", type=int, default=20)
    ap.add_argument("--states", type=int, default=3)
    ap.add_argument("--advance", type=int, default=2)
    ap.add_argument("--trace-rounds", type=int, default=2)
    ap.add_argument("--rand-pairs", type=int, default=96)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--steps-per-state", type=int, default=2)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)

    def log(msg):
        print(msg, flush=True)

    t_start = time.time()
    ctx = vnm.Context(0)
    data = build_dataset(ctx, a, log)
    ntrain_graphs = a.graphs - a.val
    train = [e for e in data if e[0] < ntrain_graphs]
    val = [e for e in data if e[0] >= ntrain_graphs]
    log(f"[train_lp] {len(train)} training states, {len(val)} validation states, "
        f"{sum(len(e[3]) for e in train)} training pairs")
    model = vnm.LPModel(ctx, seed=a.seed)
    v0 = evaluate(ctx, model, val)
    log(f"[train_lp] untrained: val loss {v0[0]:.4f}, top-5% positive-gain lift {v0[1]:.2f}")
    rng = np.random.default_rng(a.seed + 1)
    total = a.epochs * len(train) * a.steps_per_state
    step = 0
    hist = []
    t_train = time.time()
    for ep in range(a.epochs):
        ls = []
        for idx in rng.permutation(len(train)):
            _, coo, pairs, tg, _ = train[idx]
            g = vnm.Graph(ctx, coo, M)
            for _ in range(a.steps_per_state):
                sel = rng.choice(len(tg), size=min(a.batch, len(tg)), replace=False)
                lr = a.lr * 0.5 * (1 + math.cos(math.pi * step / max(total, 1)))
                ls.append(model.train_step(g, N, pairs[sel], tg[sel], max(lr, 1e-6)))
                step += 1
            g.close()
        vl = evaluate(ctx, model, val)
        hist.append({"epoch": ep + 1, "train_loss": float(np.mean(ls)), "val_loss": vl[0],
                     "val_top5_lift": vl[1]})
        log(f"[train_lp] epoch {ep + 1}: train {np.mean(ls):.4f} val {vl[0]:.4f} lift {vl[1]:.2f} "
            f"({time.time() - t_train:.1f} s)")
    path = os.path.join(a.out, "lp_rmat_deg16.f32")
    model.save(path)
    meta = {"settings": vars(a), "train_states": len(train), "val_states": len(val),
            "train_pairs": int(sum(len(e[3]) for e in train)), "steps": step,
            "untrained_val_loss": v0[0], "untrained_val_top5_lift": v0[1], "epochs": hist,
            "train_seconds": time.time() - t_train, "total_seconds": time.time() - t_start}
    json.dump(meta, open(os.path.join(a.out, "lp_rmat_deg16.json"), "w"), indent=1)
    log(f"[train_lp] wrote {path}")


if __name__ == "__main__":
    main()
