"""Train the BC7 candidate ranker on the GPU and save its weights.

    python -m fastc_amd.train_ranker [--blocks 4194304] [--epochs 5] [--batch 4096]
                                     [--lr 1e-3] [--margin 0.5] [--seed 5]
                                     [--out fastc_amd/data/bc7_ranker.npy]

Training set: `--blocks` blocks that reach the BC7 search (neither solid nor
fully transparent) from the seeded generator (seed 5), labelled on the GPU
by the exhaustive quality-0 search: every one of the 285 (mode, slot)
candidates is fitted, packed and decoded (fh_bc7_candidate_errors).  The
label is a canonical near-best candidate: among the candidates that decode
within `--margin` of the block's best error, the one most often near-best
over the whole training set (a global preference order, refined by the
label counts it produces).  With margin 0 this is the exhaustive arg-min
(fh_bc7_exhaustive_labels).  A margin lets equally good candidates share
one label, so the network can be confident where the exact arg-min is a
coin toss; the search still picks among the top-k by decoded error.

The network is trained by the library's own forward/backward/Adam kernels
(fh_ranker_train).  A held-out set from seed 6 measures the confident share
(top softmax probability of at least 0.6), how often the top-4 holds a
candidate within 2% of the best, and the decoded error of the best of the
top-4 against the best of all.  Prints one JSON summary line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "bc7_ranker.npy")


def searched_blocks(n, seed, width=8192):
    """The first n searched blocks of generator images of `width` columns."""
    from fastc_amd import synth
    rows_needed = int(np.ceil(n / (width // 4) / 0.85)) + 1
    img = synth.generate(width, 4 * rows_needed, seed)
    blocks = synth.blocks(img)
    keep = ~(blocks == blocks[:, :1]).all(axis=1) & ((blocks >> 24) != 0).any(axis=1)
    blocks = blocks[keep]
    if blocks.shape[0] < n:
        raise RuntimeError("generator image too small for the requested block count")
    return np.ascontiguousarray(blocks[:n])


def candidate_errors(fh, torch, d_blocks, n, chunk=1 << 20):
    """(n, 512) float32 decoded errors of every candidate (+inf: no candidate)."""
    out = torch.empty((n, 512), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for c0 in range(0, n, chunk):
        cn = min(chunk, n - c0)
        fh.candidate_errors_device(d_blocks[c0:].data_ptr(), cn, out[c0:].data_ptr(), stream)
    return out


def canonical_labels(torch, errs, margin, iters=4, chunk=1 << 20):
    """Canonical near-best labels (see the module text): returns (labels,
    preference order).  Near-best: error <= best * (1 + margin)."""
    n = errs.shape[0]
    best = errs.min(1, keepdim=True).values
    freq = torch.zeros(512, device=errs.device)
    for c0 in range(0, n, chunk):
        near = errs[c0:c0 + chunk] <= best[c0:c0 + chunk] * (1 + margin)
        freq += near.float().sum(0)
    pref = torch.argsort(-freq, stable=True)
    labels = torch.empty(n, dtype=torch.int64, device=errs.device)
    for _ in range(iters):
        rank = torch.empty(512, dtype=torch.float32, device=errs.device)
        rank[pref] = torch.arange(512, dtype=torch.float32, device=errs.device)
        for c0 in range(0, n, chunk):
            near = errs[c0:c0 + chunk] <= best[c0:c0 + chunk] * (1 + margin)
            score = torch.where(near, rank[None, :], torch.full_like(rank[None, :], 1e9))
            labels[c0:c0 + chunk] = score.argmin(1)
        count = torch.bincount(labels, minlength=512).float()
        pref = torch.argsort(-(count + freq * 1e-6), stable=True)
    return labels.to(torch.int32), pref


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=1 << 22)
    ap.add_argument("--epochs", type=int, default=5)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--margin", type=float, default=0.75,
                    help="near-best margin of the canonical labels (0: exact arg-min)")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--holdout", type=int, default=1 << 16)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()

    import torch  # its HIP runtime first (it shares the soname with ROCm's)
    sys.path.insert(0, os.path.dirname(HERE))
    import fastc_amd as fh

    fh.lib()
    if not torch.cuda.is_available() or not fh.device_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    t0 = time.perf_counter()
    blocks = searched_blocks(args.blocks, args.seed)
    hold = searched_blocks(args.holdout, args.seed + 1, width=2048)
    t_gen = time.perf_counter() - t0
    print(f"[train] {blocks.shape[0]} training blocks, {hold.shape[0]} held out "
          f"({t_gen:.1f} s)", file=sys.stderr, flush=True)

    d_blocks = torch.from_numpy(blocks.view(np.int32)).cuda()
    d_hold = torch.from_numpy(hold.view(np.int32)).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    errs = candidate_errors(fh, torch, d_blocks, args.blocks)
    d_lab, pref = canonical_labels(torch, errs, args.margin)
    del errs
    hold_errs = candidate_errors(fh, torch, d_hold, args.holdout)
    torch.cuda.synchronize()
    t_lab = time.perf_counter() - t0
    lab = d_lab.cpu().numpy()
    print(f"[train] labels in {t_lab:.1f} s; modes "
          f"{np.bincount(lab >> 6, minlength=8).tolist()}", file=sys.stderr, flush=True)

    rk = fh.Ranker(seed=args.init_seed)
    rep = rk.train(d_blocks.data_ptr(), d_lab.data_ptr(), args.blocks, args.epochs, args.batch,
                   args.lr, shuffle_seed=args.seed)
    print(f"[train] {rep['steps']} steps in {rep['ms'] / 1e3:.1f} s, loss "
          f"{rep['first_loss']:.3f} -> {rep['epoch_loss']}", file=sys.stderr, flush=True)

    # held-out ranking quality
    n = args.holdout
    d_log = torch.empty((n, 512), dtype=torch.float32, device="cuda")
    rk.forward(d_hold.data_ptr(), n, d_log.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    z = d_log.cpu().numpy()
    valid = np.zeros(512, bool)
    for m, k in enumerate([16, 64, 64, 64, 8, 4, 1, 64]):
        valid[64 * m:64 * m + k] = True
    z = np.where(valid, z, -np.inf)
    he = hold_errs.cpu().numpy().astype(np.float64)
    hbest = he.min(1)
    top4 = np.argsort(-z, axis=1, kind="stable")[:, :4]
    zz = z - z.max(1, keepdims=True)
    pmax = 1.0 / np.exp(zz).sum(1)
    e4 = np.take_along_axis(he, top4, 1).min(1)
    conf = pmax >= 0.6
    summary = {
        "blocks": args.blocks, "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
        "margin": args.margin,
        "steps": rep["steps"], "train_s": rep["ms"] / 1e3, "label_s": t_lab,
        "first_loss": rep["first_loss"], "epoch_loss": rep["epoch_loss"],
        "holdout_top1_is_best": float(np.mean(he[np.arange(len(he)), top4[:, 0]] == hbest)),
        "holdout_top4_within_2pct": float(np.mean(e4 <= hbest * 1.02)),
        "holdout_top4_error_ratio": float(e4.sum() / hbest.sum()),
        "holdout_confident_frac_0.6": float(np.mean(conf)),
        "holdout_top4_error_ratio_confident": float(e4[conf].sum() / hbest[conf].sum())
        if conf.any() else None,
    }
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, rk.get_params())
    summary["out"] = os.path.relpath(args.out, os.path.dirname(HERE))
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
