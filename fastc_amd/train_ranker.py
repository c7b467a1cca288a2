"""Train the BC7 candidate ranker on the GPU and save its weights.

    python -m fastc_amd.train_ranker [--blocks 4194304] [--epochs 5] [--batch 4096]
                                     [--lr 1e-3] [--seed 5] [--out fastc_amd/data/bc7_ranker.npy]

Training set: `--blocks` blocks that reach the BC7 search (neither solid nor
fully transparent) from the seeded generator (seed 5), labelled on the GPU
by the exhaustive quality-0 search (fh_bc7_exhaustive_labels: the (mode,
slot) candidate that decodes closest to the block).  The network is trained
by the library's own forward/backward/Adam kernels (fh_ranker_train).  A
held-out set from seed 6 measures how often the label is the top-1 / within
the top-4 of the trained ranking.  Prints one JSON summary line.
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


def labels_on_device(fh, torch, d_blocks, n, chunk=1 << 20):
    d_lab = torch.empty(n, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for c0 in range(0, n, chunk):
        cn = min(chunk, n - c0)
        fh.exhaustive_labels_device(d_blocks[c0:].data_ptr(), cn, d_lab[c0:].data_ptr(), 0,
                                    stream)
    return d_lab


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=1 << 22)
    ap.add_argument("--epochs", type=int, default=5)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
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
    d_lab = labels_on_device(fh, torch, d_blocks, args.blocks)
    d_hold_lab = labels_on_device(fh, torch, d_hold, args.holdout)
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
    hl = d_hold_lab.cpu().numpy()
    top4 = np.argsort(-z, axis=1, kind="stable")[:, :4]
    zz = z - z.max(1, keepdims=True)
    pmax = 1.0 / np.exp(zz).sum(1)
    summary = {
        "blocks": args.blocks, "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
        "steps": rep["steps"], "train_s": rep["ms"] / 1e3, "label_s": t_lab,
        "first_loss": rep["first_loss"], "epoch_loss": rep["epoch_loss"],
        "holdout_top1": float(np.mean(top4[:, 0] == hl)),
        "holdout_top4": float(np.mean((top4 == hl[:, None]).any(1))),
        "holdout_confident_frac_0.6": float(np.mean(pmax >= 0.6)),
    }
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, rk.get_params())
    summary["out"] = os.path.relpath(args.out, os.path.dirname(HERE))
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
