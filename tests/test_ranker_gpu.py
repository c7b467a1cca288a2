"""The candidate ranker's hand-written kernels (bf16 matrix-core products,
softmax cross-entropy, Adam) against the CPU MLP oracle: forward logits,
gradients of every parameter block, one Adam step, and a short training run
that fits labels from the exhaustive GPU search."""
import numpy as np
import pytest

from fastc_amd import synth

pytestmark = pytest.mark.gpu

# parameter blocks: (name, offset, size)
BLOCKS = [("W1", 0, 64 * 256), ("b1", 16384, 256), ("W2", 16640, 65536), ("b2", 82176, 256),
          ("W3", 82432, 131072), ("b3", 213504, 512)]


def _searched_blocks(n, seed):
    side = 4 * int(np.ceil(np.sqrt(n * 1.3)))
    blocks = synth.blocks(synth.generate(side, side, seed))
    keep = ~(blocks == blocks[:, :1]).all(axis=1) & ((blocks >> 24) != 0).any(axis=1)
    return np.ascontiguousarray(blocks[keep][:n])


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(214016, np.float32)
    for name, off, size in BLOCKS:
        fan_in = {"W1": 64, "W2": 256, "W3": 256}.get(name)
        if fan_in:
            p[off:off + size] = rng.uniform(-1, 1, size) * np.sqrt(6 / fan_in)
        else:
            p[off:off + size] = rng.standard_normal(size) * 0.05
    return p


def _dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(dtype)).cuda()


def test_forward_matches_oracle(fh, oracle):
    import torch
    blocks = _searched_blocks(256, 7)
    n = blocks.shape[0]
    p = _params(1)
    rk = fh.Ranker()
    rk.set_params(p)
    d_blk = _dev(blocks, np.int32)
    d_log = torch.zeros((n, 512), dtype=torch.float32, device="cuda")
    print(f"ranker forward n={n}", flush=True)
    rk.forward(d_blk.data_ptr(), n, d_log.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = d_log.cpu().numpy()
    x = oracle.mlp_features(blocks)
    _, _, want = oracle.mlp_forward(p, x, round_bf16=True)
    scale = np.abs(want).max() + 1.0
    # bf16 hidden activations can round one ulp apart (fp32 vs double sums)
    assert np.abs(got - want).max() <= 1e-2 * scale
    valid = oracle.mlp_valid_mask()
    top_got = np.argmax(np.where(valid, got, -np.inf), 1)
    top_want = np.argmax(np.where(valid, want, -np.inf), 1)
    assert np.mean(top_got == top_want) > 0.95


def test_gradients_match_oracle(fh, oracle):
    blocks = _searched_blocks(384, 8)
    n = blocks.shape[0]
    rng = np.random.default_rng(3)
    labels = rng.choice(np.nonzero(oracle.mlp_valid_mask())[0], n).astype(np.int32)
    p = _params(2)
    rk = fh.Ranker()
    rk.set_params(p)
    d_blk = _dev(blocks, np.int32)
    d_lab = _dev(labels, np.int32)
    print(f"ranker gradients n={n}", flush=True)
    loss, g = rk.gradients(d_blk.data_ptr(), d_lab.data_ptr(), n)
    x = oracle.mlp_features(blocks)
    wl, wg = oracle.mlp_loss_grad(p, x, labels, round_bf16=True)
    assert abs(loss - wl) <= 1e-3 * abs(wl)
    for name, off, size in BLOCKS:
        a, b = g[off:off + size], wg[off:off + size]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        print(f"  {name}: relative error {rel:.2e}", flush=True)
        assert rel < 2e-2, name


def test_adam_step_and_training(fh, oracle):
    import torch
    blocks = _searched_blocks(4096, 9)
    n = blocks.shape[0]
    d_blk = _dev(blocks, np.int32)
    d_lab = torch.zeros(n, dtype=torch.int32, device="cuda")
    fh.exhaustive_labels_device(d_blk.data_ptr(), n, d_lab.data_ptr(), 0,
                                torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    labels = d_lab.cpu().numpy()
    assert oracle.mlp_valid_mask()[labels].all()
    # one full-batch Adam step equals the oracle's Adam on the oracle's
    # gradient, to bf16 accuracy
    p = _params(4)
    rk = fh.Ranker()
    rk.set_params(p)
    print("ranker one step", flush=True)
    rk.train(d_blk.data_ptr(), d_lab.data_ptr(), n, epochs=1, batch=n, lr=1e-3)
    got = rk.get_params()
    x = oracle.mlp_features(blocks)
    _, wg = oracle.mlp_loss_grad(p, x, labels, round_bf16=True)
    w = p.copy()
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    oracle.adam_step(w, m, v, wg, 1, lr=1e-3)
    # Adam's first step is lr * sign(g) wherever |g| >> eps: compare signs
    # where the gradient is clearly above the bf16 noise
    clear = np.abs(wg) > 0.05 * np.sqrt(np.mean(wg ** 2))
    agree = np.mean(np.sign(got - p)[clear] == np.sign(w - p)[clear])
    print(f"  sign agreement {agree:.4f} over {clear.sum()} weights", flush=True)
    assert agree > 0.95
    assert np.abs(got - w).max() <= 2.1e-3
    # a short run fits the labels
    rk.init(5)
    print("ranker training", flush=True)
    rep = rk.train(d_blk.data_ptr(), d_lab.data_ptr(), n, epochs=30, batch=512, lr=2e-3,
                   shuffle_seed=1)
    print(rep, flush=True)
    assert rep["steps"] == 30 * (n // 512)
    assert rep["epoch_loss"][-1] < 0.7 * rep["first_loss"]
