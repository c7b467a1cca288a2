"""Pins for the ranking-MLP CPU oracle (no GPU): finite differences of its
own gradients in exact arithmetic, softmax cross-entropy closed forms and
Adam's known first step and convergence."""
import numpy as np
import pytest


def _params(seed, scale=0.1):
    from oracle import pyoracle as po
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(po.MLP_NPARAMS) * scale).astype(np.float32)


def test_valid_mask(oracle):
    m = oracle.mlp_valid_mask()
    # mode 0: 16 shapes, modes 1-3 and 7: 64, mode 4: 8 (rotation x index
    # mode), mode 5: 4 rotations, mode 6: 1
    assert m.sum() == 16 + 4 * 64 + 8 + 4 + 1
    assert m[0:16].all() and not m[16:64].any() and m[448:512].all()


def test_xent_closed_forms(oracle):
    B = 3
    z = np.zeros((B, 512), np.float32)
    labels = np.array([0, 70, 448], np.int32)
    loss, d = oracle.mlp_xent(z, labels)
    assert abs(loss - np.log(285)) < 1e-9
    mask = oracle.mlp_valid_mask()
    assert np.allclose(d.sum(1), 0, atol=1e-7)
    assert not d[:, ~mask].any()
    assert np.allclose(d[np.arange(B), labels], (1 / 285 - 1) / B, atol=1e-7)
    # invalid logits never take probability, however large
    z2 = z.copy()
    z2[:, 20] = 1e4
    assert abs(oracle.mlp_xent(z2, labels)[0] - loss) < 1e-9


def _np_loss(p, x, labels, mask):
    """The network's definition in float64: 64 -> ReLU 256 -> ReLU 256 ->
    512 logits, softmax cross-entropy over the valid logits."""
    o, w = 0, []
    for shp in [(64, 256), (256,), (256, 256), (256,), (256, 512), (512,)]:
        n = int(np.prod(shp))
        w.append(p[o:o + n].reshape(shp))
        o += n
    W1, b1, W2, b2, W3, b3 = w
    h1 = np.maximum(x @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    z = np.where(mask, h2 @ W3 + b3, -np.inf)
    m = z.max(1, keepdims=True)
    lse = np.log(np.exp(z - m).sum(1)) + m[:, 0]
    return float((lse - z[np.arange(len(labels)), labels]).mean())


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_finite_differences(oracle, seed):
    """Exact-arithmetic mode (no bf16 rounding): the loss equals the float64
    definition and every parameter block's gradient matches central
    differences of it."""
    rng = np.random.default_rng(seed)
    p = _params(seed, 0.2)
    x = rng.random((3, 64)).astype(np.float32)
    mask = oracle.mlp_valid_mask()
    labels = rng.choice(np.nonzero(mask)[0], 3).astype(np.int32)
    loss, g = oracle.mlp_loss_grad(p, x, labels, round_bf16=False)
    p64, x64 = p.astype(np.float64), x.astype(np.float64)
    assert abs(loss - _np_loss(p64, x64, labels, mask)) <= 1e-5 * abs(loss)
    offs = [0, 64 * 256, 64 * 256 + 256, 64 * 256 + 256 + 65536, 64 * 256 + 512 + 65536,
            64 * 256 + 512 + 65536 + 131072]
    sizes = [64 * 256, 256, 65536, 256, 131072, 512]
    nonzero = 0
    for o, n in zip(offs, sizes):
        for k in rng.integers(0, n, 6):
            i = o + int(k)
            h = 1e-6
            pp, pm = p64.copy(), p64.copy()
            pp[i] += h
            pm[i] -= h
            fd = (_np_loss(pp, x64, labels, mask) - _np_loss(pm, x64, labels, mask)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-3 * abs(fd) + 1e-6, (i, fd, g[i])
            nonzero += abs(fd) > 1e-4
    assert nonzero >= 12


def test_bf16_rounding(oracle):
    L = oracle._mlp_lib()
    assert L.ora_bf16_round(1.0) == 1.0
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even
    assert L.ora_bf16_round(1.0 + 2 ** -8) == 1.0
    assert L.ora_bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert L.ora_bf16_round(-2.5) == -2.5


def test_adam_first_step_and_convergence(oracle):
    w = np.array([1.0, -2.0, 3.0], np.float32)
    m = np.zeros(3, np.float32)
    v = np.zeros(3, np.float32)
    g = np.array([0.5, -4.0, 1e-3], np.float32)
    oracle.adam_step(w, m, v, g, 1, lr=0.01)
    # bias-corrected first step moves every weight by lr against the
    # gradient's sign (|g| >> eps)
    assert np.allclose(w, [0.99, -1.99, 2.99], atol=1e-5)
    # minimises a quadratic
    w = np.array([5.0, -3.0], np.float32)
    m = np.zeros(2, np.float32)
    v = np.zeros(2, np.float32)
    for t in range(1, 3001):
        oracle.adam_step(w, m, v, 2 * (w - np.array([1.0, 2.0], np.float32)), t, lr=0.05)
    assert np.allclose(w, [1.0, 2.0], atol=1e-2)
