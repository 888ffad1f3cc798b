"""Warm-start MLP: the numpy restatement (tests/mlp_ref.py) pinned on the CPU, then
the GPU forward / gradients / Adam through the C ABI checked against it."""
import numpy as np
import pytest

import mlp_ref as R


# ----------------------------------------------------------------- CPU pins
def test_bf16_rounding_against_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 5000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0], dtype=np.float32)])
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(R.bf16(x), ref)
    assert R.bf16(np.float32(1.0 + 2 ** -8)) == 1.0                  # tie to even
    assert R.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6      # tie to even (upwards)


def test_gelu_grad_is_derivative():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (R.gelu(z + h) - R.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - R.gelu_grad(z))) < 1e-8
    assert abs(R.gelu(np.array([0.0]))[0]) == 0.0
    assert abs(R.gelu(np.array([10.0]))[0] - 10.0) < 1e-12


def test_restatement_backward_by_finite_differences():
    """Analytic gradients of the (unrounded) restatement vs central differences."""
    rng = np.random.default_rng(7)
    shapes = [(3, 5), (5, 5), (5, 4)]
    layers = [(rng.standard_normal(s) * 0.7, rng.standard_normal(s[1]) * 0.1) for s in shapes]
    x = rng.standard_normal((6, 3))
    t = rng.standard_normal((6, 4))
    loss, grads = R.loss_and_grads(layers, x, t, round_bf16=False)
    h = 1e-6
    for li, (W, b) in enumerate(layers):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            it = np.nditer(arr, flags=["multi_index"])
            for _ in it:
                idx = it.multi_index
                old = arr[idx]
                arr[idx] = old + h
                lp, _ = R.loss_and_grads(layers, x, t, round_bf16=False)
                arr[idx] = old - h
                lm, _ = R.loss_and_grads(layers, x, t, round_bf16=False)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g[idx]) < 1e-7 * max(1.0, abs(fd)), (li, idx, fd, g[idx])


# ----------------------------------------------------------------- GPU
def _mlp(n_in, width, hidden, n_out, seed):
    import loadsg_amd as L
    return L.MLP(n_in, width, hidden, n_out, seed=seed)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("n_in,width,hidden,n_out,batch", [(24, 64, 2, 1000, 16), (24, 128, 3, 20000, 37),
                                                           (7, 96, 1, 130, 70), (24, 512, 3, 61183, 256),
                                                           (24, 192, 2, 333, 130)])
def test_forward_and_gradients_vs_restatement(n_in, width, hidden, n_out, batch):
    import torch
    print("mlp", n_in, width, hidden, n_out, batch, flush=True)
    m = _mlp(n_in, width, hidden, n_out, seed=11)
    p = m.params()
    layers = R.unflatten(p, m.layer_shapes())
    rng = np.random.default_rng(5)
    x = rng.standard_normal((batch, n_in)).astype(np.float32)
    t = rng.standard_normal((batch, n_out)).astype(np.float32)
    xd = torch.tensor(x, device="cuda")
    td = torch.tensor(t, device="cuda")
    y = m.forward(xd).cpu().numpy()
    y_ref, _, _ = R.forward(layers, x)
    assert _rel(y, y_ref) < 2e-3
    loss = m.gradients(xd, td)
    loss_ref, g_ref = R.loss_and_grads(layers, x, t)
    assert abs(loss - loss_ref) < 1e-5 * loss_ref
    g = m.grads()
    o = 0
    # Both sides round to BF16 at the same points; they differ by FP32 summation order,
    # which now and then moves a value across a BF16 rounding boundary.  Measured: weight
    # gradients within 2.8e-4 (relative 2-norm per layer), bias gradients within 4e-5, the
    # loss within 6e-7; the bounds below leave a factor of about 3.
    for (i, n), (gW, gb) in zip(m.layer_shapes(), g_ref):
        GW = g[o:o + i * n].reshape(i, n)
        o += i * n
        Gb = g[o:o + n]
        o += n
        assert np.linalg.norm(GW - gW) < 1e-3 * np.linalg.norm(gW)
        assert np.linalg.norm(Gb - gb) < 2e-4 * np.linalg.norm(gb)
    # deterministic: the same call gives the same bits
    m.gradients(xd, td)
    assert np.array_equal(m.grads(), g)


@pytest.mark.gpu
def test_full_width_forward_few_samples():
    """24 -> 512 x 3 -> 61183 (the d=5 level-7 grid size) on 4 samples."""
    import torch
    m = _mlp(24, 512, 3, 61183, seed=2)
    layers = R.unflatten(m.params(), m.layer_shapes())
    x = np.random.default_rng(1).standard_normal((4, 24)).astype(np.float32)
    y = m.forward(torch.tensor(x, device="cuda")).cpu().numpy()
    y_ref, _, _ = R.forward(layers, x)
    assert _rel(y, y_ref) < 2e-3


@pytest.mark.gpu
def test_adam_training_reduces_loss_and_params_roundtrip():
    import torch
    rng = np.random.default_rng(9)
    n_in, n_out = 8, 300
    A = rng.standard_normal((n_in, n_out)).astype(np.float32) / np.sqrt(n_in)
    x = rng.standard_normal((256, n_in)).astype(np.float32)
    t = np.tanh(x @ A).astype(np.float32)
    xd, td = torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda")
    m = _mlp(n_in, 128, 2, n_out, seed=4)
    first = m.train_step(xd, td, 1e-3)
    last = first
    for _ in range(300):
        last = m.train_step(xd, td, 1e-3)
    assert last < 0.05 * first
    p = m.params()
    m2 = _mlp(n_in, 128, 2, n_out, seed=99)
    m2.set_params(p)
    assert np.array_equal(m2.params(), p)
    assert np.array_equal(m2.forward(xd).cpu().numpy(), m.forward(xd).cpu().numpy())


@pytest.mark.gpu
def test_mix_rows_matches_batched_product():
    """Training targets sum_j w[b, j] U[rows[b], j] against a float64 product."""
    import torch
    rng = np.random.default_rng(21)
    S, J, n, B = 9, 8, 1237, 33
    U = torch.tensor(rng.standard_normal((S, J, n)), dtype=torch.float32, device="cuda")
    rows = torch.tensor(rng.integers(0, S, B), dtype=torch.int64, device="cuda")
    w = torch.tensor(rng.standard_normal((B, J)), dtype=torch.float32, device="cuda")
    out = _mlp(4, 64, 1, 10, seed=1).mix_rows(U, rows, w)
    ref = (w.double().unsqueeze(1) @ U.double()[rows]).squeeze(1)
    assert float((out.double() - ref).abs().max() / ref.abs().max()) < 1e-6
