"""The BF16-MFMA MLP (include/pwa_mlp.h) against an FP64 numpy restatement:
forward, loss and every parameter gradient, and training on a known function."""
import numpy as np
import pytest

import mlp_ref
from pwa_amd.mlp import MLP

pytestmark = pytest.mark.gpu


def _data(n, n_in, n_out, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, n_in)).astype(np.float32)
    y = rng.normal(size=(n, n_out)).astype(np.float32)
    ow = rng.uniform(0.5, 2.0, size=n_out).astype(np.float32)
    return x, y, ow


@pytest.mark.parametrize("dims", [(20, 64, 3, 9), (70, 256, 3, 71)])
def test_loss_and_gradient_match_cpu_restatement(dims):
    n_in, H, L, n_out = dims
    fan = [n_in] + [H] * L + [n_out]
    x, y, ow = _data(200, n_in, n_out, 3)
    net = MLP(n_in, n_out, H, L, seed=5)
    net.set_data(x, y, ow)
    w = net.get_params().astype(np.float64)
    loss, g = net.loss_grad(0, 200)
    ref_loss, ref_g = mlp_ref.loss_grad(w, fan, x.astype(np.float64), y.astype(np.float64), ow.astype(np.float64))
    ex_loss, ex_g = mlp_ref.loss_grad(w, fan, x.astype(np.float64), y.astype(np.float64), ow.astype(np.float64),
                                      use_bf16=False)
    rel_l = abs(loss - ref_loss) / ref_loss
    rel_g = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    rel_gx = np.linalg.norm(g - ex_g) / np.linalg.norm(ex_g)
    print(f"[mlp] {dims}: loss rel {rel_l:.2e}; grad vs bf16 restatement {rel_g:.2e}, vs exact FP64 {rel_gx:.2e}",
          flush=True)
    assert rel_l < 1e-4
    assert rel_g < 1e-3
    assert rel_gx < 5e-2
    # forward / predict agree with the training-path forward
    pred = net.predict(x[:50])
    out_ref, _, _ = mlp_ref.forward(w, fan, x[:50].astype(np.float64))
    assert np.max(np.abs(pred - out_ref)) <= 1e-3 * np.max(np.abs(out_ref)) + 1e-5
    net.close()


def test_training_fits_a_smooth_function():
    rng = np.random.default_rng(7)
    n, n_in = 4096, 8
    x = rng.uniform(-1, 1, size=(n, n_in)).astype(np.float32)
    A = rng.normal(size=(n_in, n_in))
    A = A @ A.T / n_in
    f = 0.5 * np.einsum("ni,ij,nj->n", x, A, x)
    y = np.concatenate([f[:, None], x @ A], axis=1).astype(np.float32)  # value and gradient
    net = MLP(n_in, n_in + 1, 128, 3, seed=1)
    net.set_data(x, y, np.ones(n_in + 1, np.float32))
    l0, _ = net.loss_grad(0, 1024, want_grad=False)
    l1 = net.train(60, batch=256, lr=2e-3, seed=3)
    print(f"[mlp] training loss {l0:.3e} -> {l1:.3e}", flush=True)
    assert l1 < 0.02 * l0
    net.close()
