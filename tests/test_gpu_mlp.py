"""The BF16 matrix-core MLP kernels against the CPU reference (oracle/mlp_ref.py,
BF16 rounding emulated at the same points): forward outputs, gradients of
the training loss, and a training run that lowers the loss."""
import os
import sys

import numpy as np
import pytest

from gpu_util import say

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu


def _setup(B, seed):
    import torch
    from mideq.learned import MLP
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((B, 16)) * 2 + 1).astype(np.float32)
    mean = rng.uniform(0.5, 1.5, 16).astype(np.float32)
    std = rng.uniform(1.0, 3.0, 16).astype(np.float32)
    theta = mlp_ref.init_params(seed) + (0.05 * rng.standard_normal(mlp_ref.NPARAM)).astype(np.float32)
    net = MLP(params=theta)
    net.set_norm(mean, std)
    Xd = torch.from_numpy(X).cuda()
    return net, theta, X, Xd, mean, std, rng


@pytest.mark.parametrize("B", [16, 1000, 65536 + 5])
def test_mlp_forward_matches_reference(B):
    say(f"mlp forward B={B}")
    net, theta, X, Xd, mean, std, _ = _setup(B, 3)
    Y = net.infer(Xd).cpu().numpy()
    ref, _, _ = mlp_ref.forward(theta, X, mean, std)
    err = np.abs(Y - ref)
    scale = np.abs(ref).max()
    # FP32 accumulation order and tanh ulps differ from the reference; a BF16
    # activation can round the other way, so allow a few BF16 ulps at most
    assert np.median(err) < 1e-5 * scale
    assert err.max() < 2e-2 * scale
    assert np.mean(err < 1e-4 * scale) > 0.99


def test_mlp_gradients_match_reference():
    import torch
    B = 777
    say(f"mlp gradient B={B}")
    net, theta, X, Xd, mean, std, rng = _setup(B, 4)
    target = rng.standard_normal(B).astype(np.float32)
    label = (rng.uniform(size=B) > 0.4).astype(np.float32)
    cost = (8.0 + 2.0 * rng.standard_normal(B)).astype(np.float32)
    never = (rng.uniform(size=B) > 0.7).astype(np.float32)
    for with_cost, with_never in ((True, True), (False, True), (True, False)):
        cd = torch.from_numpy(cost).cuda() if with_cost else None
        nd = torch.from_numpy(never).cuda() if with_never else None
        loss = net.grad(Xd, torch.from_numpy(target).cuda(), torch.from_numpy(label).cuda(), cd,
                        1.0, 0.5, 0.3, never=nd, w_never=0.8)
        g = net.get_grads()
        Lr, gr, _ = mlp_ref.loss_and_grad(theta, X, target.astype(np.float64),
                                          label.astype(np.float64), mean, std, 1.0, 0.5,
                                          cost=cost.astype(np.float64) if with_cost else None,
                                          w_cost=0.3,
                                          never=never.astype(np.float64) if with_never else None,
                                          w_never=0.8)
        case = (with_cost, with_never)
        assert abs(loss - Lr) < 1e-4 * abs(Lr) + 1e-6, (case, loss, Lr)
        k = 0
        for name, shp in mlp_ref.SHAPES:
            n = int(np.prod(shp))
            a, b = g[k:k + n], gr[k:k + n]
            rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
            assert rel < 1e-2, (case, name, rel)
            k += n
        # a missing target: that head's column of W4 and its bias get no gradient
        W4 = g[-4 - 4 * 64:-4].reshape(64, 4)
        if not with_cost:
            assert g[-2] == 0.0 and np.all(W4[:, 2] == 0.0)
        if not with_never:
            assert g[-1] == 0.0 and np.all(W4[:, 3] == 0.0)


def test_mlp_training_lowers_loss():
    import torch
    say("mlp training synthetic")
    from mideq.learned import MLP
    rng = np.random.default_rng(9)
    N = 65536
    X = rng.standard_normal((N, 16)).astype(np.float32)
    target = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    label = (X[:, 3] + X[:, 4] > 0).astype(np.float32)
    cost = (6.0 + X[:, 5] - 0.5 * X[:, 6] ** 2).astype(np.float32)
    never = (X[:, 7] - X[:, 8] > 0).astype(np.float32)
    net = MLP(seed=1)
    net.set_norm(np.zeros(16), np.ones(16))
    Xd = torch.from_numpy(X).cuda()
    td = torch.from_numpy(target).cuda()
    ld = torch.from_numpy(label).cuda()
    cd = torch.from_numpy(cost).cuda()
    nd = torch.from_numpy(never).cuda()
    l0 = net.grad(Xd, td, ld, cd, never=nd)
    losses = net.fit(Xd, td, ld, cd, epochs=6, batch=4096, lr=3e-3, never=nd)
    l1 = net.grad(Xd, td, ld, cd, never=nd)
    assert l1 < 0.2 * l0, (l0, l1, losses[-5:])
    # classification accuracy of both logit heads on the training data, and
    # the cost head's rank correlation with its target: the same training
    # emulated in float64 with oracle/mlp_ref.py's gradients (same data,
    # permutation, initial weights and Adam steps) reaches 0.962, 0.949 and
    # 0.888
    Y = net.infer(Xd).cpu().numpy()
    acc = np.mean((Y[:, 1] > 0) == (label > 0.5))
    assert acc > 0.95
    assert np.mean((Y[:, 3] > 0) == (never > 0.5)) > 0.93
    rk = lambda v: np.argsort(np.argsort(v))
    assert np.corrcoef(rk(Y[:, 2]), rk(cost))[0, 1] > 0.85


@pytest.mark.parametrize("B", [1, 4095, 300001])
def test_launch_order_kernel_is_sorted_permutation(B):
    """mideq_mlp_launch_order: a permutation of 0..B-1 whose keys (class,
    then decreasing log2-cost bucket, as learned.order_keys states them) are
    non-decreasing; NaN, negative and huge costs land in the end buckets."""
    import torch
    from mideq.learned import MLP, order_keys
    say(f"launch order B={B}")
    rng = np.random.default_rng(B)
    Y = np.zeros((B, 4), np.float32)
    Y[:, 1] = rng.standard_normal(B)
    Y[:, 2] = rng.uniform(-2.0, 40.0, B)
    Y[::97, 2] = np.nan
    net = MLP(seed=0)
    order = net.launch_order(torch.from_numpy(Y).cuda()).cpu().numpy()
    assert order.dtype == np.int64
    np.testing.assert_array_equal(np.sort(order), np.arange(B))
    keys = order_keys(Y[:, 1], Y[:, 2])
    assert np.all(np.diff(keys[order]) >= 0)

