"""Pins of the learned-step oracle (oracle/net_oracle.py) on the CPU.

The network oracle is checked against torch's own Conv1d and autograd in
float64 (forward, loss and every parameter gradient), and its bfloat16
rounding against torch's float32 -> bfloat16 conversion.
"""
import numpy as np
import torch

from oracle import net_oracle as NO


def _torch_model(params, X0):
    pt = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    X = torch.tensor(X0, dtype=torch.float64).permute(0, 2, 1)
    for j in range(NO.NLAYER):
        o = j * NO.LAYER_SIZE
        W = pt[o:o + NO.W_SIZE].reshape(NO.C, NO.C, NO.KW)
        b = pt[o + NO.W_SIZE:o + NO.LAYER_SIZE]
        X = torch.nn.functional.conv1d(X, W, b, padding=NO.HALF)
        if j < NO.NLAYER - 1:
            X = torch.relu(X)
    return pt, X[:, 0, :]


def test_forward_and_gradients_match_torch_autograd():
    rng = np.random.default_rng(1)
    p = NO.init_params(3).astype(np.float64) + rng.normal(0, 0.05, NO.PARAM_COUNT) * NO.trainable_mask()
    u = rng.normal(size=(2, 40)).astype(np.float32)
    X0, _ = NO.features(u, quant=False)
    X0 = X0.astype(np.float64)
    acts, y = NO.forward(p, X0, quant=False)
    ys = rng.normal(size=y.shape)
    mask = np.ones_like(y)
    mask[:, :3] = 0.0
    loss, g = NO.loss_and_grads(p, acts, y, ys, mask, quant=False)
    pt, yt = _torch_model(p, X0)
    mt = torch.tensor(mask)
    lt = (((yt - torch.tensor(ys)) * mt) ** 2).sum() / mt.sum()
    lt.backward()
    gt = pt.grad.numpy() * NO.trainable_mask()
    assert np.max(np.abs(yt.detach().numpy() - y)) < 1e-12
    assert abs(float(lt.detach()) - loss) < 1e-12 * loss
    assert np.linalg.norm(g - gt) <= 1e-12 * np.linalg.norm(gt)


def test_conv_adjoint_identity():
    # <conv(X, W), D> == <X, conv_t(D, W)> and == <W, conv_w(D, X)>
    rng = np.random.default_rng(2)
    X = rng.normal(size=(2, 30, NO.C))
    W = rng.normal(size=(NO.C, NO.C, NO.KW))
    D = rng.normal(size=(2, 30, NO.C))
    lhs = np.sum(NO.conv(X, W) * D)
    assert abs(lhs - np.sum(X * NO.conv_t(D, W))) < 1e-9 * abs(lhs)
    assert abs(lhs - np.sum(W * NO.conv_w(D, X))) < 1e-9 * abs(lhs)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(NO.bf16(x), want)


def test_trainable_mask_and_features():
    m = NO.trainable_mask()
    layers = NO.unpack(m)
    assert layers[0][0][:, :2, :].all() and not layers[0][0][:, 2:, :].any()
    assert layers[3][0][0].all() and not layers[3][0][1:].any()
    assert layers[3][1][0] == 1 and not layers[3][1][1:].any()
    u = np.array([[0.0, 2.0, -4.0, 1.0]], np.float32)
    X, A = NO.features(u, quant=False)
    assert A[0] == 4.0
    assert np.allclose(X[0, :, 0], [0.0, 0.5, -1.0, 0.25]) and np.allclose(X[0, :, 1], [0.0, 0.25, 1.0, 0.0625])
    assert not X[..., 2:].any()
    # the fallback indicator of a sampled sinusoid A sin(w t) is 2 A sin(w dT / 2) / dT (~ A w)
    dT, f0, amp = 4e-8, 1e6, 3e5
    t = np.arange(4096) * dT
    s = NO.max_dpdt((amp * np.sin(2 * np.pi * f0 * t))[None, :], dT)
    assert abs(s[0] / (2 * amp * np.sin(np.pi * f0 * dT) / dT) - 1.0) < 1e-5


def test_apply_leaves_quiet_chunks_unchanged():
    rng = np.random.default_rng(4)
    p = NO.init_params(1) + (rng.normal(0, 0.05, NO.PARAM_COUNT) * NO.trainable_mask()).astype(np.float32)
    p[3 * NO.LAYER_SIZE + NO.W_SIZE] = 0.3  # output bias: a nonzero y everywhere
    u = np.zeros((2, 1024), np.float32)
    u[:, 560:600] = rng.normal(size=(2, 40)).astype(np.float32)
    out = NO.apply(p, u, 1e-3)
    # 128-sample segments: only [512, 640) (window [496, 656)) sees the data
    assert np.array_equal(out[:, :512], u[:, :512])
    assert not np.array_equal(out[:, 512:640], u[:, 512:640])
    assert np.array_equal(out[:, 640:], u[:, 640:])


def test_apply_leaves_negligible_columns_unchanged():
    # c max|u[t+1] - u[t]| < 1e-5: the column is left as it is, network or not
    rng = np.random.default_rng(5)
    p = NO.init_params(2) + (rng.normal(0, 0.05, NO.PARAM_COUNT) * NO.trainable_mask()).astype(np.float32)
    p[3 * NO.LAYER_SIZE + NO.W_SIZE] = 0.3
    t = np.arange(256)
    u = np.stack([np.sin(0.2 * t), np.sin(0.2 * t)]).astype(np.float32)   # max step ~0.2
    c = 4e-5                                                              # c D ~ 8e-6 < 1e-5
    out = NO.apply(p, u, c)
    assert np.array_equal(out, u)
    out2 = NO.apply(p, 2 * u, c)                                          # c D ~ 1.6e-5
    assert not np.array_equal(out2, 2 * u)
