"""The warm-start U-Net's BF16 MFMA kernels against the CPU oracle
(oracle/unet_torch.py): the MFMA operand layout, the forward pass, the loss
and the weight gradients of first, middle and last layers, and a short
training run that must reduce the loss."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def unet():
    from cffc_amd import unet as U
    return U


def bf16_round(x):
    import torch
    return torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float32).numpy()


def test_mfma_operand_layout(unet):
    """lane l holds A[l%16][8*(l/16)+i], B[8*(l/16)+i][l%16], D[4*(l/16)+i][l%16]."""
    D = unet.mfma_layout_probe()
    A = bf16_round(np.array([[m + k / 64.0 for k in range(32)] for m in range(16)], dtype=np.float32))
    Bm = np.array([[1.0 if (k == n or k == n + 16) else 0.0 for n in range(16)] for k in range(32)], dtype=np.float32)
    ref = A @ Bm
    got = np.zeros((16, 16), dtype=np.float32)
    for l in range(64):
        for i in range(4):
            got[4 * (l // 16) + i, l % 16] = D[l, i]
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


def _setup(unet, H=16, W=48, B=3, seed=7):
    rng = np.random.default_rng(seed)
    net = unet.UNet(H=H, W=W, max_batch=B, seed=seed)
    X = rng.normal(0, 1, (B, H, W, 8)).astype(np.float32)
    T = rng.normal(0, 0.5, (B, H, W, 4)).astype(np.float32)
    net.set_dataset(X, T)
    # a non-trivial output layer so every layer's gradient is exercised
    l = net.nlayers - 1
    w, b = net.get_weights(l)
    net.set_weights(l, w * 20.0, b + 0.1)
    return net, X, T


def test_forward_and_loss_match_oracle(unet):
    import torch
    from oracle.unet_torch import RefUNet
    net, X, T = _setup(unet)
    shapes = [net.layer_shape(l) for l in range(net.nlayers)]
    ref = RefUNet([net.get_weights(l) for l in range(net.nlayers)], shapes)
    with torch.no_grad():
        loss_ref, y_ref = ref.loss(X, T)
    y = net.predict(X)
    err = np.abs(y - y_ref[..., :4].numpy()).max() / np.abs(y_ref[..., :4].numpy()).max()
    print("forward max rel err %.3e" % err)
    assert err < 2e-2
    loss, _, _ = net.loss_grad(np.arange(X.shape[0]), 0)
    assert abs(loss - float(loss_ref)) / float(loss_ref) < 2e-2


@pytest.mark.parametrize("layer", [0, 5, 8, 12, 14])
def test_weight_gradients_match_oracle(unet, layer):
    from oracle.unet_torch import RefUNet
    net, X, T = _setup(unet)
    shapes = [net.layer_shape(l) for l in range(net.nlayers)]
    ref = RefUNet([net.get_weights(l) for l in range(net.nlayers)], shapes)
    loss_ref, _ = ref.loss(X, T)
    loss_ref.backward()
    gw_ref = ref.w[layer].grad.numpy()
    gb_ref = ref.b[layer].grad.numpy()
    _, gw, gb = net.loss_grad(np.arange(X.shape[0]), layer)
    rel = np.linalg.norm(gw - gw_ref) / np.linalg.norm(gw_ref)
    cos = float(np.sum(gw * gw_ref) / (np.linalg.norm(gw) * np.linalg.norm(gw_ref)))
    relb = np.linalg.norm(gb - gb_ref) / max(np.linalg.norm(gb_ref), 1e-30)
    print("layer %d: |dW| rel err %.3e cos %.6f, |db| rel err %.3e" % (layer, rel, cos, relb))
    assert rel < 5e-2 and cos > 0.998 and relb < 5e-2


def test_training_reduces_loss(unet):
    net, X, T = _setup(unet, B=4)
    losses = net.train(60, batch=4, lr=3e-3)
    print("loss %.4e -> %.4e" % (losses[0], losses[-1]))
    assert losses[-1] < 0.5 * losses[0]
