"""TEST INFRASTRUCTURE ONLY -- CPU oracle of the warm-start U-Net.

A plain PyTorch (CPU) restatement of the network in cffc_amd/csrc/unet.cpp:
4 levels, channels 16-32-64-128, two 3x3 conv + ReLU per level, 2x2 average
pooling, nearest 2x upsampling, skip connections by concatenation
[upsampled, skip], final 3x3 conv to 16 channels (first 4 used), MSE loss.
With emulate_bf16 the forward rounds inputs, weights and every stored
activation to bfloat16 as the GPU kernels do (straight-through for autograd),
so forward values agree closely; the GPU's backward also rounds the
activation gradient to bfloat16 before each MFMA, which this oracle does not,
so gradients agree to bf16 accuracy.
"""
import numpy as np
import torch
import torch.nn.functional as F


def bf16(t):
    return t + (t.to(torch.bfloat16).to(torch.float32) - t).detach()


def conv_weight(w, cin):
    """[Cout][9*Cin] with k = tap*Cin + ci, tap = 3*ky + kx -> [Cout][Cin][3][3]."""
    co = w.shape[0]
    return w.reshape(co, 9, cin).permute(0, 2, 1).reshape(co, cin, 3, 3)


class RefUNet:
    def __init__(self, weights, shapes):
        # weights: list of (w [Cout][9Cin] float32 tensor, b [Cout]); shapes: (CA, CB, Cout, level)
        self.w = [torch.tensor(w, dtype=torch.float32, requires_grad=True) for w, _ in weights]
        self.b = [torch.tensor(b, dtype=torch.float32, requires_grad=True) for _, b in weights]
        self.shapes = shapes

    def forward(self, X, emulate_bf16=True):
        r = bf16 if emulate_bf16 else (lambda t: t)
        x = r(torch.tensor(X, dtype=torch.float32).permute(0, 3, 1, 2))
        li = [0]

        def conv(inp, relu=True):
            l = li[0]
            li[0] += 1
            CA, CB, Co, _ = self.shapes[l]
            wt = conv_weight(r(self.w[l]), CA + CB)
            y = F.conv2d(inp, wt, self.b[l], padding=1)
            y32 = y
            if relu:
                y = torch.relu(y)
            return r(y), y32

        skips = []
        for lev in range(4):
            x, _ = conv(x)
            x, _ = conv(x)
            if lev < 3:
                skips.append(x)
                x = r(F.avg_pool2d(x, 2))
        for lev in (2, 1, 0):
            x = F.interpolate(x, scale_factor=2, mode="nearest")
            x = torch.cat([x, skips[lev]], dim=1)
            x, _ = conv(x)
            x, _ = conv(x)
        _, y32 = conv(x, relu=False)
        return y32.permute(0, 2, 3, 1)  # (B, H, W, 16)

    def loss(self, X, T, emulate_bf16=True):
        y = self.forward(X, emulate_bf16)
        t = torch.tensor(T, dtype=torch.float32)
        return torch.mean((y[..., :4] - t) ** 2), y
