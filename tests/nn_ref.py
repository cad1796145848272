"""FP32 torch reference of the corrector network (CPU), with the BF16
roundings of the device kernels emulated: inputs, weights and the stored
activations of layers 1 and 2 are rounded to bfloat16; arithmetic is FP32."""
import numpy as np
import torch

LAYER = 27 * 16 * 16 + 16
LEAK = 0.125


def bf16(x):
    return x.to(torch.bfloat16).to(torch.float32)


class _RoundBf16(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return bf16(x)

    @staticmethod
    def backward(ctx, g):
        return g


def split(params):
    """FP32 parameter vector -> [(W [16,16,3,3,3] torch layout, b [16])] x 3."""
    out = []
    for l in range(3):
        blk = np.asarray(params[l * LAYER:(l + 1) * LAYER], np.float32)
        w = blk[:27 * 256].reshape(3, 3, 3, 16, 16)  # [dz][dy][dx][m][k]
        w = np.transpose(w, (3, 4, 0, 1, 2))  # [m][k][dz][dy][dx]
        out.append((torch.tensor(np.ascontiguousarray(w)), torch.tensor(blk[27 * 256:])))
    return out


def forward(params, x, cin=(12, 16, 16), cout=(16, 16, 12), grad=False):
    """x: (nz, ny, nx, 16) -> y (nz, ny, nx, 16) fp32 (and the torch graph)."""
    layers = split(params)
    ws = []
    for l, (w, b) in enumerate(layers):
        w = w.clone()
        w[cout[l]:] = 0
        w[:, cin[l]:] = 0
        b = b.clone()
        b[cout[l]:] = 0
        w = bf16(w).requires_grad_(grad)
        b = b.requires_grad_(grad)
        ws.append((w, b))
    h = bf16(torch.tensor(np.asarray(x, np.float32)).permute(3, 0, 1, 2)[None])
    acts = []
    for l, (w, b) in enumerate(ws):
        h = torch.nn.functional.conv3d(h, w, b, padding=1)
        if l < 2:
            h = torch.where(h > 0, h, h * LEAK)
            h = _RoundBf16.apply(h)
            acts.append(h)
    y = h[0].permute(1, 2, 3, 0)
    return y, ws


def loss_and_grads(params, x, target, mask):
    y, ws = forward(params, x, grad=True)
    m = torch.tensor(np.asarray(mask, np.float32))[..., None]
    t = torch.tensor(np.asarray(target, np.float32))
    d = (y[..., :12] - t[..., :12]) * m
    loss = (d * d).sum() / (m.sum() * 12)
    loss.backward()
    g = []
    for w, b in ws:
        gw = w.grad.permute(2, 3, 4, 0, 1).reshape(-1).numpy()  # [dz][dy][dx][m][k]
        g.append(np.concatenate([gw, b.grad.numpy()]))
    return float(loss.detach()), np.concatenate(g).astype(np.float32)
