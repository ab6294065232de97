"""Torch-CPU float32 reference of the flux-correction network (test
infrastructure): the same layers, with the inputs, the weights and the
hidden activations rounded to bfloat16 where the HIP kernels round them, so
that forward values and autograd gradients can be compared with the
library's MFMA kernels."""
import numpy as np
import torch
import torch.nn.functional as F

LAYOUT = (("W1", (27, 5, 32)), ("b1", (32,)), ("W2", (27, 32, 32)), ("b2", (32,)),
          ("W3", (27, 32, 15)), ("b3", (15,)))


def split(flat):
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], np.float32).reshape(shape)
        o += n
    return out


def bf(x):
    return x.to(torch.bfloat16).to(torch.float32)


def conv_weight(W):
    """[tap][cin][cout] (tap = (dx*3 + dy)*3 + dz) -> torch [cout][cin][3][3][3]."""
    t = torch.as_tensor(W).reshape(3, 3, 3, W.shape[1], W.shape[2])
    return t.permute(4, 3, 0, 1, 2).contiguous()


class Net:
    def __init__(self, flat):
        p = split(flat)
        self.w = [conv_weight(p[k]).requires_grad_(True) for k in ("W1", "W2", "W3")]
        self.b = [torch.as_tensor(p[k]).clone().requires_grad_(True) for k in ("b1", "b2", "b3")]

    def forward(self, x):
        """x [B, D0, D1, D2, 5] (numpy or torch) -> y [B, D0-6, D1-6, D2-6, 15]."""
        x = torch.as_tensor(np.asarray(x, np.float32)).permute(0, 4, 1, 2, 3)
        h = bf(x)
        for l in range(3):
            z = F.conv3d(h, bf(self.w[l])) + self.b[l].view(1, -1, 1, 1, 1)
            h = bf(F.gelu(z)) if l < 2 else z
        return h.permute(0, 2, 3, 4, 1)

    def grads_flat(self):
        out = []
        for l, (wk, bk) in enumerate(zip(self.w, self.b)):
            g = wk.grad.permute(2, 3, 4, 1, 0).reshape(27, wk.shape[1], wk.shape[0])
            out += [g.reshape(-1), bk.grad.reshape(-1)]
        return torch.cat(out).detach().numpy()


def divergence_loss(G, target):
    """loss = mean over cells and components of (-div G - target)^2; G on the
    (L+1)^3 faces (index c+1 = face c + e_d/2), channel 5d + q."""
    P = -((G[:, 1:, 1:, 1:, 0:5] - G[:, :-1, 1:, 1:, 0:5]) + (G[:, 1:, 1:, 1:, 5:10] - G[:, 1:, :-1, 1:, 5:10])
          + (G[:, 1:, 1:, 1:, 10:15] - G[:, 1:, 1:, :-1, 10:15]))
    return ((P - torch.as_tensor(target)) ** 2).mean()
