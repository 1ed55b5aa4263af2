"""PyTorch CPU FP32 restatement of the surrogate network (the checker of the
HIP forward and backward kernels).

It mirrors the library's storage points: weights of the 3^3 and the 1x1
inner layers are rounded to bfloat16, every stored activation is rounded to
bfloat16 (round to nearest even, as torch's cast), biases and the head stay
FP32; convolutions are periodic (circular padding), 2^3 average pooling
down, nearest upsampling added to the skip tensor on the way up.
"""
import numpy as np
import torch
import torch.nn.functional as F

LAYERS = [("e1", 4, 16, 3), ("e1b", 16, 16, 3), ("e2", 16, 32, 3), ("b", 32, 64, 3), ("q2", 64, 32, 1),
          ("d2", 32, 32, 3), ("q1", 32, 16, 1), ("d1", 16, 16, 3), ("head", 16, 1, 1)]


def split_params(flat, requires_grad=False):
    """{name: (weight, bias)} from the flat parameter vector (PyTorch order)."""
    out = {}
    k = 0
    for name, cin, cout, ks in LAYERS:
        nw = cout * cin * ks ** 3
        w = torch.tensor(np.asarray(flat[k:k + nw], dtype=np.float32).reshape(cout, cin, ks, ks, ks))
        k += nw
        b = torch.tensor(np.asarray(flat[k:k + cout], dtype=np.float32))
        k += cout
        if requires_grad:
            w.requires_grad_(True)
            b.requires_grad_(True)
        out[name] = (w, b)
    assert k == len(flat)
    return out


def flat_grads(params):
    parts = []
    for name, *_ in LAYERS:
        w, b = params[name]
        parts += [w.grad.numpy().ravel(), b.grad.numpy().ravel()]
    return np.concatenate(parts)


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def _conv(x, wb, ks, relu, round_w=True):
    w, b = wb
    if ks == 3:
        x = F.pad(x, (1, 1, 1, 1, 1, 1), mode="circular")
    y = F.conv3d(x, bf(w) if round_w else w, b)
    return torch.relu(y) if relu else y


def _up(x):
    return x.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3).repeat_interleave(2, dim=4)


def forward(params, in4):
    """in4: (4, Nz, Ny, Nx) array or tensor -> (Nz, Ny, Nx) tensor."""
    x0 = bf(torch.as_tensor(np.asarray(in4, dtype=np.float32))[None])
    ae1 = bf(_conv(x0, params["e1"], 3, True))
    s1 = bf(_conv(ae1, params["e1b"], 3, True))
    p1 = bf(F.avg_pool3d(s1, 2))
    s2 = bf(_conv(p1, params["e2"], 3, True))
    p2 = bf(F.avg_pool3d(s2, 2))
    a3 = bf(_conv(p2, params["b"], 3, True))
    q2 = bf(_conv(a3, params["q2"], 1, False))
    t2 = bf(s2 + _up(q2))
    a2 = bf(_conv(t2, params["d2"], 3, True))
    q1 = bf(_conv(a2, params["q1"], 1, False))
    t1 = bf(s1 + _up(q1))
    a1 = bf(_conv(t1, params["d1"], 3, True))
    out = _conv(a1, params["head"], 1, False, round_w=False)
    return out[0, 0]
