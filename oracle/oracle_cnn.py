"""FP64 numpy oracle of the SFS CNN -- TEST INFRASTRUCTURE ONLY.

Network (new, no counterpart in the reference): four 3×3×3 periodic
convolutions 3 → 64 → 64 → 64 → 6 with erf-GELU between them, MSE loss.
Parameter vector layout (shared with the HIP library): per layer the weights
W[o][n][k] (o = 0..26 ↔ offset (o%3−1, (o/3)%3−1, o/9−1) in (x, y, z), n output
and k input channel) followed by the bias[n].  Activations are arrays indexed
[b, z, y, x, c].  Pinned by central finite differences in tests/test_oracle_cnn.py.
"""
import numpy as np
from scipy.special import erf

CH = ((3, 64), (64, 64), (64, 64), (64, 6))


def nparams():
    return sum(27 * co * ci + co for ci, co in CH)


def unpack(p):
    out, off = [], 0
    for ci, co in CH:
        W = p[off:off + 27 * co * ci].reshape(27, co, ci)
        off += 27 * co * ci
        b = p[off:off + co]
        off += co
        out.append((W, b))
    return out


def offsets():
    return [(o % 3 - 1, (o // 3) % 3 - 1, o // 9 - 1) for o in range(27)]


def conv(x, W, b):
    """out[v, n] = b[n] + Σ_o Σ_k W[o, n, k] x[v + d_o, k] (periodic)."""
    out = np.zeros(x.shape[:4] + (W.shape[1],)) + b
    for o, (dx, dy, dz) in enumerate(offsets()):
        xs = np.roll(x, shift=(-dz, -dy, -dx), axis=(1, 2, 3))
        out += xs @ W[o].T
    return out


def gelu(z):
    return 0.5 * z * (1 + erf(z / np.sqrt(2)))


def gelu_d(z):
    return 0.5 * (1 + erf(z / np.sqrt(2))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(p, x):
    layers = unpack(p)
    acts, zs = [x], []
    h = x
    for l, (W, b) in enumerate(layers):
        z = conv(h, W, b)
        if l < 3:
            zs.append(z)
            h = gelu(z)
            acts.append(h)
        else:
            h = z
    return h, acts, zs


def loss_grad(p, x, t):
    y, acts, zs = forward(p, x)
    r = y - t
    loss = float((r * r).mean())
    dy = 2 * r / r.size
    layers = unpack(p)
    grads = []
    d = dy
    for l in range(3, -1, -1):
        W, b = layers[l]
        inp = acts[l]
        gW = np.zeros_like(W)
        for o, (dx, dy_, dz) in enumerate(offsets()):
            xs = np.roll(inp, shift=(-dz, -dy_, -dx), axis=(1, 2, 3))
            gW[o] = d.reshape(-1, d.shape[-1]).T @ xs.reshape(-1, xs.shape[-1])
        gb = d.sum(axis=(0, 1, 2, 3))
        grads.append((gW, gb))
        if l > 0:
            din = np.zeros(inp.shape)
            for o, (dx, dy_, dz) in enumerate(offsets()):
                din += np.roll(d @ W[o], shift=(dz, dy_, dx), axis=(1, 2, 3))
            d = din * gelu_d(zs[l - 1])
    grads = grads[::-1]
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def bf16_round(a):
    """Round float array to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
