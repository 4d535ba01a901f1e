"""FP64 numpy restatement of the MLP of include/pwa_mlp.h (test checker):
forward and backward with the GEMM operands rounded to bf16 exactly as the
GPU kernels round them (round to nearest even), so the GPU gradient can be
compared at accumulation precision; with bf16=False the exact FP64 network."""
import numpy as np
from scipy.special import erf


def bf16(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(w, fan):
    out, o = [], 0
    for l in range(len(fan) - 1):
        W = w[o:o + fan[l] * fan[l + 1]].reshape(fan[l], fan[l + 1])
        o += fan[l] * fan[l + 1]
        b = w[o:o + fan[l + 1]]
        o += fan[l + 1]
        out.append((np.asarray(W, np.float64), np.asarray(b, np.float64)))
    return out


def forward(w, fan, x, use_bf16=True):
    r = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(w, fan)
    a = r(x)
    acts, zs = [a], []
    for l, (W, b) in enumerate(layers):
        z = a @ r(W) + b
        zs.append(z)
        if l < len(layers) - 1:
            a = r(gelu(z))
            acts.append(a)
    return zs[-1], zs, acts


def loss_grad(w, fan, x, y, ow, use_bf16=True):
    r = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(w, fan)
    out, zs, acts = forward(w, fan, x, use_bf16)
    B = x.shape[0]
    d = out - y
    loss = float(np.sum(ow * d * d) / B)
    dz = 2.0 * ow * d / B
    grads = [None] * len(layers)
    for l in reversed(range(len(layers))):
        W, _ = layers[l]
        dzr = r(dz)
        grads[l] = (acts[l].T @ dzr, dz.sum(axis=0))
        if l == 0:
            break
        dz = (dzr @ r(W).T) * gelu_d(zs[l - 1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
