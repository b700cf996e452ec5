"""FP64 numpy restatement of the predictor network's forward pass, loss and
gradients (the network of grtm_cuda_amd/csrc/mlp.hip), emulating its BF16
rounding of inputs, weights and hidden activations.  TEST INFRASTRUCTURE
ONLY: the checker of the GPU training kernels."""
import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, shapes):
    out, o = [], 0
    for fi, fo in shapes:
        W = params[o:o + fi * fo].reshape(fi, fo)
        o += fi * fo
        b = params[o:o + fo]
        o += fo
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def loss_and_grads(params, shapes, X, T, M, nq=0, q=0.5):
    layers = unpack(params, shapes)
    h = bf16(X)
    hs, zs = [h], []
    for li, (W, b) in enumerate(layers):
        z = h @ bf16(W) + b
        zs.append(z)
        if li < len(layers) - 1:
            h = bf16(np.maximum(z, 0.0))
            hs.append(h)
    y = zs[-1]
    cnt = y.size
    d = (y - T) * M
    sq = d * d / cnt
    dz = 2.0 * d / cnt
    if nq:
        e = T[:, :nq] - y[:, :nq]
        sq[:, :nq] = M[:, :nq] * np.maximum(q * e, (q - 1.0) * e) / cnt
        dz[:, :nq] = M[:, :nq] * np.where(e > 0, -q, 1.0 - q) / cnt
    loss = float(sq.sum())
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        gW = hs[li].T @ dz
        gb = dz.sum(axis=0)
        grads[li] = (gW, gb)
        if li > 0:
            dh = dz @ bf16(W).T
            dz = dh * (zs[li - 1] > 0)
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat, y
