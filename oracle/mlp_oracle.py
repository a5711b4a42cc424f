"""FP64 restatement of the control-variate MLP (forward, backward, masked MSE).

TEST INFRASTRUCTURE ONLY (the checker of the BF16 MFMA kernels in
wobtoolbox_amd/csrc/mlp_kernels.hip).  The network has no counterpart in the reference.

Parameters are one flat FP32 vector: for layer l = 0..4, W_l (out x in, row-major) then
b_l (out); widths 128 -> 128 -> 128 -> 128 -> 128 -> 16.  Features are
[sin(2 pi F y), cos(2 pi F y)] for 64 frequencies F (64 x 3).  With bf16=True the rounding
points of the kernels are emulated (features, weights, hidden activations and the dZ
operands of the gradient products rounded to BF16; sums in FP64).
"""
import numpy as np

WIDTH, NO, LAYERS = 128, 16, 5


def layer_shapes():
    return [(WIDTH, WIDTH)] * 4 + [(NO, WIDTH)]


def unpack(params):
    out, off = [], 0
    for (o, i) in layer_shapes():
        W = params[off:off + o * i].reshape(o, i); off += o * i
        b = params[off:off + o]; off += o
        out.append((W, b))
    return out


def bf16(x):
    x = np.ascontiguousarray(np.asarray(x, np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(freq, pos, bf=True):
    pos = np.asarray(pos, np.float32)
    d = pos.shape[1]
    f = np.asarray(freq, np.float32).reshape(64, 3)[:, :d]
    a = (2.0 * (pos.astype(np.float64) @ f.astype(np.float64).T)).astype(np.float32).astype(np.float64)
    s, c = np.sin(np.pi * a), np.cos(np.pi * a)
    h = np.concatenate([s, c], 1)
    return bf16(h) if bf else h


def forward(params, freq, pos, bf=True):
    layers = unpack(np.asarray(params, np.float64))
    h = features(freq, pos, bf)
    hs = [h]
    for l, (W, b) in enumerate(layers):
        Wq = bf16(W) if bf else W
        z = h @ Wq.T + b
        if l < LAYERS - 1:
            h = np.maximum(z, 0.0)
            if bf:
                h = bf16(h)
            hs.append(h)
        else:
            out = z
    return out, hs


def loss_and_grad(params, freq, pos, idx, target, bf=True):
    """Masked MSE averaged over the batch and its gradient w.r.t. the flat parameters."""
    params = np.asarray(params, np.float64)
    layers = unpack(params)
    out, hs = forward(params, freq, pos, bf)
    n = out.shape[0]
    e = out[np.arange(n), idx] - np.asarray(target, np.float64)
    loss = float((e * e).mean())
    dz = np.zeros_like(out)
    dz[np.arange(n), idx] = 2.0 * e / n
    grads = [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        W, b = layers[l]
        dzq = bf16(dz) if bf else dz
        gW = dzq.T @ hs[l]
        gb = dz.sum(0)
        grads[l] = (gW, gb)
        if l == 0:
            break
        dh = dzq @ (bf16(W) if bf else W)
        dz = dh * (hs[l] > 0)
    flat = np.concatenate([np.concatenate([g.ravel(), gb]) for g, gb in grads])
    return loss, flat
