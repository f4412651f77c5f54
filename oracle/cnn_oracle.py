"""cnn_oracle.py — float64 numpy restatement of the pose-proposal network (forward and backward).

TEST INFRASTRUCTURE ONLY: the checker for libt2x's proposal-network kernels. Its gradients are
pinned by central finite differences (tests/test_cnn_oracle.py).

Network (channels-last activations [B][H][W][C], weights [Cout][ky][kx][Cin]):
  x [B,64,64,1] -> conv3x3/s2/p1 (32) -> ReLU -> conv (64) -> ReLU -> conv (128) -> ReLU
  -> conv (128) -> ReLU -> flatten [B, 4*4*128] (order y, x, c) -> linear 2048 -> 770
  outputs: logits [B, n_bins] (n_bins = 4 classes x 192 HEALPix order-2 directions) and a 2-D
  shift regression [B, 2] (offsets divided by the maximum shift).
Loss: mean over the batch of softmax cross-entropy on the bin + shift_weight * squared error.
"""
import numpy as np

CHANNELS = (1, 32, 64, 128, 128)
IN_SIZE = 64


def param_shapes(n_bins=768):
    shapes = []
    for i in range(4):
        shapes.append((f"w{i}", (CHANNELS[i + 1], 3, 3, CHANNELS[i])))
        shapes.append((f"b{i}", (CHANNELS[i + 1],)))
    feat = (IN_SIZE // 16) ** 2 * CHANNELS[4]
    shapes.append(("wh", (n_bins + 2, feat)))
    shapes.append(("bh", (n_bins + 2,)))
    return shapes


def im2col(a):
    """[B,H,W,C] -> [B*Ho*Wo, 9*C] for a 3x3 stride-2 pad-1 convolution, k order (ky, kx, c)."""
    B, H, W, C = a.shape
    Ho, Wo = H // 2, W // 2
    p = np.zeros((B, H + 2, W + 2, C), a.dtype)
    p[:, 1:-1, 1:-1, :] = a
    cols = np.empty((B, Ho, Wo, 3, 3, C), a.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky, kx, :] = p[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2, :]
    return cols.reshape(B * Ho * Wo, 9 * C)


def col2im(cols, shape):
    """adjoint of im2col"""
    B, H, W, C = shape
    Ho, Wo = H // 2, W // 2
    c = cols.reshape(B, Ho, Wo, 3, 3, C)
    p = np.zeros((B, H + 2, W + 2, C), cols.dtype)
    for ky in range(3):
        for kx in range(3):
            p[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2, :] += c[:, :, :, ky, kx, :]
    return p[:, 1:-1, 1:-1, :]


def forward(params, x):
    """x [B,64,64] -> (logits, shift, cache)"""
    a = x[..., None].astype(np.float64)
    cache = []
    for i in range(4):
        w, b = params[f"w{i}"], params[f"b{i}"]
        B, H, W, C = a.shape
        cols = im2col(a)
        z = cols @ w.reshape(w.shape[0], -1).T + b
        out = np.maximum(z, 0.0)
        cache.append((a.shape, cols, out))
        a = out.reshape(B, H // 2, W // 2, w.shape[0])
    flat = a.reshape(a.shape[0], -1)
    y = flat @ params["wh"].T + params["bh"]
    return y[:, :-2], y[:, -2:], (cache, flat)


def loss_and_grad(params, x, label, shift_target, shift_weight=1.0):
    logits, shift, (cache, flat) = forward(params, x)
    B = x.shape[0]
    m = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - m)
    p = e / e.sum(axis=1, keepdims=True)
    ce = -np.log(p[np.arange(B), label])
    d = shift - shift_target
    loss = float(np.mean(ce + shift_weight * (d ** 2).sum(axis=1)))
    dlog = p.copy()
    dlog[np.arange(B), label] -= 1.0
    dlog /= B
    dshift = 2.0 * shift_weight * d / B
    dy = np.concatenate([dlog, dshift], axis=1)
    g = {"wh": dy.T @ flat, "bh": dy.sum(axis=0)}
    da = (dy @ params["wh"]).reshape(B, IN_SIZE // 16, IN_SIZE // 16, CHANNELS[4])
    for i in reversed(range(4)):
        shape_in, cols, out = cache[i]
        dz = da.reshape(-1, out.shape[1]) * (out > 0)
        w = params[f"w{i}"]
        g[f"w{i}"] = (dz.T @ cols).reshape(w.shape)
        g[f"b{i}"] = dz.sum(axis=0)
        if i > 0:
            da = col2im(dz @ w.reshape(w.shape[0], -1), shape_in)
    return loss, g, (logits, shift)


def init_params(rng, n_bins=768):
    """He-uniform weights, zero biases (float64)"""
    out = {}
    for name, shape in param_shapes(n_bins):
        if name.startswith("b"):
            out[name] = np.zeros(shape)
        else:
            fan_in = int(np.prod(shape[1:]))
            lim = np.sqrt(6.0 / fan_in)
            out[name] = rng.uniform(-lim, lim, shape)
    return out
