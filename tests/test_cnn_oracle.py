"""Pins of the proposal-network oracle: backward pass against central finite differences of the
forward loss (float64), im2col/col2im adjointness, and the im2col convolution against a direct
sum over the 3x3 stride-2 window."""
import numpy as np

import cnn_oracle as N


def test_im2col_adjoint_and_direct_conv():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2, 8, 8, 3))
    cols = N.im2col(a)
    c = rng.standard_normal(cols.shape)
    # <im2col(a), c> == <a, col2im(c)>
    assert np.isclose((cols * c).sum(), (a * N.col2im(c, a.shape)).sum(), rtol=1e-12)
    w = rng.standard_normal((5, 3, 3, 3))
    z = (cols @ w.reshape(5, -1).T).reshape(2, 4, 4, 5)
    for b in range(2):
        for y in range(4):
            for x in range(4):
                s = np.zeros(5)
                for ky in range(3):
                    for kx in range(3):
                        yy, xx = 2 * y + ky - 1, 2 * x + kx - 1
                        if 0 <= yy < 8 and 0 <= xx < 8:
                            s += w[:, ky, kx, :] @ a[b, yy, xx, :]
                assert np.allclose(z[b, y, x], s, rtol=1e-12, atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    params = N.init_params(rng, n_bins=12)
    for k in params:
        if k.startswith("b"):
            params[k] = 0.05 * rng.standard_normal(params[k].shape)   # exercise the bias paths
    B = 3
    x = rng.standard_normal((B, 64, 64))
    label = np.array([1, 7, 11])
    st = rng.uniform(-1, 1, (B, 2))
    loss, g, _ = N.loss_and_grad(params, x, label, st, shift_weight=0.5)
    h = 1e-6
    for name in params:
        flat = params[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _, _ = N.loss_and_grad(params, x, label, st, shift_weight=0.5)
            flat[idx] = old - h
            lm, _, _ = N.loss_and_grad(params, x, label, st, shift_weight=0.5)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            an = g[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 + 1e-5 * abs(fd), (name, idx, fd, an)
