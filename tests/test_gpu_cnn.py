"""Proposal-network kernels on the GPU against the float64 numpy oracle (oracle/cnn_oracle.py,
itself pinned by finite differences): loss, outputs and every parameter gradient.

FP32 operand mode: outputs and gradients within 1e-4 relative (Frobenius per tensor; FP32
accumulation over reductions up to 65536 terms); this pins the kernels' indexing exactly.
BF16 operand mode (the training mode: weights, activations and back-propagated gradients are
rounded to BF16, 8-bit mantissa, before every MFMA, FP32 accumulate): outputs within 1e-2; the
rounding compounds along the backward chain (measured 0.5% on the head up to 9% on the first
layer's gradient for random weights), so gradients are held to 0.15 relative and to a cosine
similarity of 0.99 with the exact gradient, which is what the optimiser's step direction needs.
"""
import ctypes as C

import numpy as np
import pytest

import cnn_oracle as N
import estep_harness as H
import t2x_py as T

pytestmark = pytest.mark.gpu


def flat_to_dict(flat, n_bins):
    out, off = {}, 0
    for name, shape in N.param_shapes(n_bins):
        n = int(np.prod(shape))
        out[name] = flat[off:off + n].reshape(shape).astype(np.float64)
        off += n
    assert off == flat.size
    return out


@pytest.mark.parametrize("bf16,tol_out,tol_grad", [(0, 1e-4, 1e-4), (1, 1e-2, 0.15)])
def test_loss_and_gradients_match_oracle(gpu_device, bf16, tol_out, tol_grad):
    lib = T.load()
    n_bins, B = 768, 8
    T.check(lib.t2x_proposal_init(gpu_device, n_bins, 1234), "init")
    n = C.c_longlong()
    T.check(lib.t2x_proposal_num_params(gpu_device, C.byref(n)))
    flat = np.zeros(n.value, np.float32)
    T.check(lib.t2x_proposal_get_params(gpu_device, flat.ctypes.data_as(C.POINTER(C.c_float)), n.value))
    rng = np.random.default_rng(7)
    # non-zero biases exercise the bias gradients
    params = flat_to_dict(flat, n_bins)
    for k in params:
        if k.startswith("b"):
            params[k] = (0.05 * rng.standard_normal(params[k].shape)).astype(np.float32).astype(np.float64)
    flat = np.concatenate([params[k].reshape(-1) for k, _ in N.param_shapes(n_bins)]).astype(np.float32)
    T.check(lib.t2x_proposal_set_params(gpu_device, flat.ctypes.data_as(C.POINTER(C.c_float)), n.value))
    x = rng.standard_normal((B, 64, 64)).astype(np.float32)
    label = rng.integers(0, n_bins, B).astype(np.int32)
    st = rng.uniform(-1, 1, (B, 2)).astype(np.float32)
    loss = C.c_double()
    grad = np.zeros(n.value, np.float32)
    outs = np.zeros((B, n_bins + 2), np.float32)
    T.check(lib.t2x_proposal_loss_grad(gpu_device, B, x.ctypes.data_as(C.POINTER(C.c_float)),
                                       label.ctypes.data_as(C.POINTER(C.c_int)),
                                       st.ctypes.data_as(C.POINTER(C.c_float)), 0.5, bf16, C.byref(loss),
                                       grad.ctypes.data_as(C.POINTER(C.c_float)),
                                       outs.ctypes.data_as(C.POINTER(C.c_float))), "loss_grad")
    params = flat_to_dict(flat, n_bins)
    lo, go, (lg, sh) = N.loss_and_grad(params, x.astype(np.float64), label, st.astype(np.float64), 0.5)
    e_out = H.rel_norm(outs, np.concatenate([lg, sh], axis=1))
    e_loss = abs(loss.value - lo) / abs(lo)
    gg = flat_to_dict(grad, n_bins)
    errs = {k: H.rel_norm(gg[k], go[k]) for k in go}
    print(f"\n[cnn bf16={bf16}] outputs {e_out:.2e} loss {e_loss:.2e} grads " +
          " ".join(f"{k}:{v:.1e}" for k, v in errs.items()))
    assert e_out < tol_out and e_loss < tol_out
    assert max(errs.values()) < tol_grad
    cos = {k: float((gg[k] * go[k]).sum() / (np.linalg.norm(gg[k]) * np.linalg.norm(go[k]))) for k in go}
    assert min(cos.values()) > 0.99, cos
