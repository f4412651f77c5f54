"""Local angular searches (orientational priors) on the GPU against the FP64 oracle on identical
float32 inputs: per-particle direction / psi priors around the prior angles
(selectOrientationsWithNonZeroPriorProbability, healpix_sampling.cpp:769-1145), orientations
outside the range never scored, weights pdf_orientation / mean * pdf_offset / mean
(ml_optimiser.cpp:7600-7716), pdf_direction sums at the global direction (:8540-8554).

Bars as for global searches: the scored set identical, diff2 within 1e-5 relative, >= 99.9%
identical max-weight poses, weighted sums within 1e-4, backprojections within 1e-4.
"""
import numpy as np
import pytest

import estep_harness as H
import t2x_py as T

pytestmark = pytest.mark.gpu


def localise(wl, sig, sig_psi=None, noise=3.0, seed=5):
    wl.cfg.orientational_prior_mode = T.T2X_PRIOR_ROTTILT_PSI
    wl.cfg.sigma_rot = wl.cfg.sigma_tilt = sig
    wl.cfg.sigma_psi = sig if sig_psi is None else sig_psi
    rng = np.random.default_rng(seed)
    tr = wl.truth
    for i in range(wl.n):
        t = min(max(tr["tilt"][i] + rng.normal(0, noise), 0.0), 180.0)
        p = (tr["psi"][i] + rng.normal(0, noise) + 180.0) % 360.0 - 180.0
        if wl.ref_dim == 3:
            wl.pin[i].prior_rot = tr["rot"][i] + rng.normal(0, noise)
            wl.pin[i].prior_tilt = t
        # the psi prior through the current angle (prior unset): the fallback of the reference
        wl.pin[i].psi = p
    return wl


def compare(o, g, n_dbg=True):
    od, gd = o["dbg"], g["dbg"]
    if n_dbg:
        vo, vg = od["coarse_diff2"] > -998, gd["coarse_diff2"] > -998
        assert np.array_equal(vo, vg), "scored sets differ"
        assert 0 < vo.mean() < 1
        e = H.rel(gd["coarse_diff2"][vo], od["coarse_diff2"][vo])
        print("local coarse diff2 max rel", e, "scored fraction", vo.mean())
        assert e < 1e-5
        assert (od["coarse_significant"] == gd["coarse_significant"]).all(axis=1).mean() >= 0.99
    agree = (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean()
    print("local pose agreement", agree)
    assert agree >= 0.999
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-4, k
    for c in range(len(o["bp"])):
        if np.abs(o["bp"][c][1]).sum() > 0:
            assert H.rel_norm(g["bp"][c][0], o["bp"][c][0]) < 1e-4
            assert H.rel_norm(g["bp"][c][1], o["bp"][c][1]) < 1e-4


def test_local_search_parity(gpu_device):
    wl = localise(H.Workload("refine64", n=24, seed_offset=31, K=2), 10.0)
    o = wl.run_oracle(n_debug=24, fine_cap=16384)
    g = wl.run_gpu(dev=gpu_device, n_debug=24, fine_cap=16384)
    compare(o, g)
    # pdf_direction is summed at the global direction index, inside each particle's range
    assert np.count_nonzero(o["wsum"]["pdf_direction"]) > 0


def test_local_search_lazy_stores(gpu_device):
    """Without diagnostic outputs (lazy stores of negligible blocks) the results are the same."""
    wl = localise(H.Workload("refine64", n=16, seed_offset=32), 8.0, 12.0)
    o = wl.run_oracle()
    g = wl.run_gpu(dev=gpu_device)
    compare(o, g, n_dbg=False)


def test_local_search_fp16x3(gpu_device):
    wl = localise(H.Workload("refine64", n=16, seed_offset=33, precision=T.T2X_PREC_FP16X3), 10.0)
    o = wl.run_oracle(n_debug=16, fine_cap=16384)
    g = wl.run_gpu(dev=gpu_device, n_debug=16, fine_cap=16384)
    compare(o, g)


def test_wide_local_search_equals_global_gpu(gpu_device):
    wl = H.Workload("refine64", n=8, seed_offset=34)
    a = wl.run_gpu(dev=gpu_device)
    wl.cfg.orientational_prior_mode = T.T2X_PRIOR_ROTTILT_PSI
    wl.cfg.sigma_rot = wl.cfg.sigma_tilt = wl.cfg.sigma_psi = 1e8
    b = wl.run_gpu(dev=gpu_device)
    assert np.array_equal(a["out"]["best_hidden"], b["out"]["best_hidden"])
    np.testing.assert_allclose(b["out"]["dLL"], a["out"]["dLL"], rtol=1e-7)
    assert H.rel_norm(b["bp"][0][1], a["bp"][0][1]) < 1e-5


def test_local_psi_search_2d(gpu_device):
    """2D classification with a psi prior only (one direction)."""
    wl = localise(H.Workload("class2d64", n=16, seed_offset=35), 0.0, 15.0)
    o = wl.run_oracle(n_debug=16, fine_cap=16384)
    g = wl.run_gpu(dev=gpu_device, n_debug=16, fine_cap=16384)
    compare(o, g)


def test_local_with_proposal_rejected(gpu_device):
    wl = localise(H.Workload("refine64", n=2, seed_offset=36, precision=T.T2X_PREC_FP16X3), 10.0)
    wl.cfg.proposal_topk = 16
    with pytest.raises(RuntimeError, match="exclusive"):
        wl.run_gpu(dev=gpu_device)


def test_sample_lists_equal_dense_masked_pass(gpu_device, monkeypatch):
    """The local search's coarse pass on per-particle sample lists (default) and the dense pass
    over every orientation with the out-of-range ones masked (T2X_LOCAL_DENSE=1) score the same
    samples: identical scored sets and poses, diff2 within FP32 summation-order differences."""
    wl = localise(H.Workload("refine64", n=16, seed_offset=37, K=2), 9.0, 11.0)
    a = wl.run_gpu(dev=gpu_device, n_debug=16, fine_cap=16384)
    monkeypatch.setenv("T2X_LOCAL_DENSE", "1")
    b = wl.run_gpu(dev=gpu_device, n_debug=16, fine_cap=16384)
    va, vb = a["dbg"]["coarse_diff2"] > -998, b["dbg"]["coarse_diff2"] > -998
    assert np.array_equal(va, vb)
    assert H.rel(a["dbg"]["coarse_diff2"][va], b["dbg"]["coarse_diff2"][vb]) < 2e-6
    assert np.array_equal(a["out"]["best_hidden"], b["out"]["best_hidden"])
    assert H.rel(a["out"]["dLL"], b["out"]["dLL"]) < 2e-6
    for c in range(2):
        assert H.rel_norm(a["bp"][c][1], b["bp"][c][1]) < 1e-5


@pytest.fixture(scope="module")
def local360_case():
    wl = H.Workload("local360", n=3, overrides=dict(n_blobs=60))
    return wl, wl.run_oracle()


# precision -> (dLL, weighted sums / backprojection) relative bars: FP32 and FP16x3 at the FP32
# bars; BF16x3 (16 mantissa bits per operand, diff2 ~1e-5 relative) one decade looser
LOCAL360_BARS = {0: (1e-5, 1e-4), 1: (1e-4, 1e-3), 2: (1e-5, 1e-4)}


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_local360_shapes(local360_case, gpu_device, prec):
    """The local-refinement configuration's own shapes (360^2, HEALPix 5 = 2.36 M orientations,
    sigma_ang 3 deg, 113 translations, P_c 5,007) on 3 particles; 60-blob references keep the
    oracle's reference synthesis short (the shapes do not depend on the blob count). FP32: the
    operand-row GEMM; FP16x3: the grouped GEMM of gathered slices."""
    wl, o = local360_case
    wl.cfg.precision = prec
    g = wl.run_gpu(dev=gpu_device)
    wl.cfg.precision = 0
    tol_dll, tol_sum = LOCAL360_BARS[prec]
    assert np.array_equal(o["out"]["best_hidden"], g["out"]["best_hidden"])
    e_dll = H.rel(g["out"]["dLL"], o["out"]["dLL"])
    e_bd, e_bw = H.rel_norm(g["bp"][0][0], o["bp"][0][0]), H.rel_norm(g["bp"][0][1], o["bp"][0][1])
    print(f"local360 precision {prec}: dLL rel {e_dll:.2e} BP data {e_bd:.2e} weight {e_bw:.2e}")
    assert e_dll < tol_dll
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < tol_sum, k
    assert e_bd < tol_sum and e_bw < tol_sum
