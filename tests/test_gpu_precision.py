"""Split-precision coarse cross term (BF16x3, FP16x3, FP16x2) against the FP64 oracle.

BF16x3 and FP16x3 compute Re<X, A.P> as Ah.Bh + Ah.Bl + Al.Bh on 16-bit MFMA with FP32
accumulation; FP16x2 drops Ah.Bl, so the reference rows enter with their FP16 hi half alone
(11 significant bits), and keeps FP16x3 for the fine pass and every other cross term. Norms, the
constant term and all normalisation stay in FP32/FP64. The oracle is the same double-precision
CPU E-step that pins the FP32 path. Reported per variant: max relative coarse diff2 error, max
absolute error of the per-particle normalised coarse weights exp(-(d - min d)) / sum,
significant-set agreement and max-weight pose agreement.

Tolerances. FP16x3 with per-row power-of-two scaling keeps ~22 mantissa bits per operand,
below FP32 rounding of the accumulator, so it is held to the FP32 bar (diff2 1e-5 relative,
>= 99.9% identical poses). BF16x3 keeps ~16 bits per operand: its diff2 is held to 1e-4
relative and its poses to the mixed-precision bar of >= 99% agreement. The same bars, one
decade apart, hold the normalised weights (absolute), dLL and the backprojected data and weight
volumes (relative, Frobenius).

FP16x2 rounds each reference element to 2^-12 relative; the cross term, which nearly cancels
the two norms at the low frequencies where signal dominates the 1/sigma2 weighting, carries
that error into diff2 at ~1e-4 relative on this workload. Its coarse diff2 is held to 2e-4
and its normalised coarse weights to 3e-3; a sample on the edge of the significance threshold
can flip, so its significant sets are held to >= 90% agreement (the others to 100%). Its fine
pass is FP16x3, so its poses, dLL and backprojections meet the FP16x3 bars.
"""
import numpy as np
import pytest

import estep_harness as H

pytestmark = pytest.mark.gpu

# name -> (precision, diff2 rel, pose agreement, normalised weight abs, dLL rel, BP rel,
#          significant-set agreement)
VARIANTS = {
    "fp32": (0, 1e-5, 0.999, 1e-4, 1e-5, 1e-4, 1.0),
    "bf16x3": (1, 1e-4, 0.99, 1e-3, 1e-4, 1e-3, 1.0),
    "fp16x3": (2, 1e-5, 0.999, 1e-4, 1e-5, 1e-4, 1.0),
    "fp16x2": (3, 2e-4, 0.999, 3e-3, 1e-5, 1e-4, 0.9),
}


def norm_weights(d2, valid):
    w = np.zeros(d2.shape)
    for i in range(d2.shape[0]):
        v = valid[i]
        d = d2[i, v].astype(np.float64)
        e = np.exp(-(d - d.min()))
        w[i, v] = e / e.sum()
    return w


@pytest.fixture(scope="module")
def oracle_ref():
    wl = H.Workload("refine64", n=24, seed_offset=3)
    return wl.run_oracle(n_debug=24, fine_cap=16384)


@pytest.mark.parametrize("name", list(VARIANTS))
def test_split_precision_vs_oracle(name, oracle_ref, gpu_device):
    prec, tol_d2, tol_pose, tol_w, tol_dll, tol_bp, tol_sig = VARIANTS[name]
    wl = H.Workload("refine64", n=24, seed_offset=3, precision=prec)
    o = oracle_ref
    g = wl.run_gpu(dev=gpu_device, n_debug=24, fine_cap=16384)
    od, gd = o["dbg"], g["dbg"]
    valid = od["coarse_diff2"] > -998
    assert np.array_equal(valid, gd["coarse_diff2"] > -998)
    e_d2 = H.rel(gd["coarse_diff2"][valid], od["coarse_diff2"][valid])
    e_w = float(np.abs(norm_weights(gd["coarse_diff2"], valid) -
                       norm_weights(od["coarse_diff2"], valid)).max())
    sig = (od["coarse_significant"] == gd["coarse_significant"]).all(axis=1).mean()
    pose = (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean()
    e_dll = H.rel(g["out"]["dLL"], o["out"]["dLL"])
    e_bp = H.rel_norm(g["bp"][0][1], o["bp"][0][1])
    e_bd = H.rel_norm(g["bp"][0][0], o["bp"][0][0])
    print(f"\n[precision {name}] coarse diff2 max rel {e_d2:.3e}  norm weight max abs {e_w:.3e}  "
          f"sig sets {sig:.4f}  pose {pose:.4f}  dLL rel {e_dll:.3e}  BP data rel {e_bd:.3e}  "
          f"BP weight rel {e_bp:.3e}")
    assert e_d2 < tol_d2
    assert pose >= tol_pose
    assert e_w < tol_w
    assert e_dll < tol_dll
    assert e_bd < tol_bp and e_bp < tol_bp
    assert sig >= tol_sig
