"""libt2x on the GPU against the FP64 oracle on identical float32 inputs.

Tolerances: diff2 within 1e-5 relative, identical max-weight pose indices for
>= 99.9% of particles, backprojected data/weight volumes within 1e-4 relative (Frobenius norm),
weighted sums within 1e-4. Weights are exp(-diff2) so an FP32 diff2 error of ~3e-5 absolute
shows as ~3e-5 relative in the normalised weights; they are held to 1e-3.
"""
import numpy as np
import pytest

import estep_harness as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def refine64_pair(gpu_device):
    wl = H.Workload("refine64", n=24)
    o = wl.run_oracle(n_debug=24, fine_cap=16384)
    g = wl.run_gpu(dev=gpu_device, n_debug=24, fine_cap=16384)
    return wl, o, g


def test_coarse_diff2(refine64_pair):
    wl, o, g = refine64_pair
    od, gd = o["dbg"], g["dbg"]
    valid = od["coarse_diff2"] > -998
    assert np.array_equal(valid, gd["coarse_diff2"] > -998)
    err = H.rel(gd["coarse_diff2"][valid], od["coarse_diff2"][valid])
    print("coarse diff2 max rel", err)
    assert err < 1e-5
    assert H.rel(gd["coarse_min_diff2"], od["coarse_min_diff2"]) < 1e-5


def test_coarse_significance(refine64_pair):
    wl, o, g = refine64_pair
    same = (o["dbg"]["coarse_significant"] == g["dbg"]["coarse_significant"]).all(axis=1)
    assert same.mean() >= 0.99, same
    assert (o["out"]["nr_significant_coarse"] == g["out"]["nr_significant_coarse"]).mean() >= 0.99


def test_fine_diff2(refine64_pair):
    wl, o, g = refine64_pair
    od, gd = o["dbg"], g["dbg"]
    worst = 0.0
    for i in range(wl.n):
        no, ng = od["fine_count"][i], gd["fine_count"][i]
        if no != ng or not np.array_equal(od["fine_hidden"][i, :no], gd["fine_hidden"][i, :ng]):
            continue   # significance differed for this particle; covered by the test above
        e = H.rel(gd["fine_diff2"][i, :ng], od["fine_diff2"][i, :no])
        worst = max(worst, e)
        assert e < 1e-5
    print("fine diff2 max rel", worst)


def test_best_pose_and_metadata(refine64_pair):
    wl, o, g = refine64_pair
    oo, go = o["out"], g["out"]
    agree = (oo["best_hidden"] == go["best_hidden"]).mean()
    print("max-weight pose agreement", agree)
    assert agree >= 0.999
    for k in ("rot", "tilt", "psi", "xoff", "yoff", "iclass"):
        same = oo["best_hidden"] == go["best_hidden"]
        assert np.allclose(oo[k][same], go[k][same], atol=1e-9)
    assert H.rel(go["min_diff2"], oo["min_diff2"]) < 1e-5
    assert H.rel(go["dLL"], oo["dLL"]) < 1e-5
    assert H.rel(go["sum_weight"], oo["sum_weight"]) < 1e-3
    assert H.rel(go["pmax"], oo["pmax"]) < 1e-3
    # the count of significant fine samples is identical, except where the oracle holds another
    # weight within 1e-3 (relative) of its significance threshold: a near-tie that the FP32
    # rounding of diff2 (~1e-4 absolute at |diff2| ~ 400, i.e. 1e-4 in the weights) can put on
    # either side of it
    od = o["dbg"]
    for i in np.flatnonzero(oo["nr_fine_significant"] != go["nr_fine_significant"]):
        w = od["fine_weight"][i, :od["fine_count"][i]]
        near = np.abs(w / oo["significant_weight"][i] - 1.0) < 1e-3
        print(f"particle {i}: significant fine samples {oo['nr_fine_significant'][i]} (oracle) vs "
              f"{go['nr_fine_significant'][i]}; oracle weights within 1e-3 of the threshold: {near.sum()}")
        assert abs(int(oo["nr_fine_significant"][i]) - int(go["nr_fine_significant"][i])) <= near.sum() - 1
    assert (oo["nr_fine_significant"] == go["nr_fine_significant"]).mean() >= 0.9


def test_weighted_sums(refine64_pair):
    wl, o, g = refine64_pair
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            e = H.rel_norm(g["wsum"][k], v)
            print("wsum", k, e)
            assert e < 1e-4, k


def test_backprojection(refine64_pair):
    wl, o, g = refine64_pair
    (bdo, bwo), (bdg, bwg) = o["bp"][0], g["bp"][0]
    ed, ew = H.rel_norm(bdg, bdo), H.rel_norm(bwg, bwo)
    print("BP data", ed, "weight", ew)
    assert ed < 1e-4 and ew < 1e-4


def test_multiclass_and_offsets(gpu_device):
    """K=2 classes, non-zero old offsets and offset priors (integer re-centring, offset priors)."""
    wl = H.Workload("refine64", n=12, K=2, seed_offset=5, old_offsets=True)
    o = wl.run_oracle()
    g = wl.run_gpu(dev=gpu_device)
    agree = (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean()
    assert agree >= 0.999
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    for c in range(2):
        assert H.rel_norm(g["bp"][c][0], o["bp"][c][0]) < 1e-4
        assert H.rel_norm(g["bp"][c][1], o["bp"][c][1]) < 1e-4
    for k in ("pdf_class", "sigma2_offset", "sigma2_noise"):
        assert H.rel_norm(g["wsum"][k], o["wsum"][k]) < 1e-4


def test_full_size_properties(gpu_device):
    """Size-independent properties on the 256^2, K=4 workload (too large for the oracle):
    normalised weights sum to one per particle, the weight volume gains exactly the sum of
    splatted weights, and noise-light particles are assigned their true class."""
    import ctypes as C
    import t2x_py as T
    wl = H.Workload("class256", n=8, snr=1.0)
    g = wl.run_gpu(dev=gpu_device)
    out = g["out"]
    assert np.all(out["pmax"] > 0) and np.all(out["pmax"] <= 1.0 + 1e-12)
    assert np.all(out["nr_significant_coarse"] >= 1)
    # sum over particles of normalised weights = number of particles
    assert g["wsum"]["sumw_group"][0] == pytest.approx(wl.n, rel=1e-6)
    assert g["wsum"]["pdf_class"].sum() == pytest.approx(wl.n, rel=1e-6)
    assert (out["iclass"] - 1 == wl.truth["cls"]).mean() >= 0.99
    assert C is not None and T is not None


def test_radix_select_fallback_matches(gpu_device, monkeypatch):
    """The radix-select path (used when the candidate buffer overflows) gives the same
    significant sets and poses as the candidate path and the oracle."""
    monkeypatch.setenv("T2X_FORCE_RADIX_SELECT", "1")
    wl = H.Workload("refine64", n=8, seed_offset=11)
    o = wl.run_oracle(n_debug=8, fine_cap=16384)
    g = wl.run_gpu(dev=gpu_device, n_debug=8, fine_cap=16384)
    assert (o["dbg"]["coarse_significant"] == g["dbg"]["coarse_significant"]).all(axis=1).mean() >= 0.99
    assert (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean() >= 0.999
    assert H.rel_norm(g["bp"][0][1], o["bp"][0][1]) < 1e-4


def test_radix_select_per_particle(gpu_device, monkeypatch):
    """The radix select is chosen per particle on the device (a particle whose candidates
    overflow takes it, the others keep the candidate walk): with every odd-numbered particle
    forced onto it, the batch's significant sets, poses and dLL equal the all-candidate run and
    the oracle's."""
    wl = H.Workload("refine64", n=10, seed_offset=14)
    o = wl.run_oracle(n_debug=10, fine_cap=16384)
    ref = wl.run_gpu(dev=gpu_device, n_debug=10, fine_cap=16384)
    monkeypatch.setenv("T2X_FORCE_RADIX_SELECT", "odd")
    g = wl.run_gpu(dev=gpu_device, n_debug=10, fine_cap=16384)
    assert np.array_equal(g["dbg"]["coarse_significant"], ref["dbg"]["coarse_significant"])
    assert np.array_equal(g["out"]["best_hidden"], ref["out"]["best_hidden"])
    assert np.array_equal(g["out"]["dLL"], ref["out"]["dLL"])
    assert (o["dbg"]["coarse_significant"] == g["dbg"]["coarse_significant"]).all(axis=1).mean() >= 0.99
    assert (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean() >= 0.999
    assert H.rel_norm(g["bp"][0][0], ref["bp"][0][0]) < 1e-6


@pytest.mark.parametrize("n_debug", [0, 12])
def test_batch_plan_grows_its_lists(gpu_device, monkeypatch, n_debug):
    """The work lists of a batch are planned on the device; when they are too small for the
    significant pairs (forced here with a capacity of one pair) the plan reports its totals, no
    kernel after it does any work, and the batch runs again with lists of that size: the
    results equal those of a run planned right the first time (poses, dLL; weighted sums and
    backprojections up to the order of their atomic additions)."""
    wl = H.Workload("refine64", n=12, seed_offset=15)
    ref = wl.run_gpu(dev=gpu_device, n_debug=n_debug, fine_cap=16384)
    monkeypatch.setenv("T2X_PLAN_CAP", "1")
    g = wl.run_gpu(dev=gpu_device, n_debug=n_debug, fine_cap=16384)
    assert np.array_equal(g["out"]["best_hidden"], ref["out"]["best_hidden"])
    assert np.array_equal(g["out"]["dLL"], ref["out"]["dLL"])
    assert np.array_equal(g["out"]["nr_fine_samples"], ref["out"]["nr_fine_samples"])
    for k, v in ref["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-9, k
    assert H.rel_norm(g["bp"][0][0], ref["bp"][0][0]) < 1e-6
    assert H.rel_norm(g["bp"][0][1], ref["bp"][0][1]) < 1e-6


def compare_all(wl, o, g, tol_d2=1e-5):
    """The parity bars of this file: coarse diff2, significant sets, poses, fine-sample counts,
    dLL, weighted sums and both backprojected volumes."""
    od, gd = o["dbg"], g["dbg"]
    valid = od["coarse_diff2"] > -998
    assert np.array_equal(valid, gd["coarse_diff2"] > -998)
    assert H.rel(gd["coarse_diff2"][valid], od["coarse_diff2"][valid]) < tol_d2
    assert (od["coarse_significant"] == gd["coarse_significant"]).all(axis=1).mean() >= 0.99
    assert np.array_equal(o["out"]["best_hidden"], g["out"]["best_hidden"])
    assert np.array_equal(o["out"]["nr_fine_samples"], g["out"]["nr_fine_samples"])
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-4, k
    for c in range(wl.cfg.nr_classes):
        if np.abs(o["bp"][c][1]).sum() > 0:
            assert H.rel_norm(g["bp"][c][0], o["bp"][c][0]) < 1e-4, c
            assert H.rel_norm(g["bp"][c][1], o["bp"][c][1]) < 1e-4, c


@pytest.mark.parametrize("over", [0, 2])
def test_adaptive_oversampling(gpu_device, over):
    """adaptive_oversampling 0: one pass over the coarse sampling at the current size whose
    weights, normalised over every sample, are stored directly (nr_sampling_passes = 1,
    ml_optimiser.cpp:4373-4420); 2: a fine pass at HEALPix order + 2 and a quarter of the psi and
    offset steps (64 rotations x 16 translations per significant coarse sample,
    healpix_sampling.cpp:1720-1735)."""
    wl = H.Workload("refine64", n=6, seed_offset=61)
    wl.cfg.adaptive_oversampling = over
    o = wl.run_oracle(n_debug=6, fine_cap=1 << 16)
    g = wl.run_gpu(dev=gpu_device, n_debug=6, fine_cap=1 << 16)
    nf = g["out"]["nr_fine_samples"]
    print(f"\nadaptive_oversampling {over}: fine samples per particle {nf.min()}..{nf.max()}")
    assert np.all(nf % (1 if over == 0 else 1024) == 0)
    compare_all(wl, o, g)


def test_wide_offset_range(gpu_device):
    """More than 256 coarse translations (offset range 12 px at 1 px: 441 within the circle)."""
    wl = H.Workload("refine64", n=4, seed_offset=62, overrides=dict(offset_range_px=12.0, offset_step_px=1.0))
    assert wl.info.nr_trans > 256
    o = wl.run_oracle(n_debug=4, fine_cap=1 << 16)
    g = wl.run_gpu(dev=gpu_device, n_debug=4, fine_cap=1 << 16)
    print(f"\n{wl.info.nr_trans} coarse translations")
    compare_all(wl, o, g)


def test_maximum_significants(gpu_device):
    """--maxsig caps the coarse significant samples (ml_optimiser.cpp:7950-7961)."""
    wl = H.Workload("refine64", n=8, seed_offset=12, maximum_significants=3)
    o = wl.run_oracle(n_debug=8, fine_cap=16384)
    g = wl.run_gpu(dev=gpu_device, n_debug=8, fine_cap=16384)
    assert np.all(o["out"]["nr_significant_coarse"] <= 3)
    assert np.array_equal(o["out"]["nr_significant_coarse"], g["out"]["nr_significant_coarse"])
    assert (o["dbg"]["coarse_significant"] == g["dbg"]["coarse_significant"]).all(axis=1).mean() >= 0.99
    assert (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean() >= 0.999


def test_lazy_stores_leave_results_unchanged(gpu_device):
    """Without diagnostic outputs the coarse GEMM skips storing blocks whose minimum is beyond the
    negligible-weight cut of a running per-particle minimum; with them every value is stored.
    Poses and dLL are identical and backprojections agree up to atomic-add ordering."""
    wl = H.Workload("refine64", n=16, seed_offset=13)
    a = wl.run_gpu(dev=gpu_device)                        # lazy stores
    b = wl.run_gpu(dev=gpu_device, n_debug=16, fine_cap=16384)   # all values stored
    assert np.array_equal(a["out"]["best_hidden"], b["out"]["best_hidden"])
    assert np.array_equal(a["out"]["dLL"], b["out"]["dLL"])
    assert H.rel_norm(a["bp"][0][1], b["bp"][0][1]) < 1e-6


def test_empty_and_single_particle(gpu_device):
    """No particles: the call succeeds and adds nothing; one particle: parity with the oracle."""
    wl = H.Workload("refine64", n=1, seed_offset=14)
    o = wl.run_oracle()
    g = wl.run_gpu(dev=gpu_device)
    assert g["out"]["best_hidden"][0] == o["out"]["best_hidden"][0]
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    wl.n = 0
    e = wl.run_gpu(dev=gpu_device)
    assert all(np.abs(v).sum() == 0 for v in e["wsum"].values())
    assert np.abs(e["bp"][0][1]).sum() == 0


@pytest.fixture(scope="module")
def class256_case():
    """The benchmark workload's own shapes (256^2, K=4, HEALPix 3, 7.2 M coarse samples per
    particle), 4 particles: small enough for the FP64 oracle."""
    wl = H.Workload("class256", n=4)
    return wl, wl.run_oracle(n_debug=4, fine_cap=1 << 16)


# coarse diff2 bar per precision: FP16x2 rounds the reference rows to 11 significant bits
CLASS256_D2_BAR = {0: 1e-5, 2: 1e-5, 3: 2e-4}


@pytest.mark.parametrize("prec", [0, 2, 3])
def test_class256_shapes(class256_case, gpu_device, prec):
    wl, o = class256_case
    wl.cfg.precision = prec
    g = wl.run_gpu(dev=gpu_device, n_debug=4, fine_cap=1 << 16)
    wl.cfg.precision = 0
    od, gd = o["dbg"], g["dbg"]
    valid = od["coarse_diff2"] > -998
    assert np.array_equal(valid, gd["coarse_diff2"] > -998)
    e = H.rel(gd["coarse_diff2"][valid], od["coarse_diff2"][valid])
    print(f"class256 shapes precision {prec}: coarse diff2 max rel {e:.2e}")
    assert e < CLASS256_D2_BAR[prec]
    assert np.array_equal(o["out"]["best_hidden"], g["out"]["best_hidden"])
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-4, k
    for c in range(wl.cfg.nr_classes):
        if np.abs(o["bp"][c][1]).sum() > 0:
            assert H.rel_norm(g["bp"][c][0], o["bp"][c][0]) < 1e-4, c
            assert H.rel_norm(g["bp"][c][1], o["bp"][c][1]) < 1e-4, c


def test_norm_correction(gpu_device):
    """--norm: images scaled by avg_norm_correction / norm_correction before masking
    (ml_optimiser.cpp:5960-5966) and the updated norm_correction = old / avg * sqrt(2 wsum_norm)
    (ml_optimiser.cpp:8829-8841) with its average summed into wsum_model.avg_norm_correction."""
    wl = H.Workload("refine64", n=12, seed_offset=21, old_offsets=True)
    wl.cfg.do_norm_correction = 1
    wl.cfg.avg_norm_correction = 0.93
    rng = np.random.default_rng(7)
    for i in range(wl.n):
        wl.pin[i].norm_correction = float(rng.uniform(0.8, 1.25))
    o = wl.run_oracle()
    g = wl.run_gpu(dev=gpu_device)
    assert np.array_equal(o["out"]["best_hidden"], g["out"]["best_hidden"])
    nc_in = np.array([wl.pin[i].norm_correction for i in range(wl.n)])
    assert np.all(np.abs(o["out"]["norm_correction"] - nc_in) > 1e-3)   # the update ran
    assert H.rel(g["out"]["norm_correction"], o["out"]["norm_correction"]) < 1e-4
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    assert g["wsum"]["avg_norm_correction"][0] == pytest.approx(o["wsum"]["avg_norm_correction"][0], rel=1e-4)
    assert o["wsum"]["avg_norm_correction"][0] == pytest.approx(o["out"]["norm_correction"].sum(), rel=1e-12)
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-4, k
    assert H.rel_norm(g["bp"][0][0], o["bp"][0][0]) < 1e-4
    assert H.rel_norm(g["bp"][0][1], o["bp"][0][1]) < 1e-4
