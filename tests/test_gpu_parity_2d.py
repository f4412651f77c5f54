"""2D classification (2D references) on the GPU against the FP64 oracle on identical float32
inputs: in-plane sampling only (one direction, psi oversampled), references sampled and
backprojected by the 2D rotation (Projector::rotate2D, projector.cpp:904-1016;
BackProjector::backrotate2D, backprojector.cpp:471-668), class-wise offset priors
(ml_optimiser.cpp:8605-8640) and their weighted sums.

Bars as for 3D references: diff2 within 1e-5 relative, >= 99.9% identical max-weight poses,
backprojected data/weight planes within 1e-4 relative (Frobenius norm), weighted sums within
1e-4; FP16x3 held to the FP32 bar.
"""
import numpy as np
import pytest

import estep_harness as H
import t2x_py as T

pytestmark = pytest.mark.gpu


def with_class_prior(wl, xy):
    wl.prior_class = np.tile(np.asarray(xy, float), wl.cfg.nr_classes)
    wl.model = T.Model(T.dptr(wl.sigma2), T.dptr(wl.pdf_class), T.dptr(wl.pdf_dir),
                       T.dptr(wl.prior_class))
    return wl


@pytest.fixture(scope="module")
def class2d_pair(gpu_device):
    wl = H.Workload("class2d64", n=24)
    assert wl.ref_dim == 2 and wl.info.nr_dir == 1
    o = wl.run_oracle(n_debug=24, fine_cap=16384)
    g = wl.run_gpu(dev=gpu_device, n_debug=24, fine_cap=16384)
    return wl, o, g


def test_coarse_diff2_2d(class2d_pair):
    wl, o, g = class2d_pair
    od, gd = o["dbg"], g["dbg"]
    valid = od["coarse_diff2"] > -998
    assert np.array_equal(valid, gd["coarse_diff2"] > -998)
    err = H.rel(gd["coarse_diff2"][valid], od["coarse_diff2"][valid])
    print("2D coarse diff2 max rel", err)
    assert err < 1e-5
    same = (od["coarse_significant"] == gd["coarse_significant"]).all(axis=1)
    assert same.mean() >= 0.99


def test_fine_diff2_2d(class2d_pair):
    wl, o, g = class2d_pair
    od, gd = o["dbg"], g["dbg"]
    checked = 0
    for i in range(wl.n):
        no, ng = od["fine_count"][i], gd["fine_count"][i]
        if no != ng or not np.array_equal(od["fine_hidden"][i, :no], gd["fine_hidden"][i, :ng]):
            continue
        assert H.rel(gd["fine_diff2"][i, :ng], od["fine_diff2"][i, :no]) < 1e-5
        checked += 1
    assert checked >= 0.9 * wl.n


def test_best_pose_2d(class2d_pair):
    wl, o, g = class2d_pair
    oo, go = o["out"], g["out"]
    agree = (oo["best_hidden"] == go["best_hidden"]).mean()
    print("2D max-weight pose agreement", agree)
    assert agree >= 0.999
    same = oo["best_hidden"] == go["best_hidden"]
    for k in ("rot", "tilt", "psi", "xoff", "yoff", "iclass"):
        assert np.allclose(oo[k][same], go[k][same], atol=1e-9)
    assert np.all(go["rot"] == 0) and np.all(go["tilt"] == 0)
    assert np.all(np.abs(go["psi"]) <= 180)
    assert H.rel(go["dLL"], oo["dLL"]) < 1e-5
    assert H.rel(go["pmax"], oo["pmax"]) < 1e-3
    # most particles land in their true class
    assert (go["iclass"] - 1 == wl.truth["cls"]).mean() >= 0.75


def test_weighted_sums_and_planes_2d(class2d_pair):
    wl, o, g = class2d_pair
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            e = H.rel_norm(g["wsum"][k], v)
            print("2D wsum", k, e)
            assert e < 1e-4, k
    assert np.abs(o["wsum"]["prior_offset_class"]).sum() > 0
    for c in range(wl.cfg.nr_classes):
        (bdo, bwo), (bdg, bwg) = o["bp"][c], g["bp"][c]
        if np.abs(bwo).sum() == 0:
            assert np.abs(bwg).sum() == 0
            continue
        assert bdg.size == wl.pad * (wl.pad // 2 + 1)
        assert H.rel_norm(bdg, bdo) < 1e-4 and H.rel_norm(bwg, bwo) < 1e-4


def test_class_offset_prior_2d(gpu_device):
    """Equal non-zero class offset priors move the prior centre of every particle's offsets."""
    wl = with_class_prior(H.Workload("class2d64", n=12, seed_offset=4), (1.0, -2.0))
    o = wl.run_oracle()
    g = wl.run_gpu(dev=gpu_device)
    assert (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean() >= 0.999
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    for k in ("prior_offset_class", "sigma2_offset", "pdf_class"):
        assert H.rel_norm(g["wsum"][k], o["wsum"][k]) < 1e-4, k
    base = H.Workload("class2d64", n=12, seed_offset=4).run_oracle()
    assert not np.allclose(base["wsum"]["sigma2_offset"], o["wsum"]["sigma2_offset"])


@pytest.mark.parametrize("prec", [T.T2X_PREC_FP32, T.T2X_PREC_FP16X3])
def test_unequal_class_offset_priors_2d(gpu_device, prec):
    """Class-dependent offset priors (2D classification after its first iteration): one
    translation prior table per class, normalised over every (class, translation)."""
    wl = H.Workload("class2d64", n=12, seed_offset=5, precision=prec)
    pc = np.array([0.0, 0.0, 1.0, -2.0, -1.5, 0.5, 2.0, 2.0])[:2 * wl.cfg.nr_classes]
    wl.prior_class = pc
    wl.model = T.Model(T.dptr(wl.sigma2), T.dptr(wl.pdf_class), T.dptr(wl.pdf_dir), T.dptr(pc))
    o = wl.run_oracle()
    g = wl.run_gpu(dev=gpu_device)
    assert (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean() >= 0.999
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    assert H.rel(g["out"]["sum_weight"], o["out"]["sum_weight"]) < 1e-4
    for k in ("prior_offset_class", "sigma2_offset", "pdf_class"):
        assert H.rel_norm(g["wsum"][k], o["wsum"][k]) < 1e-4, k
    eq = with_class_prior(H.Workload("class2d64", n=12, seed_offset=5, precision=prec), (0.0, 0.0)).run_oracle()
    assert not np.allclose(eq["wsum"]["sigma2_offset"], o["wsum"]["sigma2_offset"])


def test_fp16x3_2d(class2d_pair, gpu_device):
    wl0, o, _ = class2d_pair
    wl = H.Workload("class2d64", n=24, precision=2)
    g = wl.run_gpu(dev=gpu_device, n_debug=24, fine_cap=16384)
    od, gd = o["dbg"], g["dbg"]
    valid = od["coarse_diff2"] > -998
    assert H.rel(gd["coarse_diff2"][valid], od["coarse_diff2"][valid]) < 1e-5
    assert (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean() >= 0.999
    for c in range(wl.cfg.nr_classes):
        if np.abs(o["bp"][c][1]).sum() > 0:
            assert H.rel_norm(g["bp"][c][0], o["bp"][c][0]) < 1e-4
            assert H.rel_norm(g["bp"][c][1], o["bp"][c][1]) < 1e-4


@pytest.fixture(scope="module")
def class2d128_case():
    """The 2D-classification configuration's own shapes: 128^2, 100 class averages, Zipf(1.2)
    class frequencies, per-particle SNR log-uniform in [0.01, 0.1]; 16 particles."""
    wl = H.Workload("class2d128", n=16)
    return wl, wl.run_oracle()


@pytest.mark.parametrize("prec", [0, 2])
def test_class2d128_shapes(class2d128_case, gpu_device, prec):
    wl, o = class2d128_case
    wl.cfg.precision = prec
    g = wl.run_gpu(dev=gpu_device)
    wl.cfg.precision = 0
    assert (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean() >= 0.999
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-4, k
    for c in range(wl.cfg.nr_classes):
        if np.abs(o["bp"][c][1]).sum() > 0:
            assert H.rel_norm(g["bp"][c][1], o["bp"][c][1]) < 1e-4, c
            assert H.rel_norm(g["bp"][c][0], o["bp"][c][0]) < 1e-4, c


@pytest.fixture(scope="module")
def class2d64_many():
    wl = H.Workload("class2d64", n=1500, seed_offset=6)
    return wl, wl.run_oracle()


@pytest.mark.parametrize("batch", [0, 700])
def test_large_and_ragged_batches(class2d64_many, gpu_device, monkeypatch, batch):
    """1,500 particles in one batch (above 1,024), or in batches of 700, 700 and 100."""
    wl, o = class2d64_many
    if batch:
        monkeypatch.setenv("T2X_BATCH", str(batch))
    g = wl.run_gpu(dev=gpu_device)
    agree = (o["out"]["best_hidden"] == g["out"]["best_hidden"]).mean()
    print("1500-particle pose agreement", agree)
    assert agree >= 0.999
    assert H.rel(g["out"]["dLL"], o["out"]["dLL"]) < 1e-5
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-4, k
    for c in range(wl.cfg.nr_classes):
        assert H.rel_norm(g["bp"][c][0], o["bp"][c][0]) < 1e-4
        assert H.rel_norm(g["bp"][c][1], o["bp"][c][1]) < 1e-4
