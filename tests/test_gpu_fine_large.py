"""The fine pass at high resolution: current sizes whose pixel count the fine pass once could not
hold (the class256 workload at current size 256, P_f ~25,700 pixels; the local360 workload at a
3.5 A state, current size 206), in the three cross-term precisions, against the FP64 oracle.

Bars: fine diff2 within 1e-5 relative (FP32, FP16x3; BF16x3 1e-4), identical max-weight poses,
dLL within 1e-5 (BF16x3 1e-4), weighted sums and backprojections within 1e-4 (BF16x3 1e-3).
"""
import numpy as np
import pytest

import estep_harness as H

pytestmark = pytest.mark.gpu

BARS = {0: (1e-5, 1e-5, 1e-4), 1: (1e-4, 1e-4, 1e-3), 2: (1e-5, 1e-5, 1e-4)}


def compare(wl, o, g, prec, fine=True):
    tol_d2, tol_dll, tol_sum = BARS[prec]
    assert np.array_equal(o["out"]["best_hidden"], g["out"]["best_hidden"])
    e_dll = H.rel(g["out"]["dLL"], o["out"]["dLL"])
    assert e_dll < tol_dll
    e_fine = 0.0
    if fine:
        od, gd = o["dbg"], g["dbg"]
        n_cmp = 0
        for i in range(wl.n):
            no, ng = od["fine_count"][i], gd["fine_count"][i]
            if no != ng or not np.array_equal(od["fine_hidden"][i, :no], gd["fine_hidden"][i, :ng]):
                continue
            e_fine = max(e_fine, H.rel(gd["fine_diff2"][i, :ng], od["fine_diff2"][i, :no]))
            n_cmp += 1
        assert n_cmp >= wl.n - 1
        assert e_fine < tol_d2
    for k, v in o["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < tol_sum, k
    for c in range(wl.cfg.nr_classes):
        if np.abs(o["bp"][c][1]).sum() > 0:
            assert H.rel_norm(g["bp"][c][0], o["bp"][c][0]) < tol_sum, c
            assert H.rel_norm(g["bp"][c][1], o["bp"][c][1]) < tol_sum, c
    return e_fine, e_dll


@pytest.fixture(scope="module")
def class256_cs256():
    """class256 (256^2, K=4, HEALPix 3) at current size 256 with references low-passed to 3 A
    (40-blob references keep the oracle's synthesis of the 515^3 padded volumes short)."""
    wl = H.Workload("class256", n=4, seed_offset=2, current_size=256, snr=0.005,
                    overrides=dict(resolution=3.0, n_blobs=40))
    return wl, wl.run_oracle(n_debug=4, fine_cap=1 << 18)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_class256_current_size_256(class256_cs256, gpu_device, prec):
    wl, o = class256_cs256
    assert wl.cfg.current_size == 256
    wl.cfg.precision = prec
    g = wl.run_gpu(dev=gpu_device, n_debug=4, fine_cap=1 << 18)
    wl.cfg.precision = 0
    e_fine, e_dll = compare(wl, o, g, prec)
    print(f"class256 cs 256 precision {prec}: fine diff2 max rel {e_fine:.2e}, dLL {e_dll:.2e}")


@pytest.fixture(scope="module")
def local360_cs206():
    """local360 (360^2, HEALPix 5, sigma_ang 3 deg) at a 3.5 A state: current size 206."""
    wl = H.Workload("local360", n=3, seed_offset=4, overrides=dict(n_blobs=60, resolution=3.5))
    return wl, wl.run_oracle(n_debug=3, fine_cap=1 << 18)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_local360_current_size_206(local360_cs206, gpu_device, prec):
    wl, o = local360_cs206
    assert wl.cfg.current_size == 206
    wl.cfg.precision = prec
    g = wl.run_gpu(dev=gpu_device, n_debug=3, fine_cap=1 << 18)
    wl.cfg.precision = 0
    e_fine, e_dll = compare(wl, o, g, prec)
    print(f"local360 cs 206 precision {prec}: fine diff2 max rel {e_fine:.2e}, dLL {e_dll:.2e}")
