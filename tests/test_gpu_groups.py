"""libt2x against the FP64 oracle with two optics groups (own noise spectra), three image groups
with scale corrections (relion_refine --scale) and the noise-filled mask background
(relion_refine without --zero_mask), on identical float32 inputs; and the rejection of group
numbers or scales the E-step cannot serve.

Bars as the other parity tests: identical max-weight poses, dLL within 1e-5 relative, weighted
sums (per optics group and per image group) and backprojected volumes within 1e-4.
"""
import ctypes as C

import numpy as np
import pytest

import estep_harness as H
import t2x_py as T

pytestmark = pytest.mark.gpu


def grouped_workload(precision):
    wl = H.Workload("refine64", n=18, seed_offset=33, old_offsets=True, precision=precision)
    xN = wl.N // 2 + 1
    base = wl.sigma2.copy()
    rows = [base, base * np.linspace(1.4, 0.8, xN)]
    dvp = np.zeros((wl.cfg.nr_classes, xN))
    dvp[:, :12] = 5.0                       # the scale sums run over shells 0..11
    wl.set_groups(rows, group_of=[i % 3 for i in range(wl.n)], optics_of=[(i // 3) % 2 for i in range(wl.n)],
                  scale=[0.85, 1.0, 1.25], dvp=dvp)
    wl.cfg.do_zero_mask = 0
    wl.cfg.random_seed = 4242
    for i in range(wl.n):
        wl.pin[i].part_id = 700 + 3 * i
    return wl


@pytest.mark.parametrize("precision", [0, 2])
def test_groups_scale_and_noise_mask(gpu_device, precision):
    wl = grouped_workload(precision)
    o = wl.run_oracle()
    g = wl.run_gpu(dev=gpu_device)
    oo, go = o["out"], g["out"]
    assert np.array_equal(oo["best_hidden"], go["best_hidden"])
    assert H.rel(go["dLL"], oo["dLL"]) < 1e-5
    assert H.rel(go["min_diff2"], oo["min_diff2"]) < 1e-5
    for k in ("sum_weight", "pmax"):
        assert H.rel(go[k], oo[k]) < 1e-3, k
    ws_o, ws_g = o["wsum"], g["wsum"]
    xN = wl.N // 2 + 1
    # both optics groups and all three image groups received sums
    assert np.all(ws_o["sumw_group"] > 0) and np.all(ws_o["wsum_signal_product"] != 0)
    assert np.all(ws_o["sigma2_noise"][:xN] > 0) and np.all(ws_o["sigma2_noise"][xN:] > 0)
    for k, v in ws_o.items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(ws_g[k], v) < 1e-4, k
    for og in range(2):
        sl = slice(og * xN, (og + 1) * xN)
        assert H.rel_norm(ws_g["sigma2_noise"][sl], ws_o["sigma2_noise"][sl]) < 1e-4
    assert H.rel(ws_g["wsum_signal_product"], ws_o["wsum_signal_product"]) < 1e-4
    assert H.rel(ws_g["wsum_reference_power"], ws_o["wsum_reference_power"]) < 1e-4
    for c in range(wl.cfg.nr_classes):
        assert H.rel_norm(g["bp"][c][0], o["bp"][c][0]) < 1e-4
        assert H.rel_norm(g["bp"][c][1], o["bp"][c][1]) < 1e-4


def test_noise_mask_changes_results(gpu_device):
    """The noise background is a different E-step from the average background (the particles'
    masks reach into the box here): dLL moves on the GPU as in the oracle."""
    wl = grouped_workload(0)
    g_noise = wl.run_gpu(dev=gpu_device)
    wl.cfg.do_zero_mask = 1
    g_avg = wl.run_gpu(dev=gpu_device)
    o_avg = wl.run_oracle()
    assert H.rel(g_avg["out"]["dLL"], o_avg["out"]["dLL"]) < 1e-5
    assert np.max(np.abs(g_noise["out"]["dLL"] - g_avg["out"]["dLL"])) > 1e-3


def test_groups_rejected(gpu_device):
    lib = T.load()
    wl = grouped_workload(0)
    K = wl.cfg.nr_classes
    for c in range(K):
        T.check(lib.t2x_projector_set(gpu_device, c, 3, wl.pad, wl.rmax, 2.0, T.fptr(wl.ref_re[c]),
                                      T.fptr(wl.ref_im[c])))
        T.check(lib.t2x_backprojector_init(gpu_device, c, 3, wl.pad, wl.rmax, 2.0))
    outs = (T.ParticleOut * wl.n)()
    ws, _ = T.make_wsum(K, wl.info.nr_dir, wl.N, 2, 3)

    def call():
        return lib.t2x_estep(gpu_device, C.byref(wl.cfg), C.byref(wl.model), wl.n,
                             wl.images.ctypes.data_as(C.c_void_p), 0, wl.pin, outs, C.byref(ws), None)
    wl.pin[5].group_id = 3                  # three image groups
    assert call() == T.T2X_EINVAL
    wl.pin[5].group_id = 2
    wl.pin[4].optics_group = 2              # two optics groups
    assert call() == T.T2X_EINVAL
    wl.pin[4].optics_group = 1
    wl._scale[1] = 0.0                      # scale corrections must lie in (0, 10000]
    assert call() == T.T2X_EINVAL
    wl._scale[1] = 2.0e4
    assert call() == T.T2X_EINVAL
    assert b"scale" in lib.t2x_last_error()
