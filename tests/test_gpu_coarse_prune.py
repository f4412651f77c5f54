"""The pruned coarse pass (t2x_estep_config.coarse_exhaustive 0) against the exhaustive one.

The coarse diff2 of a sample is a sum of non-negative pixel terms (1/2 minv |ctf F - X_t|^2,
plus a constant). A first GEMM over a prefix of the |frequency|-ordered coarse pixels (the low
frequencies), with the rest bounded by Cauchy-Schwarz and a margin for the FP16 product it
drops, gives each 64-column block's lower bound; per particle the full diff2 of its most promising block gives an upper bound of its
minimum (the seed); the full GEMM then skips the tiles whose blocks all lie more than the
negligible-weight cut (60, plus one for rounding) above the seed. Every skipped block is one the
selection already treats as negligible (its minimum beyond the cut), so poses, dLL, the
significant sets, the weighted sums and the backprojections are those of the exhaustive pass:
identical, up to the run-to-run order of the store's atomic sums (weighted sums 1e-9,
backprojections 1e-6).

Covered: the three split precisions at class256 shapes (where most tiles are pruned), a
workload where few are (refine64: the pass turns itself off after a batch that prunes under
40% of its tiles), a prefix radius forced small (T2X_PRUNE_RADIUS), ragged batches, the bounds
GEMM in the full split precision and as each of its kernels, and graph replays.
"""
import ctypes as C

import numpy as np
import pytest

import estep_harness as H
import t2x_py as T

pytestmark = pytest.mark.gpu


def run(wl, dev, exhaustive):
    lib = T.load()
    wl.cfg.coarse_exhaustive = 1 if exhaustive else 0
    T.check(lib.t2x_stats_reset(dev))
    r = wl.run_gpu(dev=dev)
    st = T.Stats()
    T.check(lib.t2x_stats_get(dev, C.byref(st)))
    wl.cfg.coarse_exhaustive = 0
    return r, (st.coarse_tiles_pruned, st.coarse_tiles_checked)


def same(g, ref):
    assert np.array_equal(g["out"]["best_hidden"], ref["out"]["best_hidden"])
    assert np.array_equal(g["out"]["dLL"], ref["out"]["dLL"])
    assert np.array_equal(g["out"]["nr_significant_coarse"], ref["out"]["nr_significant_coarse"])
    assert np.array_equal(g["out"]["nr_fine_samples"], ref["out"]["nr_fine_samples"])
    for k, v in ref["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-9, k
    for c in range(len(ref["bp"])):
        for q in (0, 1):
            if np.abs(ref["bp"][c][q]).sum() > 0:
                assert H.rel_norm(g["bp"][c][q], ref["bp"][c][q]) < 1e-6, (c, q)


@pytest.mark.parametrize("prec", [T.T2X_PREC_FP16X2, T.T2X_PREC_FP16X3, T.T2X_PREC_BF16X3])
def test_pruned_equals_exhaustive_class256(gpu_device, monkeypatch, prec):
    monkeypatch.setenv("T2X_BATCH", "6")   # 14 particles: batches of 6, 6 and 2
    wl = H.Workload("class256", n=14, seed_offset=81, precision=prec)
    ex, cnt_e = run(wl, gpu_device, True)
    pr, cnt_p = run(wl, gpu_device, False)
    print(f"\n[prune] class256 precision {prec}: {cnt_p[0]:.0f} of {cnt_p[1]:.0f} tiles pruned")
    assert cnt_e == (0, 0)
    assert cnt_p[1] > 0 and cnt_p[0] > 0.5 * cnt_p[1]
    same(pr, ex)


def test_pruned_equals_exhaustive_when_little_prunes(gpu_device, monkeypatch):
    monkeypatch.setenv("T2X_BATCH", "8")
    wl = H.Workload("refine64", n=24, K=2, seed_offset=82, precision=T.T2X_PREC_FP16X2)
    ex, _ = run(wl, gpu_device, True)
    pr, cnt = run(wl, gpu_device, False)
    print(f"\n[prune] refine64: {cnt[0]:.0f} of {cnt[1]:.0f} tiles pruned")
    same(pr, ex)


def test_pruned_small_prefix(gpu_device, monkeypatch):
    monkeypatch.setenv("T2X_BATCH", "6")
    monkeypatch.setenv("T2X_PRUNE_RADIUS", "0.5")   # a quarter of the pixels: weak bounds
    wl = H.Workload("class256", n=12, seed_offset=83, precision=T.T2X_PREC_FP16X2)
    ex, _ = run(wl, gpu_device, True)
    pr, cnt = run(wl, gpu_device, False)
    print(f"\n[prune] class256, prefix radius 0.5: {cnt[0]:.0f} of {cnt[1]:.0f} tiles pruned")
    same(pr, ex)


def test_pruned_full_split_bounds(gpu_device, monkeypatch):
    """The bounds pass in the full split precision (T2X_PRUNE_FULL_BOUNDS; the default under FP16
    takes one product per element and allows for the dropped term)."""
    monkeypatch.setenv("T2X_BATCH", "6")
    monkeypatch.setenv("T2X_PRUNE_FULL_BOUNDS", "1")
    wl = H.Workload("class256", n=12, seed_offset=84, precision=T.T2X_PREC_FP16X2)
    ex, _ = run(wl, gpu_device, True)
    pr, cnt = run(wl, gpu_device, False)
    print(f"\n[prune] class256, full-split bounds: {cnt[0]:.0f} of {cnt[1]:.0f} tiles pruned")
    assert cnt[0] > 0.5 * cnt[1]
    same(pr, ex)


def test_pruned_graph_replays_count_tiles(gpu_device, monkeypatch):
    """Batches replayed from a graph prune as the eager ones, and the tile counters see them."""
    lib = T.load()
    monkeypatch.setenv("T2X_BATCH", "4")   # 16 particles: a captured batch and two replays
    wl = H.Workload("class256", n=16, seed_offset=85, precision=T.T2X_PREC_FP16X2)
    eager, cnt_e = run(wl, gpu_device, False)
    wl.cfg.launch_graph = 1
    graph, cnt_g = run(wl, gpu_device, False)
    wl.cfg.launch_graph = 0
    st = T.Stats()
    T.check(lib.t2x_stats_get(gpu_device, C.byref(st)))
    print(f"\n[prune] graph: {st.n_graph_launches} graph launches, {cnt_g[0]:.0f} of {cnt_g[1]:.0f} tiles "
          f"pruned (eager {cnt_e[0]:.0f} of {cnt_e[1]:.0f})")
    assert st.n_graph_launches >= 2
    assert cnt_g == cnt_e
    same(graph, eager)


@pytest.mark.parametrize("tile", ["259", "258"])
def test_pruned_other_bounds_kernels(gpu_device, monkeypatch, tile):
    """The bounds GEMM as k_gemm_x2h<true> (128x256 tiles) or as k_gemm_x3p with one K chunk per
    stage (the default takes two)."""
    monkeypatch.setenv("T2X_BATCH", "6")
    monkeypatch.setenv("T2X_BOUNDS_TILE", tile)
    wl = H.Workload("class256", n=12, seed_offset=86, precision=T.T2X_PREC_FP16X2)
    ex, _ = run(wl, gpu_device, True)
    pr, cnt = run(wl, gpu_device, False)
    print(f"\n[prune] class256, bounds tile {tile}: {cnt[0]:.0f} of {cnt[1]:.0f} tiles pruned")
    assert cnt[0] > 0.5 * cnt[1]
    same(pr, ex)
