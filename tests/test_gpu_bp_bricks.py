"""Backprojection through LDS bricks (t2x_estep_config.bp_direct 0, the default) against the
direct scatter (bp_direct 1: each pixel's eight trilinear contributions, or four bilinear ones
for 2D references, added to the volume with global atomics: backproject2Dto3D and backrotate2D,
backprojector.cpp:136-357 and 568-643) and against the FP64 oracle.

Both paths insert the same FP32 contributions (positions in double without contraction, the
Friedel flip, the bounds and the trilinear weights are computed alike); only the order of the
FP32 additions differs. The volumes therefore agree to the rounding of FP32 sums: relative
Frobenius 1e-6 (two direct runs differ by ~1e-7, their atomics landing in a different order).
Every other output comes from the same kernels: poses, dLL and fine-sample counts identical,
weighted sums to 1e-9 (the run-to-run order of the store's FP64 atomics).

The brick path has three forms, each checked against the direct scatter: one workgroup per
brick that adds its LDS copy to the volume with a plain read-modify-write (T2X_BP_UNIT_ITEMS
large), the items of each brick shared among up to 256 workgroups that add theirs with global
atomics (T2X_BP_UNIT_ITEMS 1), and the default mix (shared near the origin, where every plane
passes). The items of a batch beyond the brick buffers (a capacity forced down to one item with
T2X_BP_ITEM_CAP) are scattered directly, in the same launch; t2x_stats.bp_brick_items counts
the items that took the bricks. Against the oracle the brick volumes meet the FP32 path's bar
(relative Frobenius 1e-4).
"""
import ctypes as C

import numpy as np
import pytest

import estep_harness as H
import t2x_py as T

pytestmark = pytest.mark.gpu


def run_counted(wl, dev, direct):
    lib = T.load()
    wl.cfg.bp_direct = 1 if direct else 0
    T.check(lib.t2x_stats_reset(dev))
    r = wl.run_gpu(dev=dev)
    st = T.Stats()
    T.check(lib.t2x_stats_get(dev, C.byref(st)))
    wl.cfg.bp_direct = 0
    assert st.bp_orientations > 0
    return r, (st.bp_brick_items, st.bp_orientations)


def same_but_bp_order(g, ref, bp_tol=1e-6):
    assert np.array_equal(g["out"]["best_hidden"], ref["out"]["best_hidden"])
    assert np.array_equal(g["out"]["dLL"], ref["out"]["dLL"])
    assert np.array_equal(g["out"]["nr_fine_samples"], ref["out"]["nr_fine_samples"])
    for k, v in ref["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-9, k
    worst = 0.
    for c in range(len(ref["bp"])):
        for q in (0, 1):
            assert np.abs(ref["bp"][c][q]).sum() > 0, (c, q)
            e = H.rel_norm(g["bp"][c][q], ref["bp"][c][q])
            worst = max(worst, e)
            assert e < bp_tol, (c, q, e)
    return worst


@pytest.mark.parametrize("name,n,K,batch", [("refine64", 22, 3, 8), ("class256", 12, 2, 6),
                                             ("class2d128", 40, 5, 16)])
@pytest.mark.parametrize("unit_items", ["", "1", "1000000000"])
def test_bricks_equal_direct(gpu_device, monkeypatch, name, n, K, batch, unit_items):
    monkeypatch.setenv("T2X_BATCH", str(batch))
    if unit_items:
        monkeypatch.setenv("T2X_BP_UNIT_ITEMS", unit_items)
    wl = H.Workload(name, n=n, K=K, seed_offset=71)
    direct, cnt_d = run_counted(wl, gpu_device, True)
    brick, cnt_b = run_counted(wl, gpu_device, False)
    assert cnt_d[0] == 0
    assert cnt_b[0] == cnt_b[1]   # every item through the bricks
    worst = same_but_bp_order(brick, direct)
    again, _ = run_counted(wl, gpu_device, True)
    base = max(H.rel_norm(again["bp"][c][q], direct["bp"][c][q]) for c in range(K) for q in (0, 1))
    print(f"\n[bricks] {name} K={K} unit_items={unit_items or 'default'}: bricks vs direct "
          f"{worst:.2e}, direct vs direct {base:.2e}")


def test_bricks_equal_direct_local_search(gpu_device, monkeypatch):
    """Local angular searches (each particle its own in-range orientations) insert through the
    same bricks."""
    from test_gpu_local import localise
    monkeypatch.setenv("T2X_BATCH", "8")
    wl = localise(H.Workload("refine64", n=20, seed_offset=74, K=2), 10.0)
    direct, cnt_d = run_counted(wl, gpu_device, True)
    brick, cnt_b = run_counted(wl, gpu_device, False)
    assert cnt_d[0] == 0
    assert cnt_b[0] == cnt_b[1]
    worst = same_but_bp_order(brick, direct)
    print(f"\n[bricks] local search: bricks vs direct {worst:.2e}")


def test_items_beyond_capacity_scatter_directly(gpu_device, monkeypatch):
    monkeypatch.setenv("T2X_BATCH", "8")
    wl = H.Workload("refine64", n=16, K=2, seed_offset=72)
    brick, cnt_b = run_counted(wl, gpu_device, False)
    assert cnt_b[0] == cnt_b[1]
    monkeypatch.setenv("T2X_BP_ITEM_CAP", "1")   # the first item of each batch, the rest direct
    mixed, cnt_m = run_counted(wl, gpu_device, False)
    assert cnt_m[0] == 2 and cnt_m[1] > 2
    same_but_bp_order(mixed, brick)


@pytest.mark.parametrize("name", ["refine64", "class2d128"])
def test_bricks_vs_oracle(gpu_device, monkeypatch, name):
    monkeypatch.setenv("T2X_BATCH", "6")
    wl = H.Workload(name, n=12, K=2, seed_offset=73)
    o = wl.run_oracle()
    g, cnt = run_counted(wl, gpu_device, False)
    assert cnt[0] == cnt[1]
    for c in range(2):
        e_d = H.rel_norm(g["bp"][c][0], o["bp"][c][0])
        e_w = H.rel_norm(g["bp"][c][1], o["bp"][c][1])
        print(f"\n[bricks vs oracle] {name} class {c}: data {e_d:.2e}, weight {e_w:.2e}")
        assert e_d < 1e-4 and e_w < 1e-4
