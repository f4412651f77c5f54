"""The batch graph (t2x_estep_config.launch_graph): from the second batch of a size on, a batch's
launches after its preprocessing (coarse pass and selection, the batch plan, the fine pass as a
persistent work-queue launch, weights, the store as a persistent work-queue launch) run as one
captured hipGraph, replayed for the call's further batches of that size.

Every batch of a graph-mode call, the captured one and each replay, is checked against the same
call run eagerly: identical max-weight poses, dLL and fine-sample counts per particle, weighted
sums to 1e-9 and backprojections to 1e-6 (run to run, eager or not, these differ in the order of
their atomic additions: a store item sums its significant samples in FP32 in the order its
threads claimed them, and the items add into FP64 / FP32 accumulators; ~6e-11 and ~1e-7 seen). Graph-mode calls alternate with eager calls in the same process, and the
graph is captured anew in each call (a graph serves the call that captured it).
"""
import ctypes as C

import numpy as np
import pytest

import estep_harness as H
import t2x_py as T

pytestmark = pytest.mark.gpu


def same(g, ref):
    assert np.array_equal(g["out"]["best_hidden"], ref["out"]["best_hidden"])
    assert np.array_equal(g["out"]["dLL"], ref["out"]["dLL"])
    assert np.array_equal(g["out"]["nr_fine_samples"], ref["out"]["nr_fine_samples"])
    for k, v in ref["wsum"].items():
        if np.abs(v).sum() > 0:
            assert H.rel_norm(g["wsum"][k], v) < 1e-9, k
    for c in range(len(ref["bp"])):
        for q in (0, 1):
            if np.abs(ref["bp"][c][q]).sum() > 0:
                assert H.rel_norm(g["bp"][c][q], ref["bp"][c][q]) < 1e-6, (c, q)


@pytest.mark.parametrize("prec", [T.T2X_PREC_FP32, T.T2X_PREC_FP16X2])
def test_graph_batches_equal_eager(gpu_device, monkeypatch, prec):
    monkeypatch.setenv("T2X_BATCH", "4")   # 22 particles: batches of 4, 4, 4, 4, 4 and 2
    lib = T.load()
    wl = H.Workload("refine64", n=22, seed_offset=51, precision=prec)
    eager = wl.run_gpu(dev=gpu_device)
    for rep in range(2):
        wl.cfg.launch_graph = 1
        T.check(lib.t2x_stats_reset(gpu_device))
        g = wl.run_gpu(dev=gpu_device)
        st = T.Stats()
        T.check(lib.t2x_stats_get(gpu_device, C.byref(st)))
        wl.cfg.launch_graph = 0
        # first batch of 4 eager, the second captured, three replays; the batch of 2 eager
        print(f"\n[graph] precision {prec} call {rep}: {st.n_graph_launches} graph launches, "
              f"{st.ms_graph:.3f} ms")
        assert st.n_graph_launches == 4
        same(g, eager)
        again = wl.run_gpu(dev=gpu_device)   # eager work between graph-mode calls
        same(again, eager)


def test_graph_ignored_with_diagnostics(gpu_device, monkeypatch):
    """With diagnostic outputs (or the proposal, or local searches) a call runs eagerly."""
    monkeypatch.setenv("T2X_BATCH", "4")
    lib = T.load()
    wl = H.Workload("refine64", n=12, seed_offset=52)
    ref = wl.run_gpu(dev=gpu_device, n_debug=12, fine_cap=16384)
    wl.cfg.launch_graph = 1
    T.check(lib.t2x_stats_reset(gpu_device))
    g = wl.run_gpu(dev=gpu_device, n_debug=12, fine_cap=16384)
    wl.cfg.launch_graph = 0
    st = T.Stats()
    T.check(lib.t2x_stats_get(gpu_device, C.byref(st)))
    assert st.n_graph_launches == 0
    same(g, ref)
    assert np.array_equal(g["dbg"]["coarse_diff2"], ref["dbg"]["coarse_diff2"])
