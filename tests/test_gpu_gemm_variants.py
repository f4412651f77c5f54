"""Tile variants of the split-precision coarse GEMM, and the autotuner that picks among them.

Every variant (128x128 k_gemm_x3, 256x256 k_gemm_x3w, 256x128 direct-to-LDS k_gemm_x3d, 256x256
pipelined k_gemm_x3p, each with any tile-walk group) feeds every output element the same chain
of MFMAs (lo.hi, hi.lo, hi.hi per K step of 32, K steps in order) and the same FP64 epilogue
arithmetic, so the coarse diff2 matrix (all of it stored when diagnostics are requested), the
selected poses and dLL must be bit-identical between variants; the autotuner, which times the
eligible variants on the real operands at the first launch of a shape, must report one of them
and change nothing in the results.
"""
import ctypes as C

import numpy as np
import pytest

import estep_harness as H
import t2x_py as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[T.T2X_PREC_FP16X3, T.T2X_PREC_BF16X3], ids=["fp16x3", "bf16x3"])
def case(request, gpu_device):
    wl = H.Workload("refine64", n=24, seed_offset=5, precision=request.param)
    wl.cfg.gemm_tile, wl.cfg.gemm_group = 258, 0
    ref = wl.run_gpu(dev=gpu_device, n_debug=24, fine_cap=16384)
    return wl, ref


@pytest.mark.parametrize("tile,group", [(256, 0), (257, 0), (128, 0), (258, 16), (258, 64), (0, 0)])
def test_variant_bit_identical(case, gpu_device, tile, group):
    wl, ref = case
    lib = T.load()
    wl.cfg.gemm_tile, wl.cfg.gemm_group = tile, group
    T.check(lib.t2x_stats_reset(gpu_device))
    g = wl.run_gpu(dev=gpu_device, n_debug=24, fine_cap=16384)
    st = T.Stats()
    T.check(lib.t2x_stats_get(gpu_device, C.byref(st)))
    if tile:
        assert st.gemm_tile_used == tile and (group == 0 or st.gemm_group_used == group)
    else:
        assert st.gemm_tile_used in (128, 256, 257, 258, 259) and st.gemm_group_used > 0
    print(f"\n[gemm variant] requested ({tile}, {group}) used ({st.gemm_tile_used}, {st.gemm_group_used})")
    assert np.array_equal(g["dbg"]["coarse_diff2"], ref["dbg"]["coarse_diff2"])
    assert np.array_equal(g["out"]["best_hidden"], ref["out"]["best_hidden"])
    assert np.array_equal(g["out"]["dLL"], ref["out"]["dLL"])
    wl.cfg.gemm_tile, wl.cfg.gemm_group = 258, 0


def test_bad_variant_rejected(case, gpu_device):
    wl, _ = case
    lib = T.load()
    wl.cfg.gemm_tile = 300
    try:
        outs = (T.ParticleOut * wl.n)()
        ws, _ = T.make_wsum(wl.cfg.nr_classes, wl.info.nr_dir, wl.N)
        st = lib.t2x_estep(gpu_device, C.byref(wl.cfg), C.byref(wl.model), wl.n,
                           wl.images.ctypes.data_as(C.c_void_p), 0, wl.pin, outs, C.byref(ws), None)
        assert st == T.T2X_EINVAL
    finally:
        wl.cfg.gemm_tile = 258


def test_two_product_kernel_across_k_step_counts(gpu_device):
    """The 2-product kernel (FP16X2) at coarse sizes whose K-step counts (2 ceil(P_c / 32)) fall
    in each residue mod 3 (and so end the two-stage K loop on either stage, after an odd or even
    number of step pairs) gives the oracle's coarse diff2 within the 2-product bar (2e-4
    relative: the reference rows carry 11 significant bits) and its max-weight poses; a stage
    read before its loads landed, or a wrong slice, misses by orders of magnitude."""
    lib = T.load()
    seen = {}
    for cs in (14, 16, 18, 24):
        wl = H.Workload("refine64", n=6, seed_offset=41, current_size=cs, coarse_size=cs,
                        precision=T.T2X_PREC_FP16X2)
        o = wl.run_oracle(n_debug=6, fine_cap=16384)
        T.check(lib.t2x_stats_reset(gpu_device))
        g = wl.run_gpu(dev=gpu_device, n_debug=6, fine_cap=16384)
        st = T.Stats()
        T.check(lib.t2x_stats_get(gpu_device, C.byref(st)))
        nk = 2 * -(-st.P_coarse // 32)
        valid = o["dbg"]["coarse_diff2"] > -998
        assert np.array_equal(valid, g["dbg"]["coarse_diff2"] > -998)
        e = H.rel(g["dbg"]["coarse_diff2"][valid], o["dbg"]["coarse_diff2"][valid])
        print(f"\n[fp16x2] coarse size {cs}: P_c {st.P_coarse}, {nk} K steps, coarse diff2 max rel {e:.2e}")
        assert st.gemm_tile_used in (258, 259)
        assert e < 2e-4, cs
        assert np.array_equal(g["out"]["best_hidden"], o["out"]["best_hidden"]), cs
        seen.setdefault(nk % 3, cs)
    assert set(seen) == {0, 1, 2}, seen


@pytest.fixture(scope="module")
def case_x2(gpu_device):
    wl = H.Workload("refine64", n=24, seed_offset=5, precision=T.T2X_PREC_FP16X2)
    wl.cfg.gemm_tile, wl.cfg.gemm_group = 258, 0
    ref = wl.run_gpu(dev=gpu_device, n_debug=24, fine_cap=16384)
    return wl, ref


@pytest.mark.parametrize("tile,group", [(259, 16), (259, 32), (259, 64), (258, 64), (0, 0)])
def test_two_product_variants_bit_identical(case_x2, gpu_device, tile, group):
    """The two 2-product kernels (k_gemm_x3p<f16, EPI_DIFF2, 2>, tile 258: 256x256 with 8 waves;
    k_gemm_x2h, tile 259: 128x256 with 4 waves, two workgroups per compute unit) feed every element
    the same MFMA chain and epilogue: the whole coarse diff2 matrix, poses and dLL are
    bit-identical across tiles, tile-walk groups and the autotuner's choice."""
    wl, ref = case_x2
    lib = T.load()
    wl.cfg.gemm_tile, wl.cfg.gemm_group = tile, group
    T.check(lib.t2x_stats_reset(gpu_device))
    g = wl.run_gpu(dev=gpu_device, n_debug=24, fine_cap=16384)
    st = T.Stats()
    T.check(lib.t2x_stats_get(gpu_device, C.byref(st)))
    wl.cfg.gemm_tile, wl.cfg.gemm_group = 258, 0
    print(f"\n[fp16x2 variant] requested ({tile}, {group}) used ({st.gemm_tile_used}, {st.gemm_group_used})")
    if tile:
        assert st.gemm_tile_used == tile and st.gemm_group_used == group
    else:
        assert st.gemm_tile_used in (258, 259)
    assert np.array_equal(g["dbg"]["coarse_diff2"], ref["dbg"]["coarse_diff2"])
    assert np.array_equal(g["out"]["best_hidden"], ref["out"]["best_hidden"])
    assert np.array_equal(g["out"]["dLL"], ref["out"]["dLL"])
