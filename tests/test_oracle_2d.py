"""2D references (2D classification) in the FP64 oracle, pinned by brute force and closed forms,
and the library's 2D sampling grid against the oracle's. CPU only.

Projector::rotate2D (projector.cpp:904-1016) samples the half-complex 2D reference bilinearly at
Ainv (x, y) with the Friedel flip; BackProjector::backrotate2D (backprojector.cpp:471-668) is its
adjoint on the pixels of the ellipse rows; HealpixSampling for 2D has one direction (rot = tilt =
0), in-plane angles psi = i * 360 / ceil(360 / step), and oversamples psi only, wrapping
oversampled angles above 180 to negative values (healpix_sampling.cpp:590-615, 1720-1735,
2023-2050).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
import t2x_py as T
from t2x_py import workloads as W


@pytest.fixture(scope="module")
def orc():
    return O.load()


def _bilinear_numpy(ref, pad, xp, yp):
    xdim, h = pad // 2 + 1, pad // 2
    R = (ref[0::2] + 1j * ref[1::2]).reshape(pad, xdim)
    neg = xp < 0
    if neg:
        xp, yp = -xp, -yp
    x0, y0 = int(np.floor(xp)), int(np.floor(yp))
    fx, fy = xp - x0, yp - y0
    y0 += h
    val = ((1 - fy) * ((1 - fx) * R[y0, x0] + fx * R[y0, x0 + 1]) +
           fy * ((1 - fx) * R[y0 + 1, x0] + fx * R[y0 + 1, x0 + 1]))
    return np.conj(val) if neg else val


def _euler_2d(orc, psi):
    A = np.zeros(9)
    orc.orc_euler_angles2matrix(0.0, 0.0, psi, T.dptr(A))
    return A


def test_rotate2d_brute_force(orc):
    rng = np.random.default_rng(11)
    pad, rmax, size = 21, 4, 10
    ref = rng.standard_normal(2 * pad * (pad // 2 + 1))
    v = O.Volume(2, pad, rmax, 2.0, T.dptr(ref), None)
    for psi in rng.uniform(-180, 180, 6):
        A = _euler_2d(orc, psi)
        out = np.zeros((size, size // 2 + 1), complex)
        orc.orc_project(C.byref(v), T.dptr(A), size, out.ctypes.data_as(T.PD))
        Ai = np.linalg.inv(A.reshape(3, 3)) * 2.0
        for i in range(size):
            y = i if i <= size // 2 else i - size
            for x in range(size // 2 + 1):
                if x * x + y * y > min(size // 2, rmax) ** 2:
                    assert out[i, x] == 0
                    continue
                xp, yp = Ai[0, 0] * x + Ai[0, 1] * y, Ai[1, 0] * x + Ai[1, 1] * y
                assert out[i, x] == pytest.approx(_bilinear_numpy(ref, pad, xp, yp), abs=1e-12)


def test_rotate2d_identity_is_central_plane(orc):
    """A 2D class average built as the central plane of a blob volume's transform is, at psi 0,
    the analytic transform of the blobs' z-projection."""
    rng = np.random.default_rng(12)
    blobs = W.gen_blobs(rng, 20, 8.0, (1.5, 3.0), (0.5, 1.0))
    N, pad, rmax = 32, 2 * (2 * 8 + 1) + 1, 8
    xdim = pad // 2 + 1
    data = np.zeros(2 * pad * pad * xdim)
    assert orc.orc_synth_reference(N, 2.0, pad, rmax, 2.0, blobs, 20, 0.0, T.dptr(data)) == 0
    ref = data.reshape(pad, pad * xdim * 2)[pad // 2].copy()
    v = O.Volume(2, pad, rmax, 2.0, T.dptr(ref), None)
    size = 16
    out = np.zeros((size, size // 2 + 1), complex)
    orc.orc_project(C.byref(v), T.dptr(_euler_2d(orc, 0.0)), size, out.ctypes.data_as(T.PD))
    for i in range(size):
        y = i if i <= size // 2 else i - size
        for x in range(size // 2 + 1):
            if x * x + y * y > (size // 2) ** 2:
                continue
            f = np.array([x, y]) / N
            want = sum(b.amplitude * (2 * np.pi * b.sigma ** 2) ** 1.5 *
                       np.exp(-2 * np.pi ** 2 * b.sigma ** 2 * (f @ f)) *
                       np.exp(-2j * np.pi * (f[0] * b.x + f[1] * b.y)) for b in blobs) / N ** 2
            assert out[i, x] == pytest.approx(want, rel=1e-9, abs=1e-12)


def test_backrotate2d_adjoint_and_weight_sum(orc):
    rng = np.random.default_rng(13)
    pad, rmax, size = 25, 5, 10
    ref = rng.standard_normal(2 * pad * (pad // 2 + 1))
    v = O.Volume(2, pad, rmax, 2.0, T.dptr(ref), None)
    for psi in rng.uniform(-180, 180, 5):
        A = _euler_2d(orc, psi)
        f = rng.standard_normal((size, size // 2 + 1)) + 1j * rng.standard_normal((size, size // 2 + 1))
        for i in range(size):
            y = i if i <= size // 2 else i - size
            for x in range(size // 2 + 1):
                if x * x + y * y >= (size // 2) ** 2 or (x == 0 and y <= 0):
                    f[i, x] = 0
        w = np.where(f != 0, rng.uniform(0.5, 2.0, f.shape), 0.0)
        bp, bd, bw = O.make_volume(2, pad, rmax, 2.0, True)
        orc.orc_backproject(C.byref(bp), f.ctypes.data_as(T.PD), T.dptr(np.ascontiguousarray(w)), size, T.dptr(A))
        assert bw.sum() == pytest.approx(w.sum(), rel=1e-12)
        proj = np.zeros((size, size // 2 + 1), complex)
        orc.orc_project(C.byref(v), T.dptr(A), size, proj.ctypes.data_as(T.PD))
        lhs = np.real(np.sum(np.conj(f) * proj))
        rhs = np.real(np.sum(np.conj(bd[0::2] + 1j * bd[1::2]) * (ref[0::2] + 1j * ref[1::2])))
        assert lhs == pytest.approx(rhs, rel=1e-10)


@pytest.mark.parametrize("name", ["class2d64", "class2d128"])
def test_sampling_2d(orc, name):
    lib = T.load()
    cfg, _ = W.make_config(W.WORKLOADS[name])
    oi, li = T.SamplingInfo(), T.SamplingInfo()
    assert orc.orc_sampling_info(C.byref(cfg), C.byref(oi)) == 0
    assert lib.t2x_sampling_info_get(C.byref(cfg), C.byref(li)) == 0
    assert (oi.nr_dir, oi.nr_psi, oi.nr_trans, oi.nr_over_rot, oi.nr_over_trans) == (1, 60, 81, 2, 4)
    assert [getattr(li, f) for f, _ in T.SamplingInfo._fields_] == [getattr(oi, f) for f, _ in T.SamplingInfo._fields_]
    npsi, nt = li.nr_psi, li.nr_trans
    sr, st, sp = np.zeros(1), np.zeros(1), np.zeros(npsi * 2)
    fx, fy = np.zeros(nt * 4), np.zeros(nt * 4)
    assert lib.t2x_sampling_fine_get(C.byref(cfg), T.dptr(sr), T.dptr(st), T.dptr(sp), T.dptr(fx), T.dptr(fy)) == 0
    assert sr[0] == 0 and st[0] == 0
    r, t, p = np.zeros(2), np.zeros(2), np.zeros(2)
    for ps in range(npsi):
        assert orc.orc_sampling_oversampled_orientations(C.byref(cfg), 0, ps, 1, T.dptr(r), T.dptr(t), T.dptr(p)) == 2
        assert np.array_equal(r, [0, 0]) and np.array_equal(t, [0, 0])
        assert np.array_equal(p, sp[2 * ps:2 * ps + 2])
        want = ps * 6.0 + np.array([-1.5, 1.5])
        want = np.where(want > 180, want - 360, want)
        np.testing.assert_allclose(p, want, atol=1e-12)


def test_reference_plane_is_central_plane_of_volume(orc):
    rng = np.random.default_rng(14)
    blobs = W.gen_blobs(rng, 12, 8.0, (1.5, 3.0), (0.5, 1.0))
    N, pad, rmax = 32, 2 * (2 * 8 + 1) + 1, 8
    xdim = pad // 2 + 1
    vol = np.zeros(2 * pad * pad * xdim)
    plane = np.zeros(2 * pad * xdim)
    assert orc.orc_synth_reference(N, 2.0, pad, rmax, 2.0, blobs, 12, 12.0, T.dptr(vol)) == 0
    assert orc.orc_synth_reference_plane(N, 2.0, pad, rmax, 2.0, blobs, 12, 12.0, T.dptr(plane)) == 0
    assert np.array_equal(vol.reshape(pad, -1)[pad // 2], plane)


def _set_class_priors(wl, xy):
    wl.prior_class = np.ascontiguousarray(np.asarray(xy, float).ravel())
    wl.model = T.Model(T.dptr(wl.sigma2), T.dptr(wl.pdf_class), T.dptr(wl.pdf_dir), T.dptr(wl.prior_class))


def test_class_offset_prior_normalisation():
    """The translation prior is normalised by its mean over every (class, translation)
    (ml_optimiser.cpp:7607-7660). Class 1 is excluded by its direction prior (pdf_direction 0) in
    both runs; moving its offset prior 1000 px away makes its translation prior exactly 0, which
    halves that mean: every weight of class 0 doubles (sum_weight x 2) and nothing else changes."""
    import estep_harness as H
    wl = H.Workload("class2d64", n=5, K=2, seed_offset=7)
    wl.pdf_dir[1] = 0.0
    _set_class_priors(wl, [0.0, 0.0, 0.0, 0.0])
    a = wl.run_oracle()
    _set_class_priors(wl, [0.0, 0.0, 1000.0, 0.0])
    b = wl.run_oracle()
    assert np.all(a["out"]["iclass"] == 1) and np.array_equal(a["out"]["best_hidden"], b["out"]["best_hidden"])
    assert np.allclose(b["out"]["sum_weight"] / a["out"]["sum_weight"], 2.0, rtol=1e-12, atol=0)
    for k in ("sigma2_offset", "prior_offset_class", "pdf_class", "sigma2_noise"):
        assert np.allclose(a["wsum"][k], b["wsum"][k], rtol=1e-12, atol=1e-300), k


def test_class_offset_prior_picks_own_centre():
    """sigma2_offset below 1e-4 makes the translation prior an indicator of offset == prior
    (ml_optimiser.cpp:7655-7658): each particle's offset must then sit on the prior of the class
    it is assigned to (within the fine pass's +-1/4-step oversampling), for both classes. The
    particles are made with their true class's prior as offset (an image shifted by s px has
    offset -s px). The priors are compared with old offset + coarse translation, the translation
    in the grid's units (A, ml_optimiser.cpp:7634-7635): q = (-4, 2) is the grid point of -2, +1 px
    at 2 A/px."""
    import estep_harness as H
    wl = H.Workload("class2d64", n=10, K=2, seed_offset=3)
    pix = wl.cfg.pixel_size
    q = np.array([[0.0, 0.0], [-2.0 * pix, 1.0 * pix]])

    def tie(tr):
        tr["x"][:] = -q[tr["cls"], 0] / pix
        tr["y"][:] = -q[tr["cls"], 1] / pix
    wl = H.Workload("class2d64", n=10, K=2, seed_offset=3, truth_hook=tie)
    wl.cfg.sigma2_offset = 0.0
    _set_class_priors(wl, q)
    o = wl.run_oracle()["out"]
    c = o["iclass"] - 1                      # 1-based class numbers
    assert set(c.tolist()) == {0, 1}
    assert np.all(np.abs(o["xoff"] - q[c, 0] / pix) <= 0.5) and np.all(np.abs(o["yoff"] - q[c, 1] / pix) <= 0.5)
