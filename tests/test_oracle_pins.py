"""Pins the FP64 oracle (oracle/estep_oracle.cpp) against the reference's own known answers,
the reference's HEALPix sources, closed forms and brute force. CPU only."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_py as O
import t2x_py as T

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_ctf_known_answer(orc):
    """tests/ctf.cpp:7-9: CTF::setValues(10000, 12000, 90, 300, 2.7, 0.1, 0, 1, 0); getCTF(10,10)."""
    kat = json.load(open(os.path.join(GOLDEN, "ctf_kat.json")))
    v = kat["setValues"]
    p = T.CtfParams(v["voltage"], v["Cs"], v["Q0"], v["defU"], v["defV"], v["defAng"],
                    v["Bfac"], v["phase_shift"], v["scale"])
    got = orc.orc_ctf_value(C.byref(p), kat["X"], kat["Y"])
    assert got == pytest.approx(kat["expected"], rel=kat["approx_epsilon"])


def test_ctf_closed_forms(orc):
    # at the origin gamma = -K3 so CTF(0,0) = sin(atan(Q0/sqrt(1-Q0^2))) = Q0
    for q0 in (0.07, 0.1, 0.2):
        p = T.CtfParams(300, 2.7, q0, 15000, 15000, 0, 0, 0, 1.0)
        assert orc.orc_ctf_value(C.byref(p), 0.0, 0.0) == pytest.approx(q0, rel=1e-12)
    # astigmatism angle: rotating the defocus axes by a equals evaluating at the rotated point
    p0 = T.CtfParams(300, 2.7, 0.1, 11000, 19000, 0, 0, 0, 1.0)
    a = 37.0
    pa = T.CtfParams(300, 2.7, 0.1, 11000, 19000, a, 0, 0, 1.0)
    rng = np.random.default_rng(1)
    for _ in range(20):
        X, Y = rng.uniform(-0.2, 0.2, 2)
        c, s = np.cos(np.radians(a)), np.sin(np.radians(a))
        Xr, Yr = c * X + s * Y, -s * X + c * Y
        assert orc.orc_ctf_value(C.byref(pa), X, Y) == pytest.approx(
            orc.orc_ctf_value(C.byref(p0), Xr, Yr), rel=1e-9, abs=1e-12)
    # B-factor envelope exp(-B u^2/4)
    pb = T.CtfParams(300, 2.7, 0.1, 15000, 15000, 0, 80, 0, 1.0)
    pn = T.CtfParams(300, 2.7, 0.1, 15000, 15000, 0, 0, 0, 1.0)
    X, Y = 0.05, 0.11
    ratio = orc.orc_ctf_value(C.byref(pb), X, Y) / orc.orc_ctf_value(C.byref(pn), X, Y)
    assert ratio == pytest.approx(np.exp(-80 / 4 * (X * X + Y * Y)), rel=1e-12)


def test_healpix_against_reference_fixture(orc):
    """NEST pixel centres and face coordinates equal the reference Healpix_Base outputs
    (tests/golden/healpix_ref.npz, written from src/Healpix_2.15a by make_healpix_fixture.py)."""
    f = np.load(os.path.join(GOLDEN, "healpix_ref.npz"))
    for k in range(len(f["pix"])):
        z, phi = C.c_double(), C.c_double()
        x, y, fc = C.c_int(), C.c_int(), C.c_int()
        orc.orc_healpix_pix2ang_z_phi(int(f["order"][k]), int(f["pix"][k]), C.byref(z), C.byref(phi))
        orc.orc_healpix_nest2xyf(int(f["order"][k]), int(f["pix"][k]), C.byref(x), C.byref(y), C.byref(fc))
        assert z.value == f["z"][k] and phi.value == f["phi"][k]
        assert (x.value, y.value, fc.value) == (f["ix"][k], f["iy"][k], f["face"][k])
        assert orc.orc_healpix_xyf2nest(int(f["order"][k]), x.value, y.value, fc.value) == f["pix"][k]


def test_healpix_against_reference_build(orc):
    """Same check against the reference sources compiled live (oracle/_ref), orders 0..6."""
    ref = O.load_ref_healpix()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for order in (5, 6):
        for p in range(0, 12 << (2 * order), 7):
            a, b, c, d = C.c_double(), C.c_double(), C.c_double(), C.c_double()
            orc.orc_healpix_pix2ang_z_phi(order, p, C.byref(a), C.byref(b))
            ref.ref_pix2ang_z_phi(order, p, C.byref(c), C.byref(d))
            assert (a.value, b.value) == (c.value, d.value)


def test_sampling_counts(orc):
    """Sample counts of the configs: HEALPix 2 -> 192 dirs x 24 psi, 81 translations for
    range 5 step 1; HEALPix 3 -> 768 x 48; range 6 step 1.5 -> 49."""
    from t2x_py import workloads as W
    for name, nd, npsi, nt in (("refine64", 192, 24, 81), ("class256", 768, 48, 49)):
        cfg, _ = W.make_config(W.WORKLOADS[name])
        info = T.SamplingInfo()
        assert orc.orc_sampling_info(C.byref(cfg), C.byref(info)) == 0
        assert (info.nr_dir, info.nr_psi, info.nr_trans) == (nd, npsi, nt)
        assert (info.nr_over_rot, info.nr_over_trans) == (8, 4)


def test_oversampled_orientations_are_children(orc):
    """Fine directions are the 4 NEST children of the coarse pixel, each with 2 sub-psi at
    +-psi_step/4 (healpix_sampling.cpp:1953-1975, 2023-2050)."""
    from t2x_py import workloads as W
    cfg, _ = W.make_config(W.WORKLOADS["refine64"])
    rot, tilt, psi = np.zeros(8), np.zeros(8), np.zeros(8)
    n = orc.orc_sampling_oversampled_orientations(C.byref(cfg), 37, 5, 1, T.dptr(rot), T.dptr(tilt), T.dptr(psi))
    assert n == 8
    assert np.allclose(psi[0::2], 5 * 15 - 3.75) and np.allclose(psi[1::2], 5 * 15 + 3.75)
    z, phi = C.c_double(), C.c_double()
    for k in range(4):
        orc.orc_healpix_pix2ang_z_phi(3, 37 * 4 + k, C.byref(z), C.byref(phi))
        r, t = np.degrees(phi.value), np.degrees(np.arccos(z.value))
        if r > 180:
            r -= 360
        assert rot[2 * k] == pytest.approx(r) and tilt[2 * k] == pytest.approx(t)


def _euler_closed_form(rot, tilt, psi):
    def rz(a):
        c, s = np.cos(np.radians(a)), np.sin(np.radians(a))
        return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]])

    def ry(b):
        c, s = np.cos(np.radians(b)), np.sin(np.radians(b))
        return np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])
    return rz(psi) @ ry(tilt) @ rz(rot)


def test_euler_matrix(orc):
    rng = np.random.default_rng(2)
    for _ in range(50):
        r, t, p = rng.uniform(-180, 180), rng.uniform(0, 180), rng.uniform(-180, 180)
        A = np.zeros(9)
        orc.orc_euler_angles2matrix(r, t, p, T.dptr(A))
        A = A.reshape(3, 3)
        assert np.allclose(A, _euler_closed_form(r, t, p), atol=1e-14)
        assert np.allclose(A @ A.T, np.eye(3), atol=1e-14)


@pytest.mark.parametrize("n", [16, 36, 64, 90])
def test_fft_normalisation(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n))
    out = np.zeros((n, n // 2 + 1), complex)
    orc.orc_fft2_r2c(n, T.dptr(x), out.ctypes.data_as(T.PD))
    assert np.allclose(out, np.fft.rfft2(x) / n ** 2, atol=1e-14)
    back = np.zeros((n, n))
    orc.orc_fft2_c2r(n, out.ctypes.data_as(T.PD), T.dptr(back))
    assert np.allclose(back, x, atol=1e-12)


def _random_volume(pad, rng):
    xdim = pad // 2 + 1
    v = rng.standard_normal(2 * pad * pad * xdim)
    return v


def _trilinear_numpy(vol, pad, xp, yp, zp):
    xdim, h = pad // 2 + 1, pad // 2
    V = (vol[0::2] + 1j * vol[1::2]).reshape(pad, pad, xdim)
    neg = xp < 0
    if neg:
        xp, yp, zp = -xp, -yp, -zp
    x0, y0, z0 = int(np.floor(xp)), int(np.floor(yp)), int(np.floor(zp))
    fx, fy, fz = xp - x0, yp - y0, zp - z0
    y0 += h
    z0 += h
    if x0 < 0 or x0 + 1 >= xdim or y0 < 0 or y0 + 1 >= pad or z0 < 0 or z0 + 1 >= pad:
        return 0
    val = 0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                val += w * V[z0 + dz, y0 + dy, x0 + dx]
    return np.conj(val) if neg else val


def test_project_brute_force(orc):
    """Projector::project against an independent trilinear sampler on a random volume."""
    rng = np.random.default_rng(3)
    pad, rmax, size = 21, 4, 10
    vol = _random_volume(pad, rng)
    v = O.Volume(3, pad, rmax, 2.0, T.dptr(vol), None)
    for _ in range(5):
        A = np.zeros(9)
        orc.orc_euler_angles2matrix(*rng.uniform(-180, 180, 3), T.dptr(A))
        out = np.zeros((size, size // 2 + 1), complex)
        orc.orc_project(C.byref(v), T.dptr(A), size, out.ctypes.data_as(T.PD))
        Ai = A.reshape(3, 3).T * 2.0
        for i in range(size):
            y = i if i <= size // 2 else i - size
            for x in range(size // 2 + 1):
                # r_max_out = size/2 (x loop bound) and the r_max sphere of the reference
                if x * x + y * y > min(size // 2, rmax) ** 2:
                    assert out[i, x] == 0
                    continue
                xp, yp, zp = Ai @ np.array([x, y, 0.0])
                assert out[i, x] == pytest.approx(_trilinear_numpy(vol, pad, xp, yp, zp), abs=1e-12)


def test_project_identity_is_central_section(orc):
    """At the identity orientation the slice samples grid points exactly: it equals the z=0
    plane of the analytic blob transform."""
    from t2x_py import workloads as W
    rng = np.random.default_rng(4)
    blobs = W.gen_blobs(rng, 20, 8.0, (1.5, 3.0), (0.5, 1.0))
    N, pad, rmax = 32, 2 * (2 * 8 + 1) + 1, 8
    data = np.zeros(2 * pad * pad * (pad // 2 + 1))
    assert orc.orc_synth_reference(N, 2.0, pad, rmax, 2.0, blobs, 20, 0.0, T.dptr(data)) == 0
    v = O.Volume(3, pad, rmax, 2.0, T.dptr(data), None)
    A = np.eye(3).ravel().copy()
    size = 16
    out = np.zeros((size, size // 2 + 1), complex)
    orc.orc_project(C.byref(v), T.dptr(A), size, out.ctypes.data_as(T.PD))
    for i in range(size):
        y = i if i <= size // 2 else i - size
        for x in range(size // 2 + 1):
            if x * x + y * y > (size // 2) ** 2:
                continue
            f = np.array([x, y, 0.0]) / N
            ref = sum(b.amplitude * (2 * np.pi * b.sigma ** 2) ** 1.5 *
                      np.exp(-2 * np.pi ** 2 * b.sigma ** 2 * (f @ f)) *
                      np.exp(-2j * np.pi * (f[0] * b.x + f[1] * b.y)) for b in blobs) / N ** 2
            assert out[i, x] == pytest.approx(ref, rel=1e-9, abs=1e-12)


def test_backproject_adjoint_and_weight_sum(orc):
    """Trilinear splatting is the adjoint of trilinear sampling, and the weight volume gains
    exactly the sum of the pixel weights (the 8 trilinear weights sum to 1)."""
    rng = np.random.default_rng(5)
    pad, rmax, size = 25, 5, 10
    xdim = pad // 2 + 1
    vol = _random_volume(pad, rng)
    v = O.Volume(3, pad, rmax, 2.0, T.dptr(vol), None)
    for _ in range(4):
        A = np.zeros(9)
        orc.orc_euler_angles2matrix(*rng.uniform(-180, 180, 3), T.dptr(A))
        f = rng.standard_normal((size, size // 2 + 1)) + 1j * rng.standard_normal((size, size // 2 + 1))
        # keep pixels strictly inside the disc and off the duplicated x=0, y<0 half column
        for i in range(size):
            y = i if i <= size // 2 else i - size
            for x in range(size // 2 + 1):
                if x * x + y * y >= (size // 2) ** 2 or (x == 0 and y < 0) or (x == 0 and y == 0):
                    f[i, x] = 0
        w = np.where(f != 0, rng.uniform(0.5, 2.0, f.shape), 0.0)
        bp, bd, bw = O.make_volume(3, pad, rmax, 2.0, True)
        orc.orc_backproject(C.byref(bp), f.ctypes.data_as(T.PD), T.dptr(np.ascontiguousarray(w)), size, T.dptr(A))
        assert bw.sum() == pytest.approx(w.sum(), rel=1e-12)
        proj = np.zeros((size, size // 2 + 1), complex)
        orc.orc_project(C.byref(v), T.dptr(A), size, proj.ctypes.data_as(T.PD))
        lhs = np.real(np.sum(np.conj(f) * proj))
        V = vol[0::2] + 1j * vol[1::2]
        B = bd[0::2] + 1j * bd[1::2]
        rhs = np.real(np.sum(np.conj(B) * V))
        assert lhs == pytest.approx(rhs, rel=1e-10)
        assert xdim > 0


def _threshold_numpy(w, frac, maxsig, count):
    pos = np.sort(w[w > 0])
    s = w.sum()
    acc, sig, n = 0.0, 0.0, 0
    for v in pos[::-1]:
        if maxsig > 0:
            if n < maxsig:
                if count:
                    n += 1
                sig = v
        else:
            if count:
                n += 1
            sig = v
        acc += v
        if acc > frac * s:
            break
    return s, sig, n


@pytest.mark.parametrize("seed,maxsig", [(0, 0), (1, 0), (2, 5), (3, 0)])
def test_weights_and_threshold(orc, seed, maxsig):
    """convertAllSquaredDifferencesToWeights' sorted walk, against numpy on random diff2 rows
    with ties, invalid (-999) samples and samples beyond the 700 cut-off."""
    rng = np.random.default_rng(seed)
    n = 5000
    d2 = rng.uniform(100, 130, n)
    d2[rng.integers(0, n, 50)] = -999.0
    d2[rng.integers(0, n, 50)] = 100 + 800
    d2[10:20] = d2[5]                      # ties
    prior = rng.uniform(0.5, 1.5, n)
    prior[rng.integers(0, n, 20)] = 0.0
    mind = d2[d2 >= 0].min()
    w = np.zeros(n)
    s, sig, cnt = np.zeros(1), np.zeros(1), C.c_long()
    orc.orc_weights_from_diff2(n, T.dptr(d2), T.dptr(prior), mind, 0.999, maxsig, 1, T.dptr(w),
                               T.dptr(s), T.dptr(sig), C.byref(cnt))
    ref_w = np.where(d2 < 0, 0.0, np.where(d2 - mind > 700, 0.0, prior * np.exp(-(d2 - mind))))
    assert np.allclose(w, ref_w, rtol=1e-15, atol=0)
    # the walk itself, on the oracle's own weights (exp may differ from numpy's in the last ulp)
    rs, rsig, rn = _threshold_numpy(w, 0.999, maxsig, True)
    assert s[0] == pytest.approx(rs, rel=1e-12)
    assert sig[0] == rsig and cnt.value == rn


def test_stale_projection_buffer_is_immaterial():
    """The CPU diff2 loop reuses its projection buffer (ml_optimiser.cpp:7017-7065): pixels that
    the floating-point r_max test skips keep the previous orientation's values. The oracle's
    ORC_PROJ_STALE mode reproduces that; on the refine64 workload it changes nothing beyond
    rounding, so the intended semantics (zero outside r_max) used by libt2x match the reference."""
    import estep_harness as H
    wl = H.Workload("refine64", n=3)
    a = wl.run_oracle(n_debug=3, fine_cap=8192, proj_mode=O.ORC_PROJ_CLEAN)
    b = wl.run_oracle(n_debug=3, fine_cap=8192, proj_mode=O.ORC_PROJ_STALE)
    v = a["dbg"]["coarse_diff2"] > -998
    assert H.rel(b["dbg"]["coarse_diff2"][v], a["dbg"]["coarse_diff2"][v]) < 1e-12
    assert np.array_equal(a["out"]["best_hidden"], b["out"]["best_hidden"])
    assert H.rel_norm(b["bp"][0][1], a["bp"][0][1]) < 1e-12
