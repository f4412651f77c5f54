"""Pins of the reconstruction oracle (BackProjector::reconstruct, skip_gridding branch; getFSC).

Every case has a closed form: a single Fourier component of the backprojected data turns into a
known plane wave in the reconstructed map, divided by the FFT normalisation pf^3 * ori (2D data)
and by the trilinear gridding correction sinc^2(r / (ori * pf)). The cases pin the axis order
([z][y][x]), the FFTW inverse-transform sign, CenterFFTbySign with the Xmipp origin, the factor 2
of half-complex x > 0 components, the real-part semantics of the x = 0 plane, the MAP term, the
radially averaged weight floor and both windowing branches. The FSC is pinned against numpy's
rfftn and against the invariances FSC(m, m) = 1, FSC(m, -m) = -1, FSC(m, a*m) = 1.
"""
import numpy as np
import pytest

import t2x_py as T
import oracle_py as O


def sinc2(n, pf):
    c = np.arange(n) - n // 2
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(z * z + y * y + x * x)
    rv = r / (n * pf)
    s = np.ones_like(rv)
    nz = r > 0
    s[nz] = np.sin(np.pi * rv[nz]) / (np.pi * rv[nz])
    return s * s, r, (z, y, x)


def volumes(p):
    pad = p.pad_size
    xh = pad // 2 + 1
    return np.zeros((pad, pad, xh)), np.zeros((pad, pad, xh)), np.zeros((pad, pad, xh))


def setv(arr, p, kp, ip, jp, v):
    c0 = p.pad_size // 2
    arr[kp + c0, ip + c0, jp] = v


def recon(p, re, im, w, tau2=None):
    L = O.load()
    out = np.zeros((p.ori_size,) * 3)
    t = np.ascontiguousarray(tau2, dtype=np.float64) if tau2 is not None else None
    st = L.orc_reconstruct(p, T.dptr(np.ascontiguousarray(re)), T.dptr(np.ascontiguousarray(im)),
                           T.dptr(np.ascontiguousarray(w)), T.dptr(t), T.dptr(out))
    assert st == 0, st
    return out


@pytest.mark.parametrize("r_max", [6, 8])   # zero-padding window (6) and cropping window (8)
def test_constant_component(r_max):
    p = T.recon_params(16, r_max)
    re, im, w = volumes(p)
    setv(re, p, 0, 0, 0, 3.0)
    setv(w, p, 0, 0, 0, 1.0)
    out = recon(p, re, im, w)
    g, _, _ = sinc2(16, 2.0)
    np.testing.assert_allclose(out, 3.0 / (8 * 16) / g, rtol=1e-12)


@pytest.mark.parametrize("r_max", [6, 8])
def test_plane_waves_axis_order_and_sign(r_max):
    p = T.recon_params(16, r_max)
    P = 32
    g, r, (zl, yl, xl) = sinc2(16, 2.0)
    inside = r < 8
    c = 1.0 + 2.0j
    for (kp, ip, jp) in [(0, 0, 3), (2, -1, 4), (-3, 2, 1)]:
        re, im, w = volumes(p)
        setv(re, p, kp, ip, jp, c.real)
        setv(im, p, kp, ip, jp, c.imag)
        setv(w, p, kp, ip, jp, 1.0)
        out = recon(p, re, im, w)
        phase = np.exp(2j * np.pi * (kp * zl + ip * yl + jp * xl) / P)
        expect = 2 * np.real(c * phase) / (8 * 16) / g
        np.testing.assert_allclose(out[inside], expect[inside], rtol=1e-10, atol=1e-14)


def test_x0_plane_enters_through_real_part():
    """A lone component on the x = 0 plane (no Friedel mate): half of it plus its mirror."""
    p = T.recon_params(16, 6)
    P = 32
    g, r, (zl, yl, xl) = sinc2(16, 2.0)
    inside = r < 8
    c = 0.5 - 1.5j
    re, im, w = volumes(p)
    setv(re, p, 2, -3, 0, c.real)
    setv(im, p, 2, -3, 0, c.imag)
    setv(w, p, 2, -3, 0, 1.0)
    out = recon(p, re, im, w)
    expect = np.real(c * np.exp(2j * np.pi * (2 * zl - 3 * yl) / P)) / (8 * 16) / g
    np.testing.assert_allclose(out[inside], expect[inside], rtol=1e-10, atol=1e-14)


def test_map_term():
    p = T.recon_params(16, 6, do_map=True, tau2_fudge=2.0)
    g, _, _ = sinc2(16, 2.0)
    tau = 0.25
    re, im, w = volumes(p)
    setv(re, p, 0, 0, 0, 3.0)
    setv(w, p, 0, 0, 0, 5.0)
    out = recon(p, re, im, w, tau2=np.full(9, tau))
    invtau2 = 1.0 / (8.0 * 2.0 * tau)
    np.testing.assert_allclose(out, 3.0 / (5.0 + invtau2) / 128 / g, rtol=1e-12)
    # tau2 == 0 at the origin shell: 1 / (0.001 * w)
    t0 = np.full(9, tau)
    t0[0] = 0.0
    out = recon(p, re, im, w, tau2=t0)
    np.testing.assert_allclose(out, 3.0 / (5.0 + 1000.0 / 5.0) / 128 / g, rtol=1e-12)
    # minres_map = 1 excludes the origin shell
    p1 = T.recon_params(16, 6, do_map=True, tau2_fudge=2.0, minres_map=1)
    out = recon(p1, re, im, w, tau2=np.full(9, tau))
    np.testing.assert_allclose(out, 3.0 / 5.0 / 128 / g, rtol=1e-12)


def test_radially_averaged_weight_floor():
    """A component with zero weight is divided by 1/1000 of its shell's average weight."""
    p = T.recon_params(16, 6)
    P, pf = 32, 2.0
    g, r, (zl, yl, xl) = sinc2(16, pf)
    inside = r < 8
    kq = (1, 2, 3)
    s = int(np.floor(np.sqrt(sum(k * k for k in kq)) / pf))
    re, im, w = volumes(p)
    pad, xh, c0 = p.pad_size, p.pad_size // 2 + 1, p.pad_size // 2
    round_max_r2 = int(np.floor(6 * pf * 6 * pf + 0.5))
    count = 0
    for kp in range(-c0, c0 + 1):
        for ip in range(-c0, c0 + 1):
            for jp in range(xh):
                r2 = kp * kp + ip * ip + jp * jp
                if r2 < round_max_r2 and int(np.floor(np.sqrt(r2) / pf)) == s:
                    count += 1
                    if (kp, ip, jp) != kq:
                        setv(w, p, kp, ip, jp, 1.0)
    setv(re, p, *kq, 2.0)
    out = recon(p, re, im, w)
    floor = (count - 1) / (1000.0 * count)
    expect = 2 * 2.0 / floor * np.cos(2 * np.pi * (kq[0] * zl + kq[1] * yl + kq[2] * xl) / P) / 128 / g
    np.testing.assert_allclose(out[inside], expect[inside], rtol=1e-10, atol=1e-11)


def np_fsc(m1, m2):
    n = m1.shape[0]
    a, b = np.fft.rfftn(m1), np.fft.rfftn(m2)
    k = np.fft.fftfreq(n) * n
    kx = np.arange(n // 2 + 1)
    kz, ky, kxx = np.meshgrid(k, k, kx, indexing="ij")
    idx = np.floor(np.sqrt(kz ** 2 + ky ** 2 + kxx ** 2) + 0.5).astype(int)
    xh = n // 2 + 1
    sel = idx < xh
    num = np.bincount(idx[sel], (np.conj(a) * b).real[sel], xh)
    d1 = np.bincount(idx[sel], np.abs(a[sel]) ** 2, xh)
    d2 = np.bincount(idx[sel], np.abs(b[sel]) ** 2, xh)
    return num / np.sqrt(d1 * d2)


def fsc(m1, m2):
    L = O.load()
    n = m1.shape[0]
    out = np.zeros(n // 2 + 1)
    assert L.orc_fsc(n, T.dptr(np.ascontiguousarray(m1)), T.dptr(np.ascontiguousarray(m2)), T.dptr(out)) == 0
    return out


@pytest.mark.parametrize("n", [12, 16])
def test_fsc_against_numpy_and_invariants(n):
    rng = np.random.default_rng(5)
    m1 = rng.standard_normal((n, n, n))
    m2 = m1 + 0.7 * rng.standard_normal((n, n, n))
    np.testing.assert_allclose(fsc(m1, m2), np_fsc(m1, m2), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(fsc(m1, m1), 1.0, rtol=1e-12)
    np.testing.assert_allclose(fsc(m1, -m1), -1.0, rtol=1e-12)
    np.testing.assert_allclose(fsc(m1, 3.5 * m1), 1.0, rtol=1e-12)
