"""Pins of the oracle's optics groups, image-group scale correction and noise-filled mask
background by what the mathematics fixes. CPU only.

- Noise background (ml_optimiser.cpp:6173-6226, mask.cpp:43-97): the noise image is real, its
  transform (1/N^2 forward normalisation) has no component beyond shell N/2, its power per shell
  is 2 fudge sigma2_noise[shell] on average, it scales with sqrt(sigma2), it is reproducible per
  part_id and differs between particles; the masked image equals the noise image beyond
  radius + width, the particle inside the radius, and the raised-cosine blend in between.
- Scale correction (ml_optimiser.cpp:7119-7125, 8326-8350, 8477-8486, 8855-8862): references
  scaled by s with scale_correction 1 and unscaled references with scale_correction s give the
  same diff2, weights, poses and sigma2_noise sums, while XA, AA, the backprojected data and
  weights differ by exactly 1/s, 1/s^2, 1/s and 1/s^2; with one significant sample XA and AA are
  the closed sums of Re(ctf F conj X) and |ctf F|^2 over the shells whose data_vs_prior_class
  exceeds 3, and none qualifies when it is below 3 everywhere.
- Optics groups (ml_optimiser.cpp:6703-6720, 8935-8951): a particle of optics group g is scored
  with row g of sigma2_noise and adds into row g of the sigma2_noise sums and sumw_group[g],
  exactly as a run with that spectrum alone.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
import t2x_py as T
from test_oracle_estep_pins import (N, PIX, HIGH, Exact, add_delta, base_config, centred_grid,
                                    half_freqs, mresol, particle, preprocess, relion_fft,
                                    shells, sigma2_by_shell, to_real, window)


def noise_image(orc, cfg, sigma2, part_id):
    out = np.zeros(N * N)
    assert orc.orc_background_noise(C.byref(cfg), T.dptr(np.ascontiguousarray(sigma2, float)), part_id,
                                    T.dptr(out)) == 0
    return out.reshape(N, N)


def test_noise_background_spectrum(orc):
    cfg = base_config(cs=N)
    cfg.do_zero_mask = 0
    cfg.random_seed = 2024
    cfg.sigma2_fudge = 1.5
    s2 = sigma2_by_shell(0.01)
    m = noise_image(orc, cfg, s2, 17)
    F = np.fft.rfft2(m) / N ** 2          # the half-layout components that generated it
    r = shells(N)
    assert np.all(F[r >= N // 2 + 1] == 0) or np.abs(F[r >= N // 2 + 1]).max() < 1e-15 * np.abs(F).max()
    # real image: an inverse of a Hermitian half spectrum reproduces it exactly
    assert np.abs(np.fft.irfft2(F * N ** 2, s=(N, N)) - m).max() < 1e-12 * np.abs(m).max()
    # power per shell: E|F|^2 = 2 fudge sigma2 (re and im parts each of variance fudge sigma2);
    # averaged over many particles, the shells 4..15 (>= 25 components each) within 10 %
    acc = np.zeros(N // 2 + 1)
    cnt = np.bincount(r[r < N // 2 + 1], minlength=N // 2 + 1)
    ip, jp = half_freqs(N)
    for pid in range(40):
        Fp = np.fft.rfft2(noise_image(orc, cfg, s2, 1000 + pid)) / N ** 2
        ok = r < N // 2 + 1
        acc += np.bincount(r[ok], weights=np.abs(Fp[ok]) ** 2, minlength=N // 2 + 1)
    ratio = acc[4:16] / (40 * cnt[4:16]) / (2 * cfg.sigma2_fudge * s2[4:16])
    assert np.all(np.abs(ratio - 1) < 0.1), ratio
    # the standard deviation scales with sqrt(sigma2), the instance is keyed by part_id
    assert np.abs(noise_image(orc, cfg, 4 * s2, 17) - 2 * m).max() < 1e-12 * np.abs(m).max()
    assert np.array_equal(noise_image(orc, cfg, s2, 17), m)
    assert np.abs(noise_image(orc, cfg, s2, 18) - m).max() > 0.1 * np.abs(m).max()
    # zero spectrum: zero background
    assert np.abs(noise_image(orc, cfg, 0 * s2, 17)).max() == 0.0


def test_preprocess_noise_background(orc):
    """Mask radius 10 px, cosine edge 5 px, a constant particle plus a delta: the masked image
    is the particle inside the radius, the noise image beyond 15 px and the raised-cosine blend
    between; the unmasked transform is untouched."""
    cfg = base_config(cs=N, diameter=2 * 10 * PIX)
    cfg.do_zero_mask = 0
    cfg.random_seed = 7
    s2 = sigma2_by_shell(0.02)
    c, v = 0.375, 3.0
    img = np.full((N, N), c)
    img[N // 2 + 2, N // 2 - 3] = c + v
    p = particle()
    p.part_id = 123
    r = preprocess(orc, cfg, img.astype(np.float32), p, sigma2=s2)
    noise = noise_image(orc, cfg, s2, 123)
    masked = to_real(r["Fimg"])
    x, y = centred_grid()
    rr = np.sqrt(x * x + y * y)
    inside, outside = rr < 10, rr > 15
    edge = ~inside & ~outside
    assert np.abs(masked[inside] - img[inside]).max() < 1e-12
    assert np.abs(masked[outside] - noise[outside]).max() < 1e-12
    rc = 0.5 + 0.5 * np.cos(np.pi * (15 - rr[edge]) / 5)
    assert np.abs(masked[edge] - ((1 - rc) * img[edge] + rc * noise[edge])).max() < 1e-12
    assert np.abs(r["Fnm"] - relion_fft(img)).max() < 1e-12
    # the average-background branch differs from it
    cfg.do_zero_mask = 1
    r0 = preprocess(orc, cfg, img.astype(np.float32), p)
    assert np.abs(to_real(r0["Fimg"])[outside] - noise[outside]).max() > 1e-3


class Grouped(Exact):
    """Exact particles run through the oracle with optics groups, image groups, scale
    corrections and data_vs_prior_class."""

    def run_groups(self, imgs, pins, sigma2_rows, scale=None, dvp=None, vol_scale=1.0):
        orc, cfg, info = self.orc, self.cfg, self.info
        n = len(imgs)
        cfg.nr_optics_groups = len(sigma2_rows)
        cfg.nr_groups = len(scale) if scale is not None else 1
        cfg.do_scale_correction = 1 if scale is not None else 0
        s2 = np.ascontiguousarray(np.asarray(sigma2_rows, float).ravel())
        sc = np.ascontiguousarray(scale, float) if scale is not None else None
        dv = np.ascontiguousarray(np.asarray(dvp, float).ravel()) if dvp is not None else None
        pdfc = np.full(2, 0.5)
        pdfd = np.full(2 * info.nr_dir, 1.0 / info.nr_dir)
        model = T.Model(T.dptr(s2), T.dptr(pdfc), T.dptr(pdfd), None,
                        T.dptr(sc) if sc is not None else None, T.dptr(dv) if dv is not None else None)
        projs, bps, keep, vols = (O.Volume * 2)(), (O.Volume * 2)(), [], []
        for c in range(2):
            vd = np.ascontiguousarray(self.vols[c] * vol_scale)
            vols.append(vd)
            projs[c] = O.Volume(3, self.pad, self.rmax, 2.0, T.dptr(vd), None)
            v, bd, bw = O.make_volume(3, self.pad, self.rmax, 2.0, True)
            bps[c] = v
            keep.append((bd, bw))
        outs = (T.ParticleOut * n)()
        ws, wa = T.make_wsum(2, info.nr_dir, N, cfg.nr_optics_groups, cfg.nr_groups)
        pa = (T.ParticleIn * n)(*pins)
        im = np.ascontiguousarray(np.stack(imgs), dtype=np.float32)
        st = orc.orc_estep(C.byref(cfg), C.byref(model), projs, bps, n, T.fptr(im), pa, outs, C.byref(ws),
                           None, O.ORC_PROJ_CLEAN, 1)
        assert st == 0, st
        bp = [(bd[0::2] + 1j * bd[1::2], bw) for bd, bw in keep]
        return T.outs_to_dict(outs), wa, bp


@pytest.fixture(scope="module")
def single(orc):
    """One pass at the current size, a noise-free particle at a sampled pose with zero
    translation and a flat offset prior: one significant sample of normalised weight 1."""
    cfg = base_config(ao=0)
    cfg.sigma2_offset = 1.0e30
    ex = Grouped(orc, cfg, seed=9)
    c, idir, ipsi, it = 0, 11, 2, 2
    F, A = ex.projection(c, ex.rot[idir], ex.tilt[idir], ex.psi[ipsi])
    G = add_delta(add_delta(F * ex.ctf, *HIGH, 0.5), -3, 4, 0.02 - 0.03j)
    img = ex.image(G, (0.0, 0.0))
    return ex, img, F, (c, idir, ipsi, it)


def test_scale_correction_invariances(single):
    ex, img, F, (c, idir, ipsi, it) = single
    s = 1.7
    s2 = [sigma2_by_shell(2e-6)]
    dvp = np.full((2, N // 2 + 1), 5.0)
    p = particle()
    p.ctf = ex.pin.ctf
    outA, wA, bpA = ex.run_groups([img], [p], s2, scale=[s], dvp=dvp)
    outB, wB, bpB = ex.run_groups([img], [p], s2, scale=[1.0], dvp=dvp, vol_scale=s)
    assert outA["best_hidden"][0] == outB["best_hidden"][0] == ex.hidden(c, idir, ipsi, it)
    for k in ("dLL", "min_diff2", "sum_weight", "pmax"):
        assert outA[k][0] == pytest.approx(outB[k][0], rel=1e-12), k
    assert np.abs(wA["sigma2_noise"] - wB["sigma2_noise"]).max() <= 1e-12 * np.abs(wB["sigma2_noise"]).max()
    assert wA["wsum_signal_product"][0] == pytest.approx(wB["wsum_signal_product"][0] / s, rel=1e-12)
    assert wA["wsum_reference_power"][0] == pytest.approx(wB["wsum_reference_power"][0] / s ** 2, rel=1e-12)
    assert np.linalg.norm(bpA[c][0] - bpB[c][0] * s) < 1e-12 * np.linalg.norm(bpA[c][0])
    assert np.linalg.norm(bpA[c][1] - bpB[c][1] * s ** 2) < 1e-12 * np.linalg.norm(bpA[c][1])
    # with the references unscaled and the image made of them, the true pose's diff2 does not
    # depend on s only when s = 1: the scale enters the reference side of diff2
    outC, _, _ = ex.run_groups([img], [p], s2, scale=[1.0], dvp=dvp)
    assert outC["min_diff2"][0] < outA["min_diff2"][0]


def test_scale_sums_closed_form(single):
    """XA = sum Re(ctf F conj X) and AA = sum |ctf F|^2 over the Mresol pixels of the window whose
    shell has data_vs_prior_class > 3 (here shells below 7), for the one significant sample."""
    ex, img, F, (c, idir, ipsi, it) = single
    cs = ex.cfg.current_size
    s = 0.6
    dvp = np.zeros((2, N // 2 + 1))
    dvp[c, :7] = 4.0
    p = particle()
    p.ctf = ex.pin.ctf
    out, ws, _ = ex.run_groups([img], [p], [sigma2_by_shell(2e-6)], scale=[s], dvp=dvp)
    assert out["pmax"][0] == pytest.approx(1.0, abs=1e-12)
    Fw = window(relion_fft(img.astype(np.float64)), cs)
    ref = window(F * ex.ctf, cs)
    sel = mresol(cs) & (shells(cs) < 7)
    xa = float(np.sum((ref[sel] * np.conj(Fw[sel])).real))
    aa = float(np.sum(np.abs(ref[sel]) ** 2))
    assert ws["wsum_signal_product"][0] == pytest.approx(xa, rel=1e-9)
    assert ws["wsum_reference_power"][0] == pytest.approx(aa, rel=1e-9)
    # below 3 everywhere: no shell qualifies
    _, ws0, _ = ex.run_groups([img], [p], [sigma2_by_shell(2e-6)], scale=[s], dvp=np.full((2, N // 2 + 1), 3.0))
    assert ws0["wsum_signal_product"][0] == 0.0 and ws0["wsum_reference_power"][0] == 0.0


def test_groups_route_spectra_and_sums(single):
    """Two particles, one per optics group (different spectra) and per image group (different
    scales): each equals its own single-group run; the sums land in its group's rows."""
    ex, img, F, _ = single
    sa, sb = sigma2_by_shell(2e-6), sigma2_by_shell(3e-6)
    dvp = np.full((2, N // 2 + 1), 5.0)
    p0, p1 = particle(), particle()
    p0.ctf = p1.ctf = ex.pin.ctf
    p0.optics_group, p0.group_id = 0, 1
    p1.optics_group, p1.group_id = 1, 0
    out, ws, _ = ex.run_groups([img, img], [p0, p1], [sa, sb], scale=[0.9, 1.2], dvp=dvp)
    q0, q1 = particle(), particle()
    q0.ctf = q1.ctf = ex.pin.ctf
    o0, w0, _ = ex.run_groups([img], [q0], [sa], scale=[1.2], dvp=dvp)
    o1, w1, _ = ex.run_groups([img], [q1], [sb], scale=[0.9], dvp=dvp)
    xN = N // 2 + 1
    for k in ("dLL", "min_diff2", "sum_weight", "best_hidden"):
        assert out[k][0] == pytest.approx(o0[k][0], rel=1e-12), k
        assert out[k][1] == pytest.approx(o1[k][0], rel=1e-12), k
    assert out["dLL"][0] != pytest.approx(out["dLL"][1], rel=1e-6)
    assert np.abs(ws["sigma2_noise"][:xN] - w0["sigma2_noise"]).max() <= 1e-12 * w0["sigma2_noise"].max()
    assert np.abs(ws["sigma2_noise"][xN:] - w1["sigma2_noise"]).max() <= 1e-12 * w1["sigma2_noise"].max()
    assert ws["sumw_group"][0] == pytest.approx(w0["sumw_group"][0]) and ws["sumw_group"][1] == pytest.approx(w1["sumw_group"][0])
    assert ws["wsum_signal_product"][1] == pytest.approx(w0["wsum_signal_product"][0], rel=1e-12)
    assert ws["wsum_signal_product"][0] == pytest.approx(w1["wsum_signal_product"][0], rel=1e-12)
    assert ws["wsum_reference_power"][1] == pytest.approx(w0["wsum_reference_power"][0], rel=1e-12)
    assert ws["wsum_reference_power"][0] == pytest.approx(w1["wsum_reference_power"][0], rel=1e-12)
