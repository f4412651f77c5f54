"""Pins of the oracle's preprocessing, diff2 assembly and weighted sums by closed forms. CPU only.

Every expectation here is computed from the mathematics of the step (numpy transforms of the
float32 image the oracle reads, plane waves, deltas, the projector's disc), not from the oracle's
code:

- preprocessing (getFourierTransformsAndCtfs, ml_optimiser.cpp:5689-6250): a plane wave
  A cos(2 pi k.x / N + phi) on the centred pixel grid has the single half-plane coefficient
  A/2 e^{i phi} (1/N^2 forward normalisation, CenterFFTbySign, windowing by frequency); a delta
  moved by the rounded old offset and scaled by avg_norm / norm has the plane-wave transform of
  its new position; the average-background mask keeps the inside, replaces the outside by the
  mask-weighted mean and preserves the image sum.
- diff2 (getAllSquaredDifferences, ml_optimiser.cpp:7265-7273): a particle that equals ctf times
  the projection at a sampled pose, displaced so that the sampled translation re-centres it,
  scores exactly highres_Xi2/2 plus the Mresol pixels outside the projector's disc
  (projector.cpp:535-560); a change delta at one pixel adds 1/2 Minvsigma2 |delta|^2 inside the
  pass's window, nothing at the origin, |delta|^2/2 beyond the current size (highres_Xi2).
- weighted sums (storeWeightedSums, ml_optimiser.cpp:8400-8956) with a single significant sample
  (w = 1): sigma2_noise shell sums of |ctf ref - X|^2 (origin included) plus the power beyond the
  current size, dLL = log(sum_weight) - min_diff2 - sum log(2 pi sigma2), the norm update
  old / avg * sqrt(2 wsum_norm), the pose / offset sums, and the backprojection of the unmasked
  image (a mask that changes the particle must not change what is backprojected).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
import t2x_py as T
from t2x_py import workloads as W

N = 32
PIX = 4.0


# ---------------------------------------------------------------------------------------------
# helpers: the transform conventions pinned by test_preprocess_plane_waves
# ---------------------------------------------------------------------------------------------

def half_freqs(s):
    xs = s // 2 + 1
    i = np.arange(s)
    ip = np.where(i < xs, i, i - s)[:, None]
    jp = np.arange(xs)[None, :]
    return ip, jp


def shells(s):
    ip, jp = half_freqs(s)
    return np.rint(np.sqrt(ip ** 2 + jp ** 2)).astype(int)


def mresol(s):
    """Pixels of the size-s half transform with a resolution shell (ires < s/2+1, x=0 column
    without its Hermitian duplicates), ml_optimiser.cpp:5596-5625."""
    ip, jp = half_freqs(s)
    r = shells(s)
    return (r < s // 2 + 1) & ~((jp == 0) & (ip < 0))


def in_disc(s):
    """Pixels the projector fills at output size s: x^2 + y^2 <= (s/2)^2 (projector.cpp:535-560)."""
    ip, jp = half_freqs(s)
    return ip ** 2 + jp ** 2 <= (s // 2) ** 2


def centre_sign(s):
    ip, jp = half_freqs(s)
    i = np.arange(s)[:, None]
    return np.where((i + jp) % 2, -1.0, 1.0)


def relion_fft(img):
    """Forward transform normalised by 1/N^2 with the origin at the image centre."""
    n = img.shape[0]
    return np.fft.rfft2(img) / n ** 2 * centre_sign(n)


def to_real(F):
    n = F.shape[0]
    return np.fft.irfft2(F * centre_sign(n), s=(n, n)) * n * n


def window(F, s):
    """The size-s window of a size-N half transform (same logical frequencies)."""
    ip, _ = half_freqs(s)
    rows = np.where(ip[:, 0] < 0, ip[:, 0] + F.shape[0], ip[:, 0])
    return F[rows][:, : s // 2 + 1]


def base_config(cs=24, cc=16, ao=1, diameter=1.0e5):
    cfg = T.EstepConfig()
    cfg.ori_size = N
    cfg.pixel_size = PIX
    cfg.current_size = cs
    cfg.coarse_size = cc
    cfg.nr_classes = 2
    cfg.ref_dim = 3
    cfg.healpix_order = 1
    cfg.psi_step = 30.0
    cfg.offset_range = 4 * PIX
    cfg.offset_step = 4 * PIX          # translations (0,0), (+-4, 0), (0, +-4) px; fine ones +-1 px
    cfg.adaptive_oversampling = ao
    cfg.adaptive_fraction = 0.999
    cfg.particle_diameter = diameter
    cfg.width_mask_edge = 5
    cfg.do_zero_mask = 1
    cfg.sigma2_fudge = 1.0
    cfg.avg_norm_correction = 1.0
    cfg.do_ctf_correction = 1
    cfg.sigma2_offset = 1.0e12
    return cfg


def preprocess(orc, cfg, img, pin, sigma2=None):
    cs = cfg.current_size
    nc = cs * (cs // 2 + 1)
    Fi, Fn, Fc = np.zeros(2 * nc), np.zeros(2 * nc), np.zeros(nc)
    power, hres, old = np.zeros(N // 2 + 1), C.c_double(), np.zeros(2)
    im = np.ascontiguousarray(img, dtype=np.float32)
    s2 = T.dptr(np.ascontiguousarray(sigma2, dtype=np.float64)) if sigma2 is not None else None
    assert orc.orc_preprocess(C.byref(cfg), T.fptr(im), C.byref(pin), s2, T.dptr(Fi), T.dptr(Fn), T.dptr(Fc),
                              T.dptr(power), C.byref(hres), T.dptr(old)) == 0
    shape = (cs, cs // 2 + 1)
    return dict(Fimg=(Fi[0::2] + 1j * Fi[1::2]).reshape(shape), Fnm=(Fn[0::2] + 1j * Fn[1::2]).reshape(shape),
                Fctf=Fc.reshape(shape), power=power, highres=hres.value, old=old)


def particle(ctf=None, norm=1.0):
    p = T.ParticleIn()
    p.ctf = ctf or T.CtfParams(300.0, 2.7, 0.1, 15000.0, 17000.0, 30.0, 0.0, 0.0, 1.0)
    p.prior_offset_x = p.prior_offset_y = 999.0
    p.norm_correction = norm
    p.prior_rot = p.prior_tilt = p.prior_psi = 999.0
    return p


def centred_grid():
    y, x = np.mgrid[0:N, 0:N] - N // 2
    return x.astype(float), y.astype(float)


# ---------------------------------------------------------------------------------------------
# preprocessing
# ---------------------------------------------------------------------------------------------

def test_preprocess_plane_waves(orc):
    """Two plane waves, one inside the current size and one beyond it; no mask in reach."""
    cfg = base_config(cs=24)
    x, y = centred_grid()
    (k1x, k1y, a1, p1) = (3, -5, 0.7, 0.4)      # shell 6: in the window
    (k2x, k2y, a2, p2) = (13, 2, 0.3, -1.1)     # shell 13 >= cs/2 + 1: beyond it
    img = a1 * np.cos(2 * np.pi * (k1x * x + k1y * y) / N + p1) + a2 * np.cos(2 * np.pi * (k2x * x + k2y * y) / N + p2)
    r = preprocess(orc, cfg, img.astype(np.float32), particle())
    exp = np.zeros((24, 13), complex)
    exp[k1y % 24, k1x] = a1 / 2 * np.exp(1j * p1)
    assert np.abs(r["Fimg"] - exp).max() < 2e-8
    assert np.abs(r["Fnm"] - exp).max() < 2e-8
    # power spectrum per shell and high-resolution Xi2 of the masked image
    pw = np.zeros(N // 2 + 1)
    pw[6] = a1 ** 2 / 4
    pw[13] = a2 ** 2 / 4
    assert np.abs(r["power"] - pw).max() < 1e-8
    assert r["highres"] == pytest.approx(a2 ** 2 / 4, rel=1e-6)
    # the CTF image at logical frequency (ip, jp) of the N-pixel box: CTF(jp/(N angpix), ip/(N angpix))
    p = particle()
    ip, jp = half_freqs(24)
    for (i, j) in ((0, 0), (3, 7), (20, 11), (12, 12), (23, 1)):
        want = orc.orc_ctf_value(C.byref(p.ctf), jp[0, j] / (N * PIX), ip[i, 0] / (N * PIX))
        assert r["Fctf"][i, j] == pytest.approx(want, rel=1e-12, abs=1e-15)


def test_preprocess_sine_phase_and_full_window(orc):
    """A sine pins the sign of the forward transform's exponent; at current size = N the power
    spectrum is not computed (image_current_size == image_full_size) and highres_Xi2 is zero."""
    cfg = base_config(cs=N, cc=16)
    x, y = centred_grid()
    img = 0.5 * np.sin(2 * np.pi * (4 * x + 9 * y) / N)
    r = preprocess(orc, cfg, img.astype(np.float32), particle())
    assert abs(r["Fimg"][9, 4] - (-0.25j)) < 2e-8
    assert r["highres"] == 0.0 and np.all(r["power"] == 0.0)


@pytest.mark.parametrize("old,rounded", [((1.6, -1.4), (2, -1)), ((-2.5, 2.5), (-3, 3))])
def test_preprocess_recentre_and_norm(orc, old, rounded):
    """selfTranslate by the rounded old offset (DONT_WRAP) and the norm-correction scaling."""
    cfg = base_config(cs=N)
    cfg.do_norm_correction = 1
    cfg.avg_norm_correction = 1.5
    p = particle(norm=0.75)                     # image scaled by avg / norm = 2
    p.old_offset_x, p.old_offset_y = old
    x0, y0, v = 3, -6, 1.25
    img = np.zeros((N, N))
    img[y0 + N // 2, x0 + N // 2] = v
    r = preprocess(orc, cfg, img, p)
    assert tuple(r["old"]) == rounded
    ip, jp = half_freqs(N)
    xn, yn = x0 + rounded[0], y0 + rounded[1]
    exp = 2 * v / N ** 2 * np.exp(-2j * np.pi * (jp * xn + ip * yn) / N)
    assert np.abs(r["Fnm"] - exp).max() < 1e-12
    # a pixel moved across the border is dropped, not wrapped
    img2 = np.zeros((N, N))
    img2[0, N - 1] = v                          # logical (x, y) = (N/2 - 1, -N/2)
    p2 = particle()
    p2.old_offset_x = 2.0
    assert np.abs(preprocess(orc, base_config(cs=N), img2, p2)["Fnm"]).max() == 0.0


def test_preprocess_average_background_mask(orc):
    """A constant c plus a delta v in a corner, mask radius 10 px with a 5 px cosine edge: the
    masked image is c inside the radius, c + v/S beyond radius + width and c + m(r) v/S in the
    edge, with m the raised cosine 0 -> 1 and S = sum of m; the unmasked transform keeps the delta."""
    cfg = base_config(cs=N, diameter=2 * 10 * PIX)
    c, v = 0.375, 3.0                           # exact in float32
    img = np.full((N, N), c)
    img[0, 0] = c + v                           # logical (-16, -16), r = 22.6
    r = preprocess(orc, cfg, img, particle())
    ip, jp = half_freqs(N)
    expn = v / N ** 2 * np.exp(-2j * np.pi * (jp * -16 + ip * -16) / N)
    expn[0, 0] += c
    assert np.abs(r["Fnm"] - expn).max() < 1e-12
    masked = to_real(r["Fimg"])
    x, y = centred_grid()
    rr = np.sqrt(x * x + y * y)
    inside, outside = rr < 10, rr > 15
    edge = ~inside & ~outside
    assert np.abs(masked[inside] - c).max() < 1e-12
    bg = masked[outside] - c
    assert np.ptp(bg) < 1e-12
    m = (masked - c) / bg[0]
    assert np.abs(m[edge] - (0.5 + 0.5 * np.cos(np.pi * (15 - rr[edge]) / 5))).max() < 1e-9
    S = m.sum()
    assert bg[0] == pytest.approx(v / S, rel=1e-10)
    # the average-background fill preserves the image sum (DC term)
    assert r["Fimg"][0, 0].real == pytest.approx(c + v / N ** 2, rel=1e-12)


# ---------------------------------------------------------------------------------------------
# diff2 and weighted sums on exact particles
# ---------------------------------------------------------------------------------------------

class Exact:
    """Two blob references (r_max 15, so the projection reaches beyond the current size), the
    sampling of base_config, and particles equal to ctf times a projection at a sampled pose."""

    def __init__(self, orc, cfg, seed=5):
        self.orc, self.cfg = orc, cfg
        info = T.SamplingInfo()
        assert orc.orc_sampling_info(C.byref(cfg), C.byref(info)) == 0
        self.info = info
        self.rmax = 15
        self.pad = 2 * (int(2 * self.rmax + 0.5) + 1) + 1
        rng = np.random.default_rng(seed)
        self.vols = []
        for _ in range(2):
            bl = W.gen_blobs(rng, 40, 10.0, (1.5, 3.0), (0.5, 1.0))
            d = np.zeros(2 * self.pad * self.pad * (self.pad // 2 + 1))
            assert orc.orc_synth_reference(N, PIX, self.pad, self.rmax, 2.0, bl, 40, 0.0, T.dptr(d)) == 0
            self.vols.append(d)
        nd, npsi, nt = info.nr_dir, info.nr_psi, info.nr_trans
        self.rot, self.tilt, self.psi = np.zeros(nd), np.zeros(nd), np.zeros(npsi)
        self.tx, self.ty = np.zeros(nt), np.zeros(nt)
        orc.orc_sampling_coarse(C.byref(cfg), T.dptr(self.rot), T.dptr(self.tilt), T.dptr(self.psi),
                                T.dptr(self.tx), T.dptr(self.ty))
        self.pin = particle()
        self.ctf = np.zeros(N * (N // 2 + 1))
        orc.orc_ctf_image(C.byref(self.pin.ctf), N, PIX, N, T.dptr(self.ctf))
        self.ctf = self.ctf.reshape(N, N // 2 + 1)

    def projection(self, c, rot, tilt, psi):
        A = np.zeros(9)
        self.orc.orc_euler_angles2matrix(rot, tilt, psi, T.dptr(A))
        v = O.Volume(3, self.pad, self.rmax, 2.0, T.dptr(self.vols[c]), None)
        F = np.zeros(2 * N * (N // 2 + 1))
        self.orc.orc_project(C.byref(v), T.dptr(A), N, T.dptr(F))
        return (F[0::2] + 1j * F[1::2]).reshape(N, N // 2 + 1), A

    def image(self, F, shift_px):
        """The real image of transform F displaced by -shift (a circular roll by whole pixels),
        so that translating it by shift re-centres it; as float32 (what the E-step reads)."""
        img = to_real(F)
        img = np.roll(img, shift=(-int(round(shift_px[1])), -int(round(shift_px[0]))), axis=(0, 1))
        return img.astype(np.float32)

    def run(self, img, sigma2, n_debug=1, fine_cap=1 << 17, pin=None):
        orc, cfg, info = self.orc, self.cfg, self.info
        pdfc = np.full(2, 0.5)
        pdfd = np.full(2 * info.nr_dir, 1.0 / info.nr_dir)
        model = T.Model(T.dptr(sigma2), T.dptr(pdfc), T.dptr(pdfd), None)
        projs, bps, keep = (O.Volume * 2)(), (O.Volume * 2)(), []
        for c in range(2):
            projs[c] = O.Volume(3, self.pad, self.rmax, 2.0, T.dptr(self.vols[c]), None)
            v, bd, bw = O.make_volume(3, self.pad, self.rmax, 2.0, True)
            bps[c] = v
            keep.append((bd, bw))
        outs = (T.ParticleOut * 1)()
        ws, wa = T.make_wsum(2, info.nr_dir, N)
        nh = info.nr_hidden_coarse
        arr = dict(coarse_diff2=np.zeros((1, nh), np.float32), fine_hidden=np.zeros((1, fine_cap), np.int64),
                   fine_diff2=np.zeros((1, fine_cap), np.float32), fine_count=np.zeros(1, np.int64))
        d = T.EstepDebug()
        d.n_debug = n_debug
        d.coarse_diff2 = arr["coarse_diff2"].ctypes.data_as(C.POINTER(C.c_float))
        d.fine_capacity = fine_cap
        d.fine_hidden = arr["fine_hidden"].ctypes.data_as(C.POINTER(C.c_long))
        d.fine_diff2 = arr["fine_diff2"].ctypes.data_as(C.POINTER(C.c_float))
        d.fine_count = arr["fine_count"].ctypes.data_as(C.POINTER(C.c_long))
        pins = (T.ParticleIn * 1)(pin or self.pin)
        im = np.ascontiguousarray(img, dtype=np.float32)
        st = orc.orc_estep(C.byref(cfg), C.byref(model), projs, bps, 1, T.fptr(im), pins, outs, C.byref(ws),
                           C.byref(d), O.ORC_PROJ_CLEAN, 1)
        assert st == 0, st
        bp = [(bd[0::2] + 1j * bd[1::2], bw) for bd, bw in keep]
        return T.outs_to_dict(outs), wa, arr, bp

    def hidden(self, c, idir, ipsi, it):
        i = self.info
        return ((c * i.nr_dir + idir) * i.nr_psi + ipsi) * i.nr_trans + it


def sigma2_by_shell(scale):
    return scale * (1.0 + np.arange(N // 2 + 1) / 7.0)


def highres(F, cs):
    ip, jp = half_freqs(N)
    r = shells(N)
    ok = (r > 0) & (r < N // 2 + 1) & ~((jp == 0) & (ip < 0)) & (r >= cs // 2 + 1)
    return float(np.sum(np.abs(F[ok]) ** 2))


def ring_term(F, s, sigma2):
    """1/2 sum Minv |F|^2 over the Mresol pixels (ires >= 1) of the size-s window that lie
    outside the projector's disc: the projection is zero there, the image is not."""
    ok = mresol(s) & (shells(s) >= 1) & ~in_disc(s)
    Fw = window(F, s)
    return 0.5 * float(np.sum(np.abs(Fw[ok]) ** 2 / sigma2[shells(s)[ok]]))


def add_delta(F, ip, jp, delta):
    G = F.copy()
    G[ip % N, jp] += delta
    return G


HIGH = (3, 13)    # (ip, jp), shell 13: beyond the current size 24


@pytest.fixture(scope="module")
def exact(orc):
    return Exact(orc, base_config())


def test_diff2_exact_particle_coarse_pass(exact):
    """True pose on the coarse grid (class 1, direction 7, psi 4, translation (0, -4) px): the
    coarse diff2 there is highres_Xi2/2 + the disc ring of the 16-pixel window, and it is the
    particle's minimum."""
    ex = exact
    c, idir, ipsi, it = 1, 7, 4, 1
    F, _ = ex.projection(c, ex.rot[idir], ex.tilt[idir], ex.psi[ipsi])
    F = add_delta(F * ex.ctf, *HIGH, 0.5)
    t = (ex.tx[it] / PIX, ex.ty[it] / PIX)
    img = ex.image(F, t)
    Fa = relion_fft(img.astype(np.float64))
    s2 = sigma2_by_shell(0.1)
    _, _, arr, _ = ex.run(img, s2)
    d2 = arr["coarse_diff2"][0]
    h = ex.hidden(c, idir, ipsi, it)
    want = highres(Fa, 24) / 2 + ring_term(Fa, 16, s2)
    assert highres(Fa, 24) > 0.2
    assert d2[h] == pytest.approx(want, rel=2e-6)
    valid = d2 > -998
    assert d2[h] == d2[valid].min()
    # one pixel changed by delta: inside the coarse window (shell 5) adds 1/2 Minv |delta|^2 at
    # that shell; at the origin it adds nothing; beyond the current size |delta|^2 / 2
    for (ip, jp, delta, add) in ((-3, 4, 0.3 - 0.2j, 0.5 * abs(0.3 - 0.2j) ** 2 / s2[5]),
                                 (0, 0, 0.7, 0.0),
                                 (HIGH[0], HIGH[1], 0.25j, None)):
        G = add_delta(F, ip, jp, delta)
        img2 = ex.image(G, t)
        Ga = relion_fft(img2.astype(np.float64))
        _, _, arr2, _ = ex.run(img2, s2)
        if add is None:   # highres_Xi2 grows by |G|^2 - |F|^2 at that pixel
            add = 0.5 * (abs(Ga[HIGH[0] % N, HIGH[1]]) ** 2 - abs(Fa[HIGH[0] % N, HIGH[1]]) ** 2)
        assert arr2["coarse_diff2"][0][h] - d2[h] == pytest.approx(add, rel=1e-5, abs=2e-6)


def test_diff2_exact_particle_fine_pass(exact):
    """True pose at an oversampled orientation (NEST child 1, second sub-psi of direction 7,
    psi 4) and oversampled translation (+1, -5) px of coarse translation (0, -4): its fine
    diff2 is highres_Xi2/2 + the disc ring of the 24-pixel window, the particle's minimum, and a
    delta between the coarse and the current size moves it by 1/2 Minv |delta|^2 and leaves
    the coarse diff2 unchanged."""
    ex = exact
    info = ex.info
    c, idir, ipsi, it, ir, iot = 1, 7, 4, 1, 3, 2
    orot, otilt, opsi = np.zeros(8), np.zeros(8), np.zeros(8)
    ex.orc.orc_sampling_oversampled_orientations(C.byref(ex.cfg), idir, ipsi, 1, T.dptr(orot), T.dptr(otilt), T.dptr(opsi))
    ox, oy = np.zeros(4), np.zeros(4)
    ex.orc.orc_sampling_oversampled_translations(C.byref(ex.cfg), it, 1, T.dptr(ox), T.dptr(oy))
    assert (ox[iot], oy[iot]) == (1.0, -5.0)
    F, _ = ex.projection(c, orot[ir], otilt[ir], opsi[ir])
    F = add_delta(F * ex.ctf, *HIGH, 0.5)
    img = ex.image(F, (ox[iot], oy[iot]))
    Fa = relion_fft(img.astype(np.float64))
    s2 = sigma2_by_shell(0.1)
    out, _, arr, _ = ex.run(img, s2)
    nover = info.nr_over_rot * info.nr_over_trans
    target = ex.hidden(c, idir, ipsi, it) * nover + ir * info.nr_over_trans + iot
    nf = arr["fine_count"][0]
    fh, fd = arr["fine_hidden"][0, :nf], arr["fine_diff2"][0, :nf]
    k = np.flatnonzero(fh == target)
    assert len(k) == 1
    want = highres(Fa, 24) / 2 + ring_term(Fa, 24, s2)
    assert fd[k[0]] == pytest.approx(want, rel=2e-6)
    assert fd[k[0]] == fd.min()
    assert out["best_hidden"][0] == target
    assert out["min_diff2"][0] == pytest.approx(want, rel=2e-6)
    # shell 10: in the 24-pixel window, outside the 16-pixel one
    delta = 0.2 + 0.1j
    G = add_delta(F, 6, 8, delta)
    img2 = ex.image(G, (ox[iot], oy[iot]))
    out2, _, arr2, _ = ex.run(img2, s2)
    nf2 = arr2["fine_count"][0]
    k2 = np.flatnonzero(arr2["fine_hidden"][0, :nf2] == target)
    assert arr2["fine_diff2"][0, k2[0]] - fd[k[0]] == pytest.approx(0.5 * abs(delta) ** 2 / s2[10], rel=1e-5)
    sig = arr["coarse_diff2"][0] > -998
    assert np.array_equal(sig, arr2["coarse_diff2"][0] > -998)
    assert np.abs(arr2["coarse_diff2"][0][sig] - arr["coarse_diff2"][0][sig]).max() < 1e-5 * np.abs(
        arr["coarse_diff2"][0][sig]).max()


def test_weighted_sums_single_significant_sample(orc):
    """One pass at the current size (no oversampling), noise-free particle at a sampled pose and
    zero translation, sigma2 small enough that every other sample is > 700 above the minimum:
    the one significant sample has normalised weight 1 and every weighted sum has a closed form."""
    cfg = base_config(ao=0)
    cfg.do_norm_correction = 1
    cfg.avg_norm_correction = 0.8
    cfg.sigma2_offset = 1.0e30                  # a flat offset prior: every normalised prior is 1
    ex = Exact(orc, cfg, seed=9)
    info = ex.info
    c, idir, ipsi, it = 0, 11, 2, 2
    assert (ex.tx[it], ex.ty[it]) == (0.0, 0.0)
    F, A = ex.projection(c, ex.rot[idir], ex.tilt[idir], ex.psi[ipsi])
    d0, d1 = 0.05, 0.02 - 0.03j
    G = add_delta(add_delta(add_delta(F * ex.ctf, *HIGH, 0.5), 0, 0, d0), -3, 4, d1)
    norm = 1.6                                  # the image is scaled by avg / norm = 0.5 first
    img = ex.image(G * (norm / cfg.avg_norm_correction), (0.0, 0.0))
    Fa = relion_fft(img.astype(np.float64)) * (cfg.avg_norm_correction / norm)
    s2 = sigma2_by_shell(2e-6)
    pin = particle(norm=norm)
    pin.ctf = ex.pin.ctf
    out, ws, _, bp = ex.run(img, s2, n_debug=0, pin=pin)
    assert out["best_hidden"][0] == ex.hidden(c, idir, ipsi, it)
    assert out["pmax"][0] == pytest.approx(1.0, rel=0, abs=1e-12)
    assert out["nr_fine_significant"][0] == 1
    # sum_weight = the normalised priors of the sample: 1 for uniform pdfs and a flat offset prior
    assert out["sum_weight"][0] == pytest.approx(1.0, rel=1e-14)
    cs = cfg.current_size
    mres, r24 = mresol(cs), shells(cs)
    min_d2 = highres(Fa, cs) / 2 + ring_term(Fa, cs, s2) + 0.5 * abs(d1) ** 2 / s2[5]
    assert out["min_diff2"][0] == pytest.approx(min_d2, rel=1e-5)
    # sigma2_noise: sum over the window's Mresol pixels (origin included) of |ctf ref - X|^2 =
    # the deltas and the disc ring, plus the power of the image beyond the current size
    Fw = window(Fa, cs)
    resid = np.where(in_disc(cs), 0.0, np.abs(Fw) ** 2)
    resid[0, 0] = abs(Fw[0, 0] - window(F * ex.ctf, cs)[0, 0]) ** 2
    resid[-3 % cs, 4] = abs(Fw[-3 % cs, 4] - window(F * ex.ctf, cs)[-3 % cs, 4]) ** 2
    want = np.bincount(r24[mres], weights=resid[mres], minlength=N // 2 + 1)
    ipN, jpN = half_freqs(N)
    rN = shells(N)
    beyond = (rN >= cs // 2 + 1) & (rN < N // 2 + 1) & ~((jpN == 0) & (ipN < 0))
    want += np.bincount(rN[beyond], weights=np.abs(Fa[beyond]) ** 2, minlength=N // 2 + 1)
    assert want[0] == pytest.approx(abs(d0) ** 2, rel=1e-4)
    assert np.abs(ws["sigma2_noise"] - want).max() < 1e-6 * want.max()
    # dLL and the norm update
    logsig = float(np.sum(np.log(2 * np.pi * s2[r24[mres & (r24 >= 1)]])))
    dll = np.log(out["sum_weight"][0]) - out["min_diff2"][0] - logsig
    assert out["dLL"][0] == pytest.approx(dll, rel=1e-12)
    nc = norm / cfg.avg_norm_correction * np.sqrt(2 * want.sum())
    assert out["norm_correction"][0] == pytest.approx(nc, rel=1e-6)
    assert ws["avg_norm_correction"][0] == pytest.approx(nc, rel=1e-6)
    assert ws["LL"][0] == pytest.approx(dll, rel=1e-12)
    # pose and offset sums of the one sample
    assert ws["sumw_group"][0] == pytest.approx(1.0, rel=0, abs=1e-12)
    assert ws["pdf_class"][c] == pytest.approx(1.0, rel=0, abs=1e-12) and ws["pdf_class"][1 - c] == 0.0
    pd = np.zeros(2 * info.nr_dir)
    pd[c * info.nr_dir + idir] = 1.0
    assert np.abs(ws["pdf_direction"] - pd).max() < 1e-12
    assert ws["sigma2_offset"][0] == 0.0
    assert ws["ave_Pmax"][0] == pytest.approx(1.0, rel=0, abs=1e-12)
    # backprojection: ctf Minv X and ctf^2 Minv with the origin term 1 / sigma2[0]
    # (ml_optimiser.cpp:8099-8104), at the sample's orientation
    ctfw = window(ex.ctf, cs)
    minv = np.where(mres, 1.0 / s2[np.minimum(r24, N // 2)], 0.0)
    f2d = np.ascontiguousarray(np.stack([(ctfw * minv * Fw).real, (ctfw * minv * Fw).imag], -1).ravel())
    fw = np.ascontiguousarray((ctfw * ctfw * minv).ravel())
    v, bd, bw = O.make_volume(3, ex.pad, ex.rmax, 2.0, True)
    orc.orc_backproject(C.byref(v), T.dptr(f2d), T.dptr(fw), cs, T.dptr(A))
    ref = bd[0::2] + 1j * bd[1::2]
    assert np.linalg.norm(bp[c][0] - ref) < 1e-9 * np.linalg.norm(ref)
    assert np.linalg.norm(bp[c][1] - bw) < 1e-9 * np.linalg.norm(bw)
    assert np.abs(bp[1 - c][1]).sum() == 0.0


def test_backprojection_uses_the_unmasked_image(orc):
    """With a mask that changes the particle (radius 6 px), the weighted sums still backproject
    the transform of the unmasked image; the masked one differs by far more than the tolerance."""
    cfg = base_config(ao=0, diameter=2 * 6 * PIX)
    ex = Exact(orc, cfg, seed=9)
    c, idir, ipsi, it = 0, 11, 2, 2
    F, A = ex.projection(c, ex.rot[idir], ex.tilt[idir], ex.psi[ipsi])
    img = ex.image(F * ex.ctf, (0.0, 0.0))
    s2 = sigma2_by_shell(2e-6)
    out, _, _, bp = ex.run(img, s2, n_debug=0)
    assert out["best_hidden"][0] == ex.hidden(c, idir, ipsi, it)
    assert out["pmax"][0] == pytest.approx(1.0, rel=0, abs=1e-12)
    cs = cfg.current_size
    mres, r24 = mresol(cs), shells(cs)
    ctfw = window(ex.ctf, cs)
    minv = np.where(mres, 1.0 / s2[np.minimum(r24, N // 2)], 0.0)

    def bp_of(X):
        f2d = np.ascontiguousarray(np.stack([(ctfw * minv * X).real, (ctfw * minv * X).imag], -1).ravel())
        fw = np.ascontiguousarray((ctfw * ctfw * minv).ravel())
        v, bd, bw = O.make_volume(3, ex.pad, ex.rmax, 2.0, True)
        orc.orc_backproject(C.byref(v), T.dptr(f2d), T.dptr(fw), cs, T.dptr(A))
        return bd[0::2] + 1j * bd[1::2]
    unmasked = bp_of(window(relion_fft(img.astype(np.float64)), cs))
    r = preprocess(orc, cfg, img, ex.pin)
    masked = bp_of(r["Fimg"])
    assert np.linalg.norm(bp[c][0] - unmasked) < 1e-9 * np.linalg.norm(unmasked)
    assert np.linalg.norm(masked - unmasked) > 1e-2 * np.linalg.norm(unmasked)
