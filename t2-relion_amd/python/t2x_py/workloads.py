"""Synthetic E-step workloads (Gaussian-blob references, CTF-modulated noisy particles).

The iteration state follows MlOptimiser::updateImageSizeAndResolutionPointers
(ml_optimiser.cpp:5532-5600): current_size = 2*(pixel(current_resolution) + incr_size),
coarse_size = 2*CEIL(N*angpix / (rotated_distance/keepsafe)) capped by current_size, and the
projector pad_size = 2*(ROUND(pad*r_max)+1)+1 with r_max = current_size/2 (projector.cpp:55-85).
Random numbers come from numpy's PCG64 with the seeds fixed below.
"""
import math

import numpy as np

from . import Blob, CtfParams, EstepConfig, ParticleIn

T2X_PREC_FP32 = 0

# name -> workload description (box, sampling, data)
WORKLOADS = {
    # Plain 3D refinement, CPU-runnable: 64^3 @2 A, 200 blobs, 1000 particles, K=1,
    # HEALPix 2, offsets 5/1 px, mask 120 A, SNR 0.05, initial model low-passed to 30 A.
    "refine64": dict(ori_size=64, pixel_size=2.0, n_blobs=200, blob_radius=24.0,
                     blob_sigma=(1.5, 3.0), blob_amp=(0.5, 1.0), K=1, n_particles=1000,
                     max_shift=4.0, snr=0.05, healpix_order=2, offset_range_px=5.0,
                     offset_step_px=1.0, particle_diameter=120.0, resolution=30.0,
                     seed=20240),
    # 3D classification: 4 refs 256^3 @1.1 A (800 blobs, r<=100 px), 20000 particles,
    # shifts +-6 px, SNR 0.03, HEALPix 3, offsets 6/1.5 px.
    "class256": dict(ori_size=256, pixel_size=1.1, n_blobs=800, blob_radius=100.0,
                     blob_sigma=(1.5, 3.0), blob_amp=(0.5, 1.0), K=4, n_particles=20000,
                     max_shift=6.0, snr=0.03, healpix_order=3, offset_range_px=6.0,
                     offset_step_px=1.5, particle_diameter=240.0, resolution=10.0,
                     seed=20241),
    # Local refinement: 360^3 @0.9 A (1500 blobs, r<=150 px), 50000 particles 360^2, SNR 0.02,
    # local searches at HEALPix 5 (12288 directions x 192 psi) with sigma_ang 3 deg around prior
    # angles 1 deg (rms) from the true poses, offsets 3/0.5 px; resolution 7 A (current size 112).
    "local360": dict(ori_size=360, pixel_size=0.9, n_blobs=1500, blob_radius=150.0,
                     blob_sigma=(1.5, 3.0), blob_amp=(0.5, 1.0), K=1, n_particles=50000,
                     max_shift=3.0, snr=0.02, healpix_order=5, offset_range_px=3.0,
                     offset_step_px=0.5, particle_diameter=300.0, resolution=7.0,
                     sigma_ang=3.0, prior_noise=1.0, seed=20244),
    # 2D classification (parity size): 4 class averages = z-projections of 64^3 blob volumes
    # @2 A, psi step 6 deg (60 in-plane angles), offsets 5/1 px, SNR 0.05.
    "class2d64": dict(ori_size=64, pixel_size=2.0, n_blobs=150, blob_radius=24.0,
                      blob_sigma=(1.5, 3.0), blob_amp=(0.5, 1.0), K=4, n_particles=2000,
                      max_shift=4.0, snr=0.05, ref_dim=2, psi_step=6.0, healpix_order=0,
                      offset_range_px=5.0, offset_step_px=1.0, particle_diameter=110.0,
                      resolution=20.0, seed=20242),
    # 2D classification with uneven workloads: 100 class averages 128^2 @2 A (z-projections of
    # 300-blob volumes), 100,000 particles with Zipf(1.2) class frequencies, psi step 6 deg,
    # offsets 5/1 px, SNR log-uniform 0.01-0.1.
    "class2d128": dict(ori_size=128, pixel_size=2.0, n_blobs=300, blob_radius=48.0,
                       blob_sigma=(1.5, 3.0), blob_amp=(0.5, 1.0), K=100, n_particles=100000,
                       max_shift=5.0, snr=(0.01, 0.1), zipf=1.2, ref_dim=2, psi_step=6.0,
                       healpix_order=0, offset_range_px=5.0, offset_step_px=1.0,
                       particle_diameter=200.0, resolution=12.0, seed=20243),
}


def pixel_from_resolution(res_A, N, angpix):
    """MlModel::getPixelFromResolution: ROUND(N*angpix/res)."""
    x = N * angpix / res_A
    return int(x + 0.5)


def derived_sizes(N, angpix, healpix_order, particle_diameter, resolution, incr_size=10,
                  ref_dim=3, psi_step=None):
    maxres = pixel_from_resolution(resolution, N, angpix) + incr_size
    current = min(2 * maxres, N)
    ang = 360.0 / (6 * 2 ** healpix_order) if ref_dim == 3 else psi_step
    rotated_distance = ang / 360.0 * math.pi * particle_diameter
    keepsafe = 1.2 if ref_dim == 3 else 1.5
    coarse_res = rotated_distance / keepsafe
    x = angpix * N / coarse_res
    coarse = 2 * (int(x) if x == int(x) else int(x) + 1)
    coarse = min(coarse, current)
    r_max = current // 2
    pad = 2 * (int(2 * r_max + 0.5) + 1) + 1
    return current, coarse, r_max, pad


def make_config(w, *, current_size=None, coarse_size=None, precision=T2X_PREC_FP32,
                maximum_significants=0, K=None):
    N, angpix = w["ori_size"], w["pixel_size"]
    ref_dim = w.get("ref_dim", 3)
    cur, coarse, r_max, pad = derived_sizes(N, angpix, w["healpix_order"],
                                            w["particle_diameter"], w["resolution"],
                                            ref_dim=ref_dim, psi_step=w.get("psi_step"))
    if current_size:
        cur = current_size
        r_max = cur // 2
        pad = 2 * (int(2 * r_max + 0.5) + 1) + 1
        coarse = min(coarse, cur)
    if coarse_size:
        coarse = coarse_size
    ang = 360.0 / (6 * 2 ** w["healpix_order"]) if ref_dim == 3 else w["psi_step"]
    cfg = EstepConfig()
    cfg.ori_size = N
    cfg.pixel_size = angpix
    cfg.current_size = cur
    cfg.coarse_size = coarse
    cfg.nr_classes = K or w["K"]
    cfg.ref_dim = ref_dim
    cfg.healpix_order = w["healpix_order"]
    cfg.psi_step = ang
    cfg.offset_range = w["offset_range_px"] * angpix
    cfg.offset_step = w["offset_step_px"] * angpix
    cfg.adaptive_oversampling = 1
    cfg.adaptive_fraction = 0.999
    cfg.maximum_significants = maximum_significants
    cfg.particle_diameter = w["particle_diameter"]
    cfg.width_mask_edge = 5
    cfg.do_zero_mask = 1                   # relion_refine --zero_mask (average background)
    cfg.sigma2_fudge = 1.0
    cfg.do_norm_correction = 0
    cfg.avg_norm_correction = 1.0
    cfg.do_ctf_correction = 1
    cfg.sigma2_offset = 10.0 ** 2          # relion_refine --offset 10 (A)
    cfg.orientational_prior_mode = 0
    if w.get("sigma_ang"):   # local searches (relion_refine --sigma_ang)
        cfg.orientational_prior_mode = 1
        cfg.sigma_rot = cfg.sigma_tilt = cfg.sigma_psi = float(w["sigma_ang"])
    cfg.precision = precision
    return cfg, dict(current_size=cur, coarse_size=coarse, r_max=r_max, pad_size=pad)


def gen_blobs(rng, n, radius, sig, amp):
    """Centres uniform in a ball of `radius` px, sigma U[sig], amplitude U[amp]."""
    out = (Blob * n)()
    for k in range(n):
        while True:
            c = rng.uniform(-radius, radius, 3)
            if c @ c <= radius * radius:
                break
        out[k] = Blob(c[0], c[1], c[2], rng.uniform(*sig), rng.uniform(*amp))
    return out


def gen_particles(rng, n, K, max_shift, def_range=(10000.0, 30000.0), voltage=300.0, cs=2.7,
                  q0=0.1, integer_shifts=True, class_weights=None):
    """Uniform SO(3) poses (ZYZ Haar measure), shifts, CTF parameters and class labels."""
    rot = rng.uniform(-180.0, 180.0, n)
    tilt = np.degrees(np.arccos(rng.uniform(-1.0, 1.0, n)))
    psi = rng.uniform(-180.0, 180.0, n)
    if integer_shifts:
        sx = rng.integers(-int(max_shift), int(max_shift) + 1, n).astype(np.float64)
        sy = rng.integers(-int(max_shift), int(max_shift) + 1, n).astype(np.float64)
    else:
        sx = rng.uniform(-max_shift, max_shift, n)
        sy = rng.uniform(-max_shift, max_shift, n)
    du = rng.uniform(*def_range, n)
    dv = rng.uniform(*def_range, n)
    ang = rng.uniform(0.0, 180.0, n)
    if class_weights is None:
        cls = rng.integers(0, K, n).astype(np.int32)
    else:
        p = np.asarray(class_weights, float)
        cls = rng.choice(K, size=n, p=p / p.sum()).astype(np.int32)
    pin = (ParticleIn * n)()
    for i in range(n):
        pin[i].ctf = CtfParams(voltage, cs, q0, du[i], dv[i], ang[i], 0.0, 0.0, 1.0)
        pin[i].old_offset_x = 0.0
        pin[i].old_offset_y = 0.0
        pin[i].prior_offset_x = 999.0
        pin[i].prior_offset_y = 999.0
        pin[i].norm_correction = 1.0
        pin[i].rot = pin[i].tilt = pin[i].psi = 0.0
        pin[i].prior_rot = pin[i].prior_tilt = pin[i].prior_psi = 999.0
    truth = dict(rot=rot, tilt=tilt, psi=psi, x=sx, y=sy, cls=cls)
    return pin, truth


def workload_particles(rng, w, n, K):
    """Particles of a workload: poses and CTFs (gen_particles), class labels with Zipf(s) class
    frequencies when the workload names one, per-particle SNR (a constant, or log-uniform over a
    (lo, hi) range), in-plane poses only (rot = tilt = 0) for 2D references."""
    cw = None
    if w.get("zipf"):
        cw = 1.0 / np.arange(1, K + 1) ** float(w["zipf"])
    pin, tr = gen_particles(rng, n, K, w["max_shift"], class_weights=cw)
    if w.get("ref_dim", 3) == 2:
        tr["rot"][:] = 0.0
        tr["tilt"][:] = 0.0
    if w.get("sigma_ang"):   # prior angles near the true poses
        e = float(w.get("prior_noise", 1.0))
        for i in range(n):
            pin[i].prior_rot = tr["rot"][i] + rng.normal(0.0, e)
            pin[i].prior_tilt = min(max(tr["tilt"][i] + rng.normal(0.0, e), 0.0), 180.0)
            pin[i].prior_psi = (tr["psi"][i] + rng.normal(0.0, e) + 180.0) % 360.0 - 180.0
    s = w["snr"]
    if isinstance(s, (tuple, list)):
        snr = np.exp(rng.uniform(np.log(s[0]), np.log(s[1]), n))
    else:
        snr = np.full(n, float(s))
    return pin, tr, snr


def mask_power_fraction(N, angpix, particle_diameter, width=5):
    """Mean square of the soft circular mask (mask.cpp:43-95 profile)."""
    radius = particle_diameter / (2 * angpix)
    y, x = np.mgrid[0:N, 0:N] - N // 2
    r = np.sqrt(x * x + y * y)
    m = np.where(r < radius, 1.0, 0.0)
    edge = (r >= radius) & (r <= radius + width)
    m[edge] = 1.0 - (0.5 + 0.5 * np.cos(np.pi * (radius + width - r[edge]) / width))
    return float((m * m).mean())


def model_arrays(cfg, n_dir, noise_var):
    """sigma2_noise from the known white-noise variance (per-component variance of the
    normalised FFT, scaled by the mask power), uniform pdf_class and pdf_direction."""
    N, K = cfg.ori_size, cfg.nr_classes
    f = mask_power_fraction(N, cfg.pixel_size, cfg.particle_diameter, cfg.width_mask_edge)
    s2 = float(np.mean(noise_var)) * f / (2.0 * N * N)
    sigma2 = np.full(N // 2 + 1, s2)
    pdf_class = np.full(K, 1.0 / K)
    pdf_dir = np.full(K * n_dir, 1.0 / n_dir)
    return sigma2, pdf_class, pdf_dir
