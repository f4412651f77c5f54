"""Local angular searches (orientational priors) in the FP64 oracle, pinned on the CPU.

HealpixSampling::selectOrientationsWithNonZeroPriorProbability (healpix_sampling.cpp:769-1145)
keeps the directions within 3 sigma of the prior direction with a Gaussian weight on the angle
between them, the psi angles within 3 sigma_psi on the circular difference, normalises both, and
falls back to the nearest sample when none is in range; the E-step then weighs orientation
(idir, ipsi) by directions_prior * psi_prior / mean (ml_optimiser.cpp:7031-7045, 7600-7716).

Pins: the selected sets and prior ratios against angles computed independently (atan2 of the
cross and dot products of unit vectors); the nearest-sample fallback; and the whole local E-step
with very wide sigmas against the global search (every orientation in range with equal prior,
so both must give the same weights, poses and sums).
"""
import ctypes as C

import numpy as np
import pytest

import estep_harness as H
import oracle_py as O
import t2x_py as T
from t2x_py import workloads as W


@pytest.fixture(scope="module")
def orc():
    return O.load()


def local_cfg(name="refine64", sig=(10.0, 10.0, 10.0)):
    cfg, _ = W.make_config(W.WORKLOADS[name])
    cfg.orientational_prior_mode = T.T2X_PRIOR_ROTTILT_PSI
    cfg.sigma_rot, cfg.sigma_tilt, cfg.sigma_psi = sig
    return cfg


def particle(rot, tilt, psi, prior=None):
    p = T.ParticleIn()
    p.rot, p.tilt, p.psi = rot, tilt, psi
    if prior is None:
        p.prior_rot = p.prior_tilt = p.prior_psi = 999.0
    else:
        p.prior_rot, p.prior_tilt, p.prior_psi = prior
    return p


def unit(rot, tilt):
    r, t = np.radians(rot), np.radians(tilt)
    return np.stack([np.sin(t) * np.cos(r), np.sin(t) * np.sin(r), np.cos(t)], axis=-1)


def grid(orc, cfg):
    info = T.SamplingInfo()
    assert orc.orc_sampling_info(C.byref(cfg), C.byref(info)) == 0
    nd, npsi = info.nr_dir, info.nr_psi
    rot, tilt, psi = np.zeros(nd), np.zeros(nd), np.zeros(npsi)
    tx, ty = np.zeros(info.nr_trans), np.zeros(info.nr_trans)
    assert orc.orc_sampling_coarse(C.byref(cfg), T.dptr(rot), T.dptr(tilt), T.dptr(psi), T.dptr(tx), T.dptr(ty)) == 0
    return rot, tilt, psi


def priors(orc, cfg, p, nd, npsi):
    dpr, ppr = np.zeros(nd), np.zeros(npsi)
    a, b = C.c_long(), C.c_long()
    assert orc.orc_local_priors(C.byref(cfg), C.byref(p), T.dptr(dpr), T.dptr(ppr), C.byref(a), C.byref(b)) == 0
    return dpr, ppr, a.value, b.value


@pytest.mark.parametrize("sig,sig_psi", [(10.0, 10.0), (6.0, 20.0), (25.0, 4.0)])
def test_selection_geometry(orc, sig, sig_psi):
    cfg = local_cfg(sig=(sig, sig, sig_psi))
    rot, tilt, psi = grid(orc, cfg)
    rng = np.random.default_rng(int(sig * 10 + sig_psi))
    for _ in range(6):
        pr, pt, pp = rng.uniform(-180, 180), np.degrees(np.arccos(rng.uniform(-1, 1))), rng.uniform(-180, 180)
        dpr, ppr, nd, npsi = priors(orc, cfg, particle(0, 0, 0, (pr, pt, pp)), len(rot), len(psi))
        u, v = unit(rot, tilt), unit(pr, pt)
        ang = np.degrees(np.arctan2(np.linalg.norm(np.cross(u, v), axis=1), u @ v))
        clear = np.abs(ang - 3 * sig) > 1e-7
        assert np.array_equal((dpr > 0)[clear], (ang < 3 * sig)[clear])
        assert nd == (dpr > 0).sum() and dpr.sum() == pytest.approx(1.0, rel=1e-13)
        sel = np.flatnonzero(dpr > 0)
        ref = np.exp(-0.5 * (ang[sel] / sig) ** 2)
        np.testing.assert_allclose(dpr[sel], ref / ref.sum(), rtol=1e-9)
        dpsi = np.abs((psi - pp + 180.0) % 360.0 - 180.0)
        clear = np.abs(dpsi - 3 * sig_psi) > 1e-7
        assert np.array_equal((ppr > 0)[clear], (dpsi < 3 * sig_psi)[clear])
        assert npsi == (ppr > 0).sum() and ppr.sum() == pytest.approx(1.0, rel=1e-13)
        sel = np.flatnonzero(ppr > 0)
        ref = np.exp(-0.5 * (dpsi[sel] / sig_psi) ** 2)
        np.testing.assert_allclose(ppr[sel], ref / ref.sum(), rtol=1e-9)


def test_unset_priors_use_current_angles_and_fallback(orc):
    cfg = local_cfg(sig=(10.0, 10.0, 10.0))
    rot, tilt, psi = grid(orc, cfg)
    a = priors(orc, cfg, particle(30.0, 70.0, -45.0), len(rot), len(psi))
    b = priors(orc, cfg, particle(0.0, 0.0, 0.0, (30.0, 70.0, -45.0)), len(rot), len(psi))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    # sigmas far below the sampling: the nearest direction and psi angle alone, prior 1
    cfg = local_cfg(sig=(1e-4, 1e-4, 1e-4))
    dpr, ppr, nd, npsi = priors(orc, cfg, particle(0, 0, 0, (33.3, 71.1, -44.4)), len(rot), len(psi))
    u, v = unit(rot, tilt), unit(33.3, 71.1)
    assert nd == 1 and dpr[np.argmax(u @ v)] == 1.0
    assert npsi == 1 and ppr[np.argmin(np.abs((psi + 44.4 + 180) % 360 - 180))] == 1.0
    # no direction prior (sigma_rot = sigma_tilt = 0): every direction, equal prior
    cfg = local_cfg(sig=(0.0, 0.0, 10.0))
    dpr, ppr, nd, npsi = priors(orc, cfg, particle(0, 0, 0, (10.0, 20.0, 30.0)), len(rot), len(psi))
    assert nd == len(rot) and np.allclose(dpr, 1.0 / len(rot), rtol=1e-14)


def test_rot_only_prior_rejected(orc):
    cfg = local_cfg(sig=(10.0, 0.0, 10.0))
    info = T.SamplingInfo()
    assert orc.orc_sampling_info(C.byref(cfg), C.byref(info)) != 0


def test_wide_local_search_equals_global_search():
    """With sigmas of 1e8 degrees every orientation is in range with a prior equal to within
    2e-12, so the local E-step must reproduce the global one (uniform pdf_direction)."""
    g = H.Workload("refine64", n=4, seed_offset=21)
    og = g.run_oracle(n_debug=4, fine_cap=16384)
    loc = H.Workload("refine64", n=4, seed_offset=21)
    loc.cfg.orientational_prior_mode = T.T2X_PRIOR_ROTTILT_PSI
    loc.cfg.sigma_rot = loc.cfg.sigma_tilt = loc.cfg.sigma_psi = 1e8
    ol = loc.run_oracle(n_debug=4, fine_cap=16384)
    assert np.array_equal(og["out"]["best_hidden"], ol["out"]["best_hidden"])
    assert np.array_equal(og["dbg"]["coarse_significant"], ol["dbg"]["coarse_significant"])
    for k in ("dLL", "pmax", "sum_weight", "min_diff2"):
        np.testing.assert_allclose(ol["out"][k], og["out"][k], rtol=1e-9)
    for k, v in og["wsum"].items():
        np.testing.assert_allclose(ol["wsum"][k], v, rtol=1e-8, atol=1e-12 * np.abs(v).max())
    np.testing.assert_allclose(ol["bp"][0][1], og["bp"][0][1], rtol=1e-8, atol=1e-12 * np.abs(og["bp"][0][1]).max())


def test_local_search_stays_in_range():
    """Narrow priors around the true poses: every particle's best direction and psi lie within
    3 sigma of its prior, and unscored orientations carry no diff2."""
    wl = H.Workload("refine64", n=4, seed_offset=22, snr=1.0)
    wl.cfg.orientational_prior_mode = T.T2X_PRIOR_ROTTILT_PSI
    wl.cfg.sigma_rot = wl.cfg.sigma_tilt = wl.cfg.sigma_psi = 8.0
    rng = np.random.default_rng(5)
    tr = wl.truth
    pri = []
    for i in range(wl.n):
        pr = (tr["rot"][i] + rng.normal(0, 3), tr["tilt"][i] + rng.normal(0, 3), tr["psi"][i] + rng.normal(0, 3))
        pr = (pr[0], min(max(pr[1], 0.0), 180.0), (pr[2] + 180.0) % 360.0 - 180.0)
        wl.pin[i].prior_rot, wl.pin[i].prior_tilt, wl.pin[i].prior_psi = pr
        pri.append(pr)
    o = wl.run_oracle(n_debug=4, fine_cap=16384)
    out = o["out"]
    for i in range(wl.n):
        u, v = unit(out["rot"][i], out["tilt"][i]), unit(pri[i][0], pri[i][1])
        assert np.degrees(np.arccos(np.clip(u @ v, -1, 1))) < 3 * 8.0 + 7.5   # + half a coarse step
        dpsi = abs((out["psi"][i] - pri[i][2] + 180.0) % 360.0 - 180.0)
        assert dpsi < 3 * 8.0 + 7.5
    scored = (o["dbg"]["coarse_diff2"] > -998).mean()
    assert 0 < scored < 0.2
