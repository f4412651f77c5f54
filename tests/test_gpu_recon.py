"""t2x_reconstruct / t2x_fsc on the GPU against the FP64 oracle, and the half-map FSC of the
whole GPU path (E-step + reconstruction) against the CPU path (oracle E-step + oracle
reconstruction) on the same particles.

Tolerances: reconstruction and FSC are FP64 on both sides and differ only by FFT algorithm and
summation order: 1e-9 relative (Frobenius) on maps, 1e-9 absolute on FSC values. The half-map
FSC of the plain FP32 GPU path is held to a bar of 0.01 at every shell.
"""
import numpy as np
import pytest

import estep_harness as H
import oracle_py as O
import t2x_py as T

pytestmark = pytest.mark.gpu


def rand_volumes(p, rng):
    pad = p.pad_size
    xh = pad // 2 + 1
    shape = (pad, pad, xh)
    re, im = rng.standard_normal(shape), rng.standard_normal(shape)
    w = rng.uniform(0.0, 3.0, shape)
    w[rng.uniform(size=shape) < 0.05] = 0.0     # some empty voxels exercise the weight floor
    return re, im, w


def gpu_recon(dev, p, re, im, w, tau2=None):
    lib = T.load()
    out = np.zeros((p.ori_size,) * 3)
    t = None if tau2 is None else np.ascontiguousarray(tau2, dtype=np.float64)
    T.check(lib.t2x_reconstruct(dev, p, T.dptr(np.ascontiguousarray(re)), T.dptr(np.ascontiguousarray(im)),
                                T.dptr(np.ascontiguousarray(w)), T.dptr(t), T.dptr(out)), "reconstruct")
    return out


def gpu_recon_class(dev, iclass, p, tau2=None):
    out = np.zeros((p.ori_size,) * 3)
    t = None if tau2 is None else np.ascontiguousarray(tau2, dtype=np.float64)
    T.check(T.load().t2x_reconstruct_class(dev, iclass, p, T.dptr(t), T.dptr(out)), "reconstruct_class")
    return out


def orc_recon(p, re, im, w, tau2=None):
    out = np.zeros((p.ori_size,) * 3)
    t = None if tau2 is None else np.ascontiguousarray(tau2, dtype=np.float64)
    assert O.load().orc_reconstruct(p, T.dptr(np.ascontiguousarray(re)), T.dptr(np.ascontiguousarray(im)),
                                    T.dptr(np.ascontiguousarray(w)), T.dptr(t), T.dptr(out)) == 0
    return out


def gpu_fsc(dev, m1, m2):
    n = m1.shape[0]
    out = np.zeros(n // 2 + 1)
    T.check(T.load().t2x_fsc(dev, n, T.dptr(np.ascontiguousarray(m1)), T.dptr(np.ascontiguousarray(m2)),
                             T.dptr(out)), "fsc")
    return out


def orc_fsc(m1, m2):
    n = m1.shape[0]
    out = np.zeros(n // 2 + 1)
    assert O.load().orc_fsc(n, T.dptr(np.ascontiguousarray(m1)), T.dptr(np.ascontiguousarray(m2)), T.dptr(out)) == 0
    return out


@pytest.mark.parametrize("ori,r_max,do_map", [(16, 6, False), (16, 8, True), (32, 12, True)])
def test_reconstruct_matches_oracle(gpu_device, ori, r_max, do_map):
    rng = np.random.default_rng(ori + r_max)
    p = T.recon_params(ori, r_max, do_map=do_map, tau2_fudge=1.5, minres_map=1)
    re, im, w = rand_volumes(p, rng)
    tau2 = rng.uniform(0.1, 2.0, ori // 2 + 1) if do_map else None
    if do_map:
        tau2[3] = 0.0                     # the tau2 == 0 rule
    g = gpu_recon(gpu_device, p, re, im, w, tau2)
    o = orc_recon(p, re, im, w, tau2)
    err = H.rel_norm(g, o)
    print(f"\n[recon ori={ori} r_max={r_max} map={do_map}] rel {err:.2e}")
    assert err < 1e-9


def test_fsc_matches_oracle(gpu_device):
    rng = np.random.default_rng(3)
    m1 = rng.standard_normal((24, 24, 24))
    m2 = m1 + rng.standard_normal((24, 24, 24))
    assert np.abs(gpu_fsc(gpu_device, m1, m2) - orc_fsc(m1, m2)).max() < 1e-9


def test_full_box_closed_forms(gpu_device):
    """256 box, pf 2 (a 512^3 inverse transform): a constant component gives
    c / (pf^3 ori) / sinc^2(r / (ori pf)); FSC(m, m) = 1 and FSC(m, -m) = -1 at every shell."""
    ori, r_max = 256, 38
    p = T.recon_params(ori, r_max)
    pad = p.pad_size
    xh = pad // 2 + 1
    re, im, w = np.zeros((pad, pad, xh)), np.zeros((pad, pad, xh)), np.zeros((pad, pad, xh))
    re[pad // 2, pad // 2, 0] = 5.0
    w[pad // 2, pad // 2, 0] = 2.0
    out = gpu_recon(gpu_device, p, re, im, w)
    c = np.arange(ori) - ori // 2
    r = np.sqrt(c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2)
    rv = r / (ori * 2.0)
    s = np.where(r > 0, np.sin(np.pi * rv) / np.where(r > 0, np.pi * rv, 1.0), 1.0)
    expect = 2.5 / (8.0 * ori) / (s * s)
    assert H.rel_norm(out, expect) < 1e-12
    m = np.random.default_rng(1).standard_normal((ori, ori, ori))
    f = gpu_fsc(gpu_device, m, m)
    assert np.abs(f - 1.0).max() < 1e-12
    assert np.abs(gpu_fsc(gpu_device, m, -m) + 1.0).max() < 1e-12


def test_half_map_fsc_gpu_vs_cpu_path(gpu_device):
    """Two half-sets of refine64 particles: E-step + reconstruction on the GPU (plain FP32 path)
    and on the CPU oracle; the half-map FSC curves agree within 0.01 at every shell."""
    n = 32
    wl1 = H.Workload("refine64", n=n, seed_offset=21)
    wl2 = H.Workload("refine64", n=n, seed_offset=22)
    p = T.recon_params(wl1.cfg.ori_size, wl1.rmax)
    maps = {}
    for side, wl in (("a", wl1), ("b", wl2)):
        g = wl.run_gpu(dev=gpu_device)
        o = wl.run_oracle()
        (dg, wg), (do, wo) = g["bp"][0], o["bp"][0]
        shape = (p.pad_size, p.pad_size, p.pad_size // 2 + 1)
        maps["gpu" + side] = gpu_recon(gpu_device, p, dg.real.reshape(shape), dg.imag.reshape(shape),
                                       wg.reshape(shape))
        maps["cpu" + side] = orc_recon(p, do.real.reshape(shape), do.imag.reshape(shape), wo.reshape(shape))
        # the same reconstruction read straight from the device backprojector the E-step filled
        dm = gpu_recon_class(gpu_device, 0, p)
        print(f"\n[reconstruct_class {side}] rel vs host-array path {H.rel_norm(dm, maps['gpu' + side]):.2e}")
        assert H.rel_norm(dm, maps["gpu" + side]) < 1e-12
    f_gpu = gpu_fsc(gpu_device, maps["gpua"], maps["gpub"])
    f_cpu = orc_fsc(maps["cpua"], maps["cpub"])
    d = np.abs(f_gpu - f_cpu)
    print("\n[half-map FSC] max |FSC_gpu - FSC_cpu|", d.max(), "map rel", H.rel_norm(maps["gpua"], maps["cpua"]))
    assert d.max() <= 0.01


def test_reconstruct_class_from_backprojector(gpu_device):
    """t2x_reconstruct_class on a backprojector holding known volumes (written by a refine64
    E-step, MAP term on) equals t2x_reconstruct on the volumes t2x_backprojector_get returns;
    parameters that disagree with the backprojector and an uninitialised class are rejected."""
    lib = T.load()
    wl = H.Workload("refine64", n=8, seed_offset=5)
    g = wl.run_gpu(dev=gpu_device)
    p = T.recon_params(wl.cfg.ori_size, wl.rmax, do_map=True, tau2_fudge=1.0, minres_map=2)
    tau2 = np.random.default_rng(9).uniform(0.05, 1.0, wl.cfg.ori_size // 2 + 1)
    dg, wg = g["bp"][0]
    shape = (p.pad_size, p.pad_size, p.pad_size // 2 + 1)
    ref = gpu_recon(gpu_device, p, dg.real.reshape(shape), dg.imag.reshape(shape), wg.reshape(shape), tau2)
    out = gpu_recon_class(gpu_device, 0, p, tau2)
    assert H.rel_norm(out, ref) < 1e-12
    bad = T.recon_params(wl.cfg.ori_size, wl.rmax - 1)
    o = np.zeros((p.ori_size,) * 3)
    assert lib.t2x_reconstruct_class(gpu_device, 0, bad, None, T.dptr(o)) == T.T2X_EINVAL
    # a class no test initialises (the shared device holds classes of earlier tests)
    assert lib.t2x_reconstruct_class(gpu_device, 1 << 20, p, T.dptr(tau2), T.dptr(o)) == T.T2X_ESTATE
