"""The C-ABI library without a GPU: it loads, exports every entry point include/*.h declares,
rejects bad arguments with status codes, and its host-side sampling grids equal the oracle's
bit for bit."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import t2x_py as T
from t2x_py import workloads as W

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(glob.glob(os.path.join(REPO, "include", "*.h"))):
        src = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(t2x_\w+)\s*\(", src, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = T.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert lib.t2x_abi_version() == 6


def test_bad_arguments_return_status():
    lib = T.load()
    assert lib.t2x_device_open(0, None) == T.T2X_EINVAL
    assert lib.t2x_projector_set(None, 0, 3, 21, 4, 2.0, None, None) == T.T2X_EINVAL
    assert lib.t2x_estep(None, None, None, 0, None, 0, None, None, None, None) == T.T2X_EINVAL
    assert b"null" in lib.t2x_last_error()
    cfg, _ = W.make_config(W.WORKLOADS["refine64"])
    cfg.coarse_size = cfg.current_size + 2
    info = T.SamplingInfo()
    assert lib.t2x_sampling_info_get(C.byref(cfg), C.byref(info)) == T.T2X_EINVAL
    cfg2, _ = W.make_config(W.WORKLOADS["refine64"])
    cfg2.orientational_prior_mode = 2
    assert lib.t2x_sampling_info_get(C.byref(cfg2), C.byref(info)) == T.T2X_EINVAL
    cfg2.orientational_prior_mode = T.T2X_PRIOR_ROTTILT_PSI
    cfg2.sigma_rot, cfg2.sigma_tilt = 10.0, 0.0   # rot-only prior: not supported
    assert lib.t2x_sampling_info_get(C.byref(cfg2), C.byref(info)) == T.T2X_EINVAL
    cfg2.sigma_tilt = 10.0
    assert lib.t2x_sampling_info_get(C.byref(cfg2), C.byref(info)) == 0


@pytest.mark.parametrize("name", ["refine64", "class256"])
def test_sampling_grids_match_oracle(orc, name):
    lib = T.load()
    cfg, _ = W.make_config(W.WORKLOADS[name])
    info = T.SamplingInfo()
    assert lib.t2x_sampling_info_get(C.byref(cfg), C.byref(info)) == 0
    oi = T.SamplingInfo()
    assert orc.orc_sampling_info(C.byref(cfg), C.byref(oi)) == 0
    assert [getattr(info, f) for f, _ in T.SamplingInfo._fields_] == \
        [getattr(oi, f) for f, _ in T.SamplingInfo._fields_]
    nd, npsi, nt = info.nr_dir, info.nr_psi, info.nr_trans
    g = [np.zeros(k) for k in (nd, nd, npsi, nt, nt)]
    o = [np.zeros(k) for k in (nd, nd, npsi, nt, nt)]
    assert lib.t2x_sampling_coarse_get(C.byref(cfg), *[T.dptr(a) for a in g]) == 0
    assert orc.orc_sampling_coarse(C.byref(cfg), *[T.dptr(a) for a in o]) == 0
    for a, b in zip(g, o):
        assert np.array_equal(a, b)
    sr, st, sp = np.zeros(nd * 4), np.zeros(nd * 4), np.zeros(npsi * 2)
    fx, fy = np.zeros(nt * 4), np.zeros(nt * 4)
    assert lib.t2x_sampling_fine_get(C.byref(cfg), T.dptr(sr), T.dptr(st), T.dptr(sp), T.dptr(fx), T.dptr(fy)) == 0
    r, t, p = np.zeros(8), np.zeros(8), np.zeros(8)
    for d in range(0, nd, max(1, nd // 40)):
        for ps in (0, npsi - 1):
            orc.orc_sampling_oversampled_orientations(C.byref(cfg), d, ps, 1, T.dptr(r), T.dptr(t), T.dptr(p))
            for k in range(8):
                assert r[k] == sr[d * 4 + k // 2] and t[k] == st[d * 4 + k // 2]
                assert p[k] == sp[ps * 2 + k % 2]
    x, y = np.zeros(4), np.zeros(4)
    for it in range(nt):
        orc.orc_sampling_oversampled_translations(C.byref(cfg), it, 1, T.dptr(x), T.dptr(y))
        assert np.array_equal(x, fx[it * 4:it * 4 + 4]) and np.array_equal(y, fy[it * 4:it * 4 + 4])
