"""t2x_projector_set_noalloc, the upload that allocates nothing (its planes pass through the
library's static device staging array), against t2x_projector_set: the device volumes must be
equal bit for bit on two calls in a row, the second with a smaller volume, and the shapes the
no-allocation call cannot serve are rejected with status codes."""
import ctypes as C

import numpy as np
import pytest

import t2x_py as T

pytestmark = pytest.mark.gpu

FP = C.POINTER(C.c_float)


def _fp(a):
    return a.ctypes.data_as(FP)


def _elems(ref_dim, pad):
    return pad * pad * (pad // 2 + 1) if ref_dim == 3 else pad * (pad // 2 + 1)


def _planes(seed, n):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)


def _get(lib, h, iclass, n):
    re, im = np.empty(n, np.float32), np.empty(n, np.float32)
    T.check(lib.t2x_debug_projector_get(h, iclass, _fp(re), _fp(im)), "t2x_debug_projector_get")
    return re, im


@pytest.fixture(scope="module")
def dev():
    lib = T.load()
    h = C.c_void_p()
    T.check(lib.t2x_device_open(0, C.byref(h)), "t2x_device_open")
    yield h
    lib.t2x_device_close(h)


def test_noalloc_equals_projector_set(dev):
    lib = T.load()
    # (class of the no-allocation upload, class of the reference upload, ref_dim, pad, r_max):
    # the 259^3 half volume (8.7 M elements) spans two staging chunks; the second call is smaller
    cases = [(40, 42, 3, 259, 64), (41, 43, 3, 131, 32)]
    for k, (ca, cr, ref_dim, pad, rmax) in enumerate(cases):
        n = _elems(ref_dim, pad)
        z = np.zeros(n, np.float32)
        T.check(lib.t2x_projector_set(dev, ca, ref_dim, pad, rmax, 2.0, _fp(z), _fp(z)), "set (allocate)")
    expect = {}
    for k, (ca, cr, ref_dim, pad, rmax) in enumerate(cases):
        n = _elems(ref_dim, pad)
        re, im = _planes(100 + k, n)
        print(f"no-alloc upload: class {ca}, ref_dim {ref_dim}, pad {pad}", flush=True)
        T.check(lib.t2x_projector_set_noalloc(dev, ca, ref_dim, pad, rmax, 2.0, _fp(re), _fp(im)), "noalloc")
        T.check(lib.t2x_projector_set(dev, cr, ref_dim, pad, rmax, 2.0, _fp(re), _fp(im)), "set (reference)")
        ga, gr = _get(lib, dev, ca, n), _get(lib, dev, cr, n)
        assert np.array_equal(ga[0], gr[0]) and np.array_equal(ga[1], gr[1])
        assert np.array_equal(ga[0], re) and np.array_equal(ga[1], im)
        expect[ca] = (re, im)
    # the second call reused the staging array: the first class's volume is untouched
    ca, _, ref_dim, pad, _ = cases[0]
    g = _get(lib, dev, ca, _elems(ref_dim, pad))
    assert np.array_equal(g[0], expect[ca][0]) and np.array_equal(g[1], expect[ca][1])


def test_noalloc_2d_plane_and_device_input(dev):
    lib = T.load()
    pad, rmax, n = 131, 32, _elems(2, 131)
    z = np.zeros(n, np.float32)
    T.check(lib.t2x_projector_set(dev, 44, 2, pad, rmax, 2.0, _fp(z), _fp(z)), "set (allocate)")
    re, im = _planes(7, n)
    # planes already in HBM
    dre, dim = C.c_void_p(), C.c_void_p()
    T.check(lib.t2x_device_alloc(dev, 4 * n, C.byref(dre)))
    T.check(lib.t2x_device_alloc(dev, 4 * n, C.byref(dim)))
    try:
        T.check(lib.t2x_memcpy_h2d(dev, dre, re.ctypes.data_as(C.c_void_p), 4 * n))
        T.check(lib.t2x_memcpy_h2d(dev, dim, im.ctypes.data_as(C.c_void_p), 4 * n))
        T.check(lib.t2x_projector_set_noalloc(dev, 44, 2, pad, rmax, 2.0, C.cast(dre, FP), C.cast(dim, FP)), "noalloc")
    finally:
        lib.t2x_device_free(dev, dre)
        lib.t2x_device_free(dev, dim)
    T.check(lib.t2x_projector_set(dev, 45, 2, pad, rmax, 2.0, _fp(re), _fp(im)), "set (reference)")
    ga, gr = _get(lib, dev, 44, n), _get(lib, dev, 45, n)
    assert np.array_equal(ga[0], gr[0]) and np.array_equal(ga[1], gr[1])


def test_noalloc_rejects_what_needs_an_allocation(dev):
    lib = T.load()
    n = _elems(3, 131)
    re, im = _planes(3, n)
    # no volume for this class yet
    assert lib.t2x_projector_set_noalloc(dev, 46, 3, 131, 32, 2.0, _fp(re), _fp(im)) == T.T2X_ESTATE
    assert b"allocation" in lib.t2x_last_error()
    # a volume of another shape
    T.check(lib.t2x_projector_set(dev, 46, 3, 131, 32, 2.0, _fp(re), _fp(im)), "set")
    r2, i2 = _planes(4, _elems(3, 67))
    assert lib.t2x_projector_set_noalloc(dev, 46, 3, 67, 16, 2.0, _fp(r2), _fp(i2)) == T.T2X_EINVAL
    assert lib.t2x_projector_set_noalloc(dev, 46, 2, 131, 32, 2.0, _fp(re), _fp(im)) == T.T2X_EINVAL
    assert lib.t2x_projector_set_noalloc(dev, 46, 3, 131, 32, 2.0, None, None) == T.T2X_EINVAL
