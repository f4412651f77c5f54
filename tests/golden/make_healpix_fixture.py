"""Writes healpix_ref.npz: pixel centres (z, phi) and face coordinates of HEALPix NEST pixels
for orders 0..4, computed by the reference's own Healpix_Base (src/Healpix_2.15a, compiled by
oracle/Makefile into oracle/_ref/libhealpix_ref.so). Run from the repository root."""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle_py  # noqa: E402

lib = oracle_py.load_ref_healpix()
if lib is None:
    sys.exit("oracle/_ref/libhealpix_ref.so missing: run make -C oracle")
rows = []
for order in range(0, 5):
    for p in range(12 << (2 * order)):
        z, phi = C.c_double(), C.c_double()
        x, y, f = C.c_int(), C.c_int(), C.c_int()
        lib.ref_pix2ang_z_phi(order, p, C.byref(z), C.byref(phi))
        lib.ref_nest2xyf(order, p, C.byref(x), C.byref(y), C.byref(f))
        rows.append((order, p, z.value, phi.value, x.value, y.value, f.value))
a = np.array(rows)
np.savez_compressed(os.path.join(HERE, "healpix_ref.npz"), order=a[:, 0].astype(np.int32),
                    pix=a[:, 1].astype(np.int64), z=a[:, 2], phi=a[:, 3],
                    ix=a[:, 4].astype(np.int32), iy=a[:, 5].astype(np.int32),
                    face=a[:, 6].astype(np.int32))
print("wrote", len(rows), "pixels")
