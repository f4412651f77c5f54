"""ctypes loader of oracle/liboracle.so — TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker and
CPU baseline. The product path never imports this module.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "t2-relion_amd", "python"))
import t2x_py as T  # noqa: E402

LIB = os.path.join(HERE, "liboracle.so")
REF_HEALPIX = os.path.join(HERE, "_ref", "libhealpix_ref.so")

ORC_PROJ_CLEAN = 0
ORC_PROJ_STALE = 1


class Volume(C.Structure):
    _fields_ = [("ref_dim", C.c_int), ("pad_size", C.c_int), ("r_max", C.c_int),
                ("padding_factor", C.c_double), ("data", T.PD), ("weight", T.PD)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} not built: run `make -C oracle`")
    lib = C.CDLL(LIB)
    PD, PI = T.PD, C.POINTER(C.c_int)
    sig = {
        "orc_healpix_pix2ang_z_phi": (None, [C.c_int, C.c_long, PD, PD]),
        "orc_healpix_nest2xyf": (None, [C.c_int, C.c_long, PI, PI, PI]),
        "orc_healpix_xyf2nest": (C.c_long, [C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_sampling_info": (C.c_int, [C.POINTER(T.EstepConfig), C.POINTER(T.SamplingInfo)]),
        "orc_sampling_coarse": (C.c_int, [C.POINTER(T.EstepConfig), PD, PD, PD, PD, PD]),
        "orc_sampling_oversampled_orientations": (C.c_int, [C.POINTER(T.EstepConfig), C.c_long,
                                                            C.c_long, C.c_int, PD, PD, PD]),
        "orc_sampling_oversampled_translations": (C.c_int, [C.POINTER(T.EstepConfig), C.c_long,
                                                            C.c_int, PD, PD]),
        "orc_local_priors": (C.c_int, [C.POINTER(T.EstepConfig), C.POINTER(T.ParticleIn), PD, PD,
                                       C.POINTER(C.c_long), C.POINTER(C.c_long)]),
        "orc_euler_angles2matrix": (None, [C.c_double, C.c_double, C.c_double, PD]),
        "orc_ctf_value": (C.c_double, [C.POINTER(T.CtfParams), C.c_double, C.c_double]),
        "orc_ctf_image": (None, [C.POINTER(T.CtfParams), C.c_int, C.c_double, C.c_int, PD]),
        "orc_fft2_r2c": (None, [C.c_int, PD, PD]),
        "orc_fft2_c2r": (None, [C.c_int, PD, PD]),
        "orc_project": (None, [C.POINTER(Volume), PD, C.c_int, PD]),
        "orc_backproject": (None, [C.POINTER(Volume), PD, PD, C.c_int, PD]),
        "orc_synth_reference": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                          C.POINTER(T.Blob), C.c_int, C.c_double, PD]),
        "orc_synth_reference_plane": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                                C.POINTER(T.Blob), C.c_int, C.c_double, PD]),
        "orc_synth_particles": (C.c_int, [C.c_int, C.c_double, C.c_int, C.POINTER(T.Blob), PI,
                                          C.POINTER(T.ParticleIn), PD, PD, PD, PD, PD, PI, PD,
                                          C.c_uint64, C.POINTER(C.c_float), PD]),
        "orc_noise_sample": (C.c_double, [C.c_uint64, C.c_long, C.c_long]),
        "orc_estep": (C.c_int, [C.POINTER(T.EstepConfig), C.POINTER(T.Model), C.POINTER(Volume),
                                C.POINTER(Volume), C.c_int, C.POINTER(C.c_float),
                                C.POINTER(T.ParticleIn), C.POINTER(T.ParticleOut),
                                C.POINTER(T.Wsum), C.POINTER(T.EstepDebug), C.c_int, C.c_int]),
        "orc_preprocess": (C.c_int, [C.POINTER(T.EstepConfig), C.POINTER(C.c_float),
                                     C.POINTER(T.ParticleIn), PD, PD, PD, PD, PD, PD, PD]),
        "orc_background_noise": (C.c_int, [C.POINTER(T.EstepConfig), PD, C.c_long, PD]),
        "orc_reconstruct": (C.c_int, [C.POINTER(T.ReconParams), PD, PD, PD, PD, PD]),
        "orc_fsc": (C.c_int, [C.c_int, PD, PD, PD]),
        "orc_weights_from_diff2": (C.c_int, [C.c_long, PD, PD, C.c_double, C.c_double, C.c_long,
                                             C.c_int, PD, PD, PD, C.POINTER(C.c_long)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def load_ref_healpix():
    """The reference's own HEALPix sources built by oracle/Makefile, or None if absent."""
    if not os.path.exists(REF_HEALPIX):
        return None
    lib = C.CDLL(REF_HEALPIX)
    PD, PI = T.PD, C.POINTER(C.c_int)
    lib.ref_pix2ang_z_phi.argtypes = [C.c_int, C.c_long, PD, PD]
    lib.ref_nest2xyf.argtypes = [C.c_int, C.c_long, PI, PI, PI]
    lib.ref_xyf2nest.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.ref_xyf2nest.restype = C.c_long
    return lib


def make_volume(ref_dim, pad, r_max, pf, with_weight):
    xdim = pad // 2 + 1
    n = pad * pad * xdim if ref_dim == 3 else pad * xdim
    data = np.zeros(2 * n)
    weight = np.zeros(n) if with_weight else None
    v = Volume(ref_dim, pad, r_max, pf, T.dptr(data), T.dptr(weight))
    return v, data, weight
