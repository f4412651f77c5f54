"""ctypes bindings of the t2x C ABI (include/t2x.h).

The structs mirror include/t2x.h field by field. `load()` opens the in-tree libt2x.so and
raises if it is missing: there is no CPU fallback for the E-step.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(os.path.dirname(HERE))          # .../t2-relion_amd
REPO_ROOT = os.path.dirname(PKG_ROOT)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libt2x.so")

T2X_OK = 0
T2X_EINVAL = -1
T2X_ENOMEM = -2
T2X_EDEVICE = -3
T2X_ESTATE = -4
T2X_EZEROSUM = -5
T2X_ENOSIG = -6

T2X_PREC_FP32 = 0
T2X_PREC_BF16X3 = 1
T2X_PREC_FP16X3 = 2
T2X_PREC_FP16X2 = 3
T2X_NOPRIOR = 0
T2X_PRIOR_ROTTILT_PSI = 1
T2X_GATE_TOP1 = 0
T2X_GATE_TOPK_MASS = 1


class CtfParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "voltage_kv", "cs_mm", "q0", "defocus_u", "defocus_v", "defocus_angle",
        "bfactor", "phase_shift", "scale")]


class EstepConfig(C.Structure):
    _fields_ = [
        ("ori_size", C.c_int), ("pixel_size", C.c_double),
        ("current_size", C.c_int), ("coarse_size", C.c_int),
        ("nr_classes", C.c_int), ("ref_dim", C.c_int),
        ("healpix_order", C.c_int), ("psi_step", C.c_double),
        ("offset_range", C.c_double), ("offset_step", C.c_double),
        ("adaptive_oversampling", C.c_int), ("adaptive_fraction", C.c_double),
        ("maximum_significants", C.c_long), ("particle_diameter", C.c_double),
        ("width_mask_edge", C.c_int), ("do_zero_mask", C.c_int), ("random_seed", C.c_uint64),
        ("sigma2_fudge", C.c_double),
        ("do_norm_correction", C.c_int), ("avg_norm_correction", C.c_double),
        ("do_scale_correction", C.c_int), ("nr_groups", C.c_int), ("nr_optics_groups", C.c_int),
        ("do_ctf_correction", C.c_int), ("sigma2_offset", C.c_double),
        ("orientational_prior_mode", C.c_int), ("precision", C.c_int),
        ("proposal_topk", C.c_int), ("proposal_gate", C.c_int),
        ("proposal_min_conf", C.c_double), ("gemm_tile", C.c_int), ("gemm_group", C.c_int),
        ("sigma_rot", C.c_double), ("sigma_tilt", C.c_double), ("sigma_psi", C.c_double),
        ("launch_graph", C.c_int), ("bp_direct", C.c_int), ("coarse_exhaustive", C.c_int),
    ]


class Model(C.Structure):
    _fields_ = [("sigma2_noise", C.POINTER(C.c_double)), ("pdf_class", C.POINTER(C.c_double)),
                ("pdf_direction", C.POINTER(C.c_double)),
                ("prior_offset_class", C.POINTER(C.c_double)),
                ("scale_correction", C.POINTER(C.c_double)),
                ("data_vs_prior_class", C.POINTER(C.c_double))]


class ParticleIn(C.Structure):
    _fields_ = [("ctf", CtfParams),
                ("old_offset_x", C.c_double), ("old_offset_y", C.c_double),
                ("prior_offset_x", C.c_double), ("prior_offset_y", C.c_double),
                ("norm_correction", C.c_double),
                ("rot", C.c_double), ("tilt", C.c_double), ("psi", C.c_double),
                ("prior_rot", C.c_double), ("prior_tilt", C.c_double), ("prior_psi", C.c_double),
                ("group_id", C.c_int), ("optics_group", C.c_int), ("part_id", C.c_long)]


class ParticleOut(C.Structure):
    _fields_ = [("rot", C.c_double), ("tilt", C.c_double), ("psi", C.c_double),
                ("xoff", C.c_double), ("yoff", C.c_double),
                ("iclass", C.c_int), ("pad0", C.c_int),
                ("pmax", C.c_double), ("dLL", C.c_double), ("norm_correction", C.c_double),
                ("sum_weight", C.c_double), ("min_diff2", C.c_double),
                ("significant_weight", C.c_double),
                ("nr_significant_coarse", C.c_long), ("nr_fine_samples", C.c_long),
                ("nr_fine_significant", C.c_long), ("best_hidden", C.c_long)]


PD = C.POINTER(C.c_double)


class Wsum(C.Structure):
    _fields_ = [(n, PD) for n in (
        "sigma2_noise", "sumw_group", "pdf_class", "pdf_direction", "prior_offset_class",
        "sigma2_offset", "avg_norm_correction", "LL", "ave_Pmax",
        "wsum_signal_product", "wsum_reference_power")]


class EstepDebug(C.Structure):
    _fields_ = [("n_debug", C.c_int),
                ("coarse_diff2", C.POINTER(C.c_float)),
                ("coarse_significant", C.POINTER(C.c_ubyte)),
                ("fine_capacity", C.c_long),
                ("fine_hidden", C.POINTER(C.c_long)),
                ("fine_diff2", C.POINTER(C.c_float)),
                ("fine_count", C.POINTER(C.c_long)),
                ("coarse_min_diff2", PD), ("coarse_sum_weight", PD), ("coarse_sig_weight", PD),
                ("fine_weight", PD)]


class SamplingInfo(C.Structure):
    _fields_ = [(n, C.c_long) for n in (
        "nr_dir", "nr_psi", "nr_trans", "nr_over_rot", "nr_over_trans",
        "nr_coarse_per_class", "nr_hidden_coarse")]


class Stats(C.Structure):
    _fields_ = [("n_calls", C.c_int), ("n_particles", C.c_long)] + \
        [(n, C.c_double) for n in ("ms_preprocess", "ms_coarse_project", "ms_coarse_diff2",
                                   "ms_coarse_weights", "ms_fine_diff2", "ms_fine_weights",
                                   "ms_store", "ms_coarse_gemm")] + \
        [("n_select_fallbacks", C.c_long), ("n_coarse_diff2_launches", C.c_long)] + \
        [(n, C.c_double) for n in ("coarse_diff2_flop", "coarse_diff2_bytes", "fine_samples",
                                   "fine_flop", "bp_orientations", "sig_coarse")] + \
        [(n, C.c_int) for n in ("coarse_size", "current_size", "P_coarse", "P_fine")] + \
        [("ms_proposal", C.c_double), ("n_proposal_particles", C.c_long), ("n_proposal_fallback", C.c_long),
         ("proposal_bins_scored", C.c_double), ("gemm_tile_used", C.c_int), ("gemm_group_used", C.c_int),
                ("local_rows", C.c_double), ("local_slices", C.c_double),
                ("ms_graph", C.c_double), ("n_graph_launches", C.c_long),
                ("bp_brick_items", C.c_double), ("coarse_tiles_checked", C.c_double),
                ("coarse_tiles_pruned", C.c_double), ("coarse_bound_flop", C.c_double)]


class MrcInfo(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("mode", C.c_int),
                ("angpix", C.c_double), ("data_offset", C.c_long), ("swapped", C.c_int),
                ("dmin", C.c_double), ("dmax", C.c_double), ("dmean", C.c_double), ("rms", C.c_double)]


class ReconParams(C.Structure):
    _fields_ = [("ori_size", C.c_int), ("pad_size", C.c_int), ("r_max", C.c_int),
                ("padding_factor", C.c_double), ("ref_dim", C.c_int), ("data_dim", C.c_int),
                ("do_map", C.c_int), ("tau2_fudge", C.c_double), ("minres_map", C.c_int),
                ("skip_gridding", C.c_int), ("reserved", C.c_int * 6)]


def recon_params(ori_size, r_max, padding_factor=2.0, do_map=False, tau2_fudge=1.0, minres_map=0):
    """Parameters of BackProjector::reconstruct for a 3D reference from 2D images."""
    pad = 2 * (int(np.floor(padding_factor * r_max + 0.5)) + 1) + 1
    return ReconParams(ori_size=ori_size, pad_size=pad, r_max=r_max, padding_factor=padding_factor,
                       ref_dim=3, data_dim=2, do_map=int(do_map), tau2_fudge=tau2_fudge,
                       minres_map=minres_map, skip_gridding=1)


class ProposalTrainParams(C.Structure):
    _fields_ = [("steps", C.c_int), ("batch", C.c_int), ("n_projections", C.c_int), ("bf16", C.c_int),
                ("lr", C.c_double), ("shift_weight", C.c_double), ("max_shift_px", C.c_double),
                ("defocus_min", C.c_double), ("defocus_max", C.c_double), ("voltage", C.c_double),
                ("cs", C.c_double), ("q0", C.c_double), ("seed", C.c_uint64), ("reserved", C.c_int * 6)]


class ProposalTrainStats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_setup", C.c_double), ("loss_first", C.c_double),
                ("loss_last", C.c_double), ("top1_last", C.c_double), ("top16_last", C.c_double),
                ("steps", C.c_int),
                ("n_bins", C.c_int), ("reserved", C.c_int * 6)]


class Blob(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "z", "sigma", "amplitude")]


def dptr(a):
    """double* of a contiguous float64 numpy array (or NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PD)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


def make_wsum(K, n_dir, ori_size, n_optics_groups=1, n_groups=1):
    """numpy accumulators + a Wsum struct pointing at them."""
    arrs = {
        "sigma2_noise": np.zeros(n_optics_groups * (ori_size // 2 + 1)), "sumw_group": np.zeros(n_optics_groups),
        "pdf_class": np.zeros(K), "pdf_direction": np.zeros(K * n_dir),
        "prior_offset_class": np.zeros(2 * K), "sigma2_offset": np.zeros(1),
        "avg_norm_correction": np.zeros(1), "LL": np.zeros(1), "ave_Pmax": np.zeros(1),
        "wsum_signal_product": np.zeros(n_groups), "wsum_reference_power": np.zeros(n_groups)}
    ws = Wsum(**{k: dptr(v) for k, v in arrs.items()})
    return ws, arrs


def outs_to_dict(outs):
    names = [f[0] for f in ParticleOut._fields_ if f[0] != "pad0"]
    return {n: np.array([getattr(o, n) for o in outs]) for n in names}


_lib = None


def load(path=None):
    """Open libt2x.so; raise if it is not built (the E-step has no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libt2x.so not built at {p}: run __graft_entry__.build()")
    lib = C.CDLL(p)
    vp = C.c_void_p
    sig = {
        "t2x_abi_version": (C.c_int, []),
        "t2x_last_error": (C.c_char_p, []),
        "t2x_device_open": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "t2x_device_close": (C.c_int, [vp]),
        "t2x_projector_set": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "t2x_projector_set_noalloc": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                                C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "t2x_debug_projector_get": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "t2x_backprojector_init": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
        "t2x_backprojector_get": (C.c_int, [vp, C.c_int, PD, PD, PD]),
        "t2x_backprojector_clear": (C.c_int, [vp, C.c_int]),
        "t2x_sampling_info_get": (C.c_int, [C.POINTER(EstepConfig), C.POINTER(SamplingInfo)]),
        "t2x_sampling_coarse_get": (C.c_int, [C.POINTER(EstepConfig), PD, PD, PD, PD, PD]),
        "t2x_sampling_fine_get": (C.c_int, [C.POINTER(EstepConfig), PD, PD, PD, PD, PD]),
        "t2x_estep": (C.c_int, [vp, C.POINTER(EstepConfig), C.POINTER(Model), C.c_int, vp, C.c_int,
                                C.POINTER(ParticleIn), C.POINTER(ParticleOut), C.POINTER(Wsum),
                                C.POINTER(EstepDebug)]),
        "t2x_stats_get": (C.c_int, [vp, C.POINTER(Stats)]),
        "t2x_stats_reset": (C.c_int, [vp]),
        "t2x_device_alloc": (C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
        "t2x_device_free": (C.c_int, [vp, vp]),
        "t2x_memcpy_h2d": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "t2x_memcpy_d2h": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "t2x_device_synchronize": (C.c_int, [vp]),
        "t2x_device_copy": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "t2x_debug_copy": (C.c_int, [vp, C.c_int, C.c_size_t, C.c_size_t, C.POINTER(C.c_float)]),
        "t2x_proposal_init": (C.c_int, [vp, C.c_int, C.c_uint64]),
        "t2x_proposal_num_params": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
        "t2x_proposal_get_params": (C.c_int, [vp, C.POINTER(C.c_float), C.c_longlong]),
        "t2x_proposal_set_params": (C.c_int, [vp, C.POINTER(C.c_float), C.c_longlong]),
        "t2x_proposal_mark_trained": (C.c_int, [vp, C.POINTER(EstepConfig)]),
        "t2x_proposal_gate_values": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "t2x_proposal_loss_grad": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int),
                                             C.POINTER(C.c_float), C.c_double, C.c_int, PD,
                                             C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "t2x_proposal_train": (C.c_int, [vp, C.POINTER(EstepConfig), C.POINTER(Model),
                                         C.POINTER(ProposalTrainParams), C.POINTER(ProposalTrainStats)]),
        "t2x_reconstruct": (C.c_int, [vp, C.POINTER(ReconParams), PD, PD, PD, PD, PD]),
        "t2x_reconstruct_class": (C.c_int, [vp, C.c_int, C.POINTER(ReconParams), PD, PD]),
        "t2x_mrc_info_read": (C.c_int, [C.c_char_p, C.POINTER(MrcInfo)]),
        "t2x_mrc_read_sections": (C.c_int, [C.c_char_p, C.c_long, C.c_long, C.POINTER(C.c_float)]),
        "t2x_mrc_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_float)]),
        "t2x_star_read": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "t2x_star_free": (C.c_int, [vp]),
        "t2x_star_write": (C.c_int, [vp, C.c_char_p]),
        "t2x_star_num_rows": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_long)]),
        "t2x_star_get_double": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_long, PD]),
        "t2x_star_get_string": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_long, C.c_char_p, C.c_long]),
        "t2x_star_set_double": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_long, C.c_double]),
        "t2x_star_set_int": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_long, C.c_long]),
        "t2x_star_particles": (C.c_int, [vp, C.c_long, C.c_long, C.POINTER(ParticleIn), PD]),
        "t2x_star_image_ref": (C.c_int, [vp, C.c_long, C.POINTER(C.c_long), C.c_char_p, C.c_long]),
        "t2x_star_store_results": (C.c_int, [vp, C.c_long, C.c_long, C.POINTER(ParticleOut), C.c_int]),
        "t2x_fsc": (C.c_int, [vp, C.c_int, PD, PD, PD]),
        "t2x_synth_reference_device": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, C.c_int,
                                                 C.c_double, C.POINTER(Blob), C.c_int, C.c_double,
                                                 vp, vp]),
        "t2x_synth_particles_device": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, C.POINTER(Blob),
                                                 C.POINTER(C.c_int), C.POINTER(ParticleIn), PD, PD,
                                                 PD, PD, PD, C.POINTER(C.c_int), PD, C.c_uint64,
                                                 vp, PD]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status, what="t2x call"):
    if status != T2X_OK:
        msg = load().t2x_last_error()
        raise RuntimeError(f"{what} failed ({status}): {msg.decode() if msg else ''}")
