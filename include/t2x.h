/*
 * t2x.h — C ABI of the MI355X-native E-step of RELION's ML optimiser.
 *
 * The library (libt2x.so) replaces the accelerated expectation step of relion_refine:
 * the per-thread E-step driver (MlOptimiserCuda), the device bundle (MlDeviceBundle),
 * the projector textures (AccProjector) and the backprojector volumes (AccBackprojector).
 * Every entry point takes plain pointers and sizes, returns an integer status
 * (T2X_OK or a negative T2X_E* code) and never throws; t2x_last_error() returns a
 * thread-local description of the last failure.
 *
 * Conventions shared with relion_refine's CPU path:
 *   - Fourier images use the FFTW half layout: size x (size/2+1), row i holds
 *     frequency ip = (i <= size/2) ? i : i - size, column j holds jp = j.
 *   - Forward FFTs are normalised by 1/(number of real pixels) (fftw.cpp:335-358).
 *   - Real-space images are N x N, row-major, pixel (i,j) has logical
 *     coordinates (i - N/2, j - N/2) (Xmipp origin).
 *   - Projector / backprojector volumes are pad x pad x (pad/2+1) arrays,
 *     x from 0, y and z centred: index = (z + pad/2)*pad*(pad/2+1) + (y + pad/2)*(pad/2+1) + x
 *     (projector.cpp:55-85).
 *   - Offsets in the particle metadata are in pixels, the translational sampling
 *     grid is in Angstrom (ml_optimiser.cpp:343-350), sigma2_offset is in A^2.
 */
#ifndef T2X_H
#define T2X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define T2X_ABI_VERSION 6

enum {
    T2X_OK = 0,
    T2X_EINVAL = -1,      /* bad argument or unsupported configuration            */
    T2X_ENOMEM = -2,      /* device or host allocation failed                     */
    T2X_EDEVICE = -3,     /* HIP runtime / kernel launch failure                  */
    T2X_ESTATE = -4,      /* call out of order (e.g. projector not set)           */
    T2X_EZEROSUM = -5,    /* zero or NaN sum of weights (ml_optimiser.cpp:7834)   */
    T2X_ENOSIG = -6       /* no significant coarse sample (ml_optimiser.cpp:7986) */
};

/* Which arithmetic the coarse/fine cross terms use. */
enum {
    T2X_PREC_FP32 = 0,        /* plain FP32 products, FP32 accumulation                */
    T2X_PREC_BF16X3 = 1,      /* BF16 hi+lo split, 3 MFMA products, FP32 accumulation  */
    T2X_PREC_FP16X3 = 2,      /* FP16 hi+lo split, 3 MFMA products, FP32 accumulation  */
    T2X_PREC_FP16X2 = 3       /* the dense coarse cross term as FP16 hi+lo image rows x  */
                              /* FP16 reference rows (2 MFMA products, FP32 accumulation); */
                              /* every other cross term as FP16X3                          */
};

/* Learned-proposal gate (t2x_estep_config.proposal_gate): the value compared with
 * proposal_min_conf to decide between the top-k restricted and the exhaustive coarse search. */
enum {
    T2X_GATE_TOP1 = 0,        /* softmax probability of the most probable bin            */
    T2X_GATE_TOPK_MASS = 1    /* softmax probability mass of the top-k bins              */
};

/* Orientational prior modes (ml_model.h NOPRIOR / PRIOR_ROTTILT_PSI). */
enum { T2X_NOPRIOR = 0, T2X_PRIOR_ROTTILT_PSI = 1 };

/* CTF parameters of one image (ctf.cpp:150-260, CTF::setValuesByGroup + initialise). */
typedef struct t2x_ctf_params {
    double voltage_kv;      /* acceleration voltage (kV)                 */
    double cs_mm;           /* spherical aberration (mm)                 */
    double q0;              /* amplitude contrast                        */
    double defocus_u;       /* A                                         */
    double defocus_v;       /* A                                         */
    double defocus_angle;   /* degrees                                   */
    double bfactor;         /* A^2                                       */
    double phase_shift;     /* degrees                                   */
    double scale;           /* CTF scale factor                          */
} t2x_ctf_params;

/* E-step configuration: the MlOptimiser / HealpixSampling state the E-step reads. */
typedef struct t2x_estep_config {
    int    ori_size;               /* box size N of the particle images                  */
    double pixel_size;             /* A/px                                               */
    int    current_size;           /* image_current_size (even)                          */
    int    coarse_size;            /* image_coarse_size (even, <= current_size)          */
    int    nr_classes;             /* K                                                  */
    int    ref_dim;                /* 3: 3D references; 2: 2D class averages             */
    int    healpix_order;          /* coarse HEALPix order (ref_dim 3)                   */
    double psi_step;               /* coarse in-plane step in degrees                    */
    double offset_range;           /* A                                                  */
    double offset_step;            /* A                                                  */
    int    adaptive_oversampling;  /* 0 or 1                                             */
    double adaptive_fraction;      /* 0.999 by default                                   */
    long   maximum_significants;   /* <= 0: unlimited                                    */
    double particle_diameter;      /* A, circular soft mask                              */
    int    width_mask_edge;        /* px                                                 */
    int    do_zero_mask;           /* 1 (--zero_mask): the mask's background is its      */
                                   /* average value; 0 (relion_refine's default): noise  */
                                   /* with the optics group's sigma2_noise spectrum,     */
                                   /* drawn per particle from (random_seed, part_id)     */
    uint64_t random_seed;          /* seed of that background noise                      */
    double sigma2_fudge;           /* 1                                                  */
    int    do_norm_correction;     /* 0/1                                                */
    double avg_norm_correction;    /* mymodel.avg_norm_correction                        */
    int    do_scale_correction;    /* 0/1 (--scale): references scaled per image group   */
    int    nr_groups;              /* image groups (scale correction); 0 means 1         */
    int    nr_optics_groups;       /* optics groups (sigma2_noise); 0 means 1. Every     */
                                   /* group has the box ori_size and pixel_size          */
    int    do_ctf_correction;      /* 0/1                                                */
    double sigma2_offset;          /* mymodel.sigma2_offset (A^2)                        */
    int    orientational_prior_mode; /* T2X_NOPRIOR: global search over pdf_direction;   */
                                   /* T2X_PRIOR_ROTTILT_PSI: local searches around each  */
                                   /* particle's prior angles (sigma_rot/tilt/psi)       */
    int    precision;              /* T2X_PREC_* for the cross terms                     */
    int    proposal_topk;          /* > 0: coarse pass scores only the top-k bins of the */
                                   /* trained proposal network (t2x_proposal_train)      */
    int    proposal_gate;          /* T2X_GATE_*: what proposal_min_conf is compared with */
    double proposal_min_conf;      /* gate value below this: exhaustive coarse search    */
    int    gemm_tile;              /* split-precision GEMM variant: 0 autotuned (fastest  */
                                   /* on first use per shape), 128, 256, 257, 258, or    */
                                   /* 259 (FP16X2 only: 128x256 tiles, 2 per CU)         */
    int    gemm_group;             /* its tile-walk group in tile rows (0: default)      */
    /* local searches: sqrt(mymodel.sigma2_rot / sigma2_tilt / sigma2_psi) in degrees;
     * sigma_rot and sigma_tilt both > 0 (a Gaussian on the angle between the directions) or
     * both 0 (every direction); sigma_psi 0: every psi angle */
    double sigma_rot, sigma_tilt, sigma_psi;
    int    launch_graph;           /* 1: from the second batch of a size on, each batch's */
                                   /* coarse pass, selection, fine pass and store (all    */
                                   /* launches after the preprocessing) run as one        */
                                   /* captured hipGraph, replayed for the call's further  */
                                   /* batches of that size; their stages are then timed   */
                                   /* together (t2x_stats.ms_graph). Global searches      */
                                   /* without the proposal and without diagnostics only;  */
                                   /* elsewhere ignored                                   */
    int    bp_direct;              /* 1: every pixel inserted into the backprojector with */
                                   /* global atomics; 0: through LDS bricks (16^3 voxels */
                                   /* of a 3D volume, 32^2 of a 2D one), as many of a    */
                                   /* batch's items as the brick buffers hold, the rest  */
                                   /* directly (t2x_stats.bp_brick_items)                */
    int    coarse_exhaustive;      /* 1: the coarse GEMM computes every tile. 0 (global  */
                                   /* searches in a split precision, without the        */
                                   /* proposal or diagnostics): a first GEMM over the   */
                                   /* lowest frequencies bounds each block's diff2 from */
                                   /* below (every pixel adds a non-negative term), and */
                                   /* the full GEMM skips the tiles whose blocks all lie */
                                   /* beyond the negligible-weight cut above a computed */
                                   /* value; the weights are those of the exhaustive    */
                                   /* pass (t2x_stats.coarse_tiles_*)                    */
} t2x_estep_config;

/* Model arrays of the current iteration (host, caller-owned, read-only). */
typedef struct t2x_model {
    const double* sigma2_noise;    /* [nr_optics_groups][ori_size/2+1]                   */
                                   /*                            mymodel.sigma2_noise    */
    const double* pdf_class;       /* [K]                        mymodel.pdf_class       */
    const double* pdf_direction;   /* [K * n_dir]                mymodel.pdf_direction   */
    const double* prior_offset_class; /* [2K] x,y per class for 2D refs (may differ per   */
                                   /* class), compared with old offset + translation as  */
                                   /* ml_optimiser.cpp:7627-7653 does; may be NULL       */
    const double* scale_correction;   /* [nr_groups] mymodel.scale_correction (used with   */
                                   /* do_scale_correction; each in (0, 10000])           */
    const double* data_vs_prior_class; /* [K][ori_size/2+1] mymodel.data_vs_prior_class: */
                                   /* the scale sums run over shells where it is > 3     */
                                   /* (do_scale_correction; NULL: no shell qualifies)    */
} t2x_model;

/* Per-particle input metadata (one image per particle). */
typedef struct t2x_particle_in {
    t2x_ctf_params ctf;
    double old_offset_x, old_offset_y;       /* px (METADATA_XOFF/YOFF)                    */
    double prior_offset_x, prior_offset_y;   /* px, 999 = unset (METADATA_XOFF_PRIOR)     */
    double norm_correction;                  /* METADATA_NORM                              */
    double rot, tilt, psi;                   /* current angles (METADATA_ROT/TILT/PSI)     */
    double prior_rot, prior_tilt, prior_psi; /* METADATA_ROT/TILT/PSI_PRIOR, 999 = unset:  */
                                             /* the local search centres on rot/tilt/psi  */
    int    group_id;                         /* image group, 0-based (getGroupId)          */
    int    optics_group;                     /* optics group, 0-based (getOpticsGroup)     */
    long   part_id;                          /* particle number in the data set: the key   */
                                             /* of its mask-background noise               */
} t2x_particle_in;

/* Per-particle results (the exp_metadata row written by storeWeightedSums). */
typedef struct t2x_particle_out {
    double rot, tilt, psi;        /* best pose (degrees)                                  */
    double xoff, yoff;            /* best offset (px) = rounded old offset + translation  */
    int    iclass;                /* best class, 1-based                                  */
    int    pad0;
    double pmax;                  /* METADATA_PMAX                                        */
    double dLL;                   /* METADATA_DLL                                         */
    double norm_correction;       /* METADATA_NORM after the update                       */
    double sum_weight;            /* exp_sum_weight after the fine pass                   */
    double min_diff2;             /* exp_min_diff2 after the fine pass                    */
    double significant_weight;    /* exp_significant_weight after the fine pass           */
    long   nr_significant_coarse; /* METADATA_NR_SIGN                                     */
    long   nr_fine_samples;       /* fine (orientation, translation) pairs scored         */
    long   nr_fine_significant;   /* fine samples with weight >= significant_weight       */
    long   best_hidden;           /* fine hidden index (ihidden_over) of the best pose    */
} t2x_particle_out;

/* Weighted sums the E-step adds into (host, caller-owned, double). Any member may be NULL. */
typedef struct t2x_wsum {
    double* sigma2_noise;         /* [nr_optics_groups][ori_size/2+1] wsum_model.sigma2_noise */
    double* sumw_group;           /* [nr_optics_groups] wsum_model.sumw_group             */
    double* pdf_class;            /* [K]                                                  */
    double* pdf_direction;        /* [K * n_dir]                                          */
    double* prior_offset_class;   /* [2K]  (2D references only)                           */
    double* sigma2_offset;        /* [1]                                                  */
    double* avg_norm_correction;  /* [1]                                                  */
    double* LL;                   /* [1]                                                  */
    double* ave_Pmax;             /* [1]                                                  */
    double* wsum_signal_product;  /* [nr_groups] XA sums of the scale correction          */
    double* wsum_reference_power; /* [nr_groups] AA sums of the scale correction          */
} t2x_wsum;

/* Optional inspection buffers for parity tests (host, caller-owned; NULL = skip).
 * Only the first n_debug particles of a call are reported. */
typedef struct t2x_estep_debug {
    int     n_debug;
    float*  coarse_diff2;     /* [n_debug][K*n_dir*n_psi*n_trans] ihidden order; -999 where not computed */
    unsigned char* coarse_significant; /* [n_debug][same] Mcoarse_significant                          */
    long    fine_capacity;    /* capacity (entries) per particle of the two arrays below                 */
    long*   fine_hidden;      /* [n_debug][fine_capacity] ihidden_over of each fine sample                */
    float*  fine_diff2;       /* [n_debug][fine_capacity] fine diff2                                      */
    long*   fine_count;       /* [n_debug] number of fine samples                                         */
    double* coarse_min_diff2; /* [n_debug]                                                                */
    double* coarse_sum_weight;/* [n_debug]                                                                */
    double* coarse_sig_weight;/* [n_debug]                                                                */
    double* fine_weight;      /* [n_debug][fine_capacity] fine weights (the CPU oracle fills it; libt2x  */
                              /* leaves it untouched)                                                    */
} t2x_estep_debug;

/* Sampling sizes implied by a configuration (HealpixSampling::NrDirections etc.). */
typedef struct t2x_sampling_info {
    long nr_dir, nr_psi, nr_trans;          /* coarse                                 */
    long nr_over_rot, nr_over_trans;        /* oversampling factors of the fine pass  */
    long nr_coarse_per_class;               /* nr_dir * nr_psi                        */
    long nr_hidden_coarse;                  /* K * nr_dir * nr_psi * nr_trans         */
} t2x_sampling_info;

/* Per-stage timing and work counters accumulated over t2x_estep calls (for reporting). */
typedef struct t2x_stats {
    int    n_calls;
    long   n_particles;
    double ms_preprocess, ms_coarse_project, ms_coarse_diff2, ms_coarse_weights;
    double ms_fine_diff2, ms_fine_weights, ms_store;     /* GPU event time per stage          */
    double ms_coarse_gemm;       /* the coarse cross-term GEMM alone (HIP events on its stream) */
    long   n_select_fallbacks;   /* batches whose coarse selection used the radix-select path   */
    long   n_coarse_diff2_launches;
    double coarse_diff2_flop;    /* algorithmic 4*T*O*P per particle, summed                  */
    double coarse_diff2_bytes;   /* algorithmic HBM bytes of the coarse GEMM                  */
    double fine_samples;         /* fine (o,t) pairs scored                                   */
    double fine_flop;            /* algorithmic 4*S*P_f                                       */
    double bp_orientations;      /* orientations backprojected                                */
    double sig_coarse;           /* significant coarse samples                                */
    int    coarse_size, current_size, P_coarse, P_fine;
    double ms_proposal;          /* network inference + restricted work lists (learned path)  */
    long   n_proposal_particles; /* particles that went through the proposal network         */
    long   n_proposal_fallback;  /* of those, searched exhaustively (low confidence)          */
    double proposal_bins_scored; /* sum over particles of the bins scored in the coarse pass  */
    int    gemm_tile_used;       /* split-precision GEMM variant of the last launch (tile, group) */
    int    gemm_group_used;
    double local_rows;           /* local searches: (particle, orientation) rows of the coarse pass */
    double local_slices;         /* of those, distinct slices projected (shared within a batch)    */
    double ms_graph;             /* launch_graph: GPU time of the batches run as a graph           */
    long   n_graph_launches;     /* batches run as a graph (captured, or replayed)                 */
    double bp_brick_items;       /* of bp_orientations, those inserted through LDS bricks          */
    double coarse_tiles_checked; /* coarse GEMM tiles tested against the low-frequency bounds      */
    double coarse_tiles_pruned;  /* of those, skipped (every block beyond the cut)                 */
    double coarse_bound_flop;    /* the bounds GEMMs' work, 4 flops per (row, orientation, pixel)  */
} t2x_stats;

/* ---------------------------------------------------------------------------------------
 * Device bundle (MlDeviceBundle, src/acc/cuda/cuda_ml_optimiser.h:18-81)
 * ------------------------------------------------------------------------------------- */
typedef struct t2x_device t2x_device;

/* Replaces MlDeviceBundle::MlDeviceBundle + setDevice (cuda_ml_optimiser.cu:1153-1169). */
int t2x_device_open(int device_id, t2x_device** out);
/* Replaces the bundle teardown at the end of MlOptimiser::expectation (ml_optimiser.cpp:3505-3543). */
int t2x_device_close(t2x_device* dev);
/* Thread-local text of the last error (CRITICAL/RelionError text in cuda_settings.h:50-84). */
const char* t2x_last_error(void);
int t2x_abi_version(void);

/* ---------------------------------------------------------------------------------------
 * Projector (AccProjector, src/acc/acc_projector.h:14-87)
 * Replaces AccProjector::setMdlDim + initMdl (acc_projector_impl.h:10-223) for class iclass:
 * re/im are pad*pad*(pad/2+1) (3D) or pad*(pad/2+1) (2D) floats of PPref[iclass].data.
 * --------------------------------------------------------------------------------------- */
int t2x_projector_set(t2x_device* dev, int iclass, int ref_dim, int pad_size, int r_max,
                      double padding_factor, const float* re, const float* im);

/* No counterpart in the reference. t2x_projector_set for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * it calls no hipMalloc, hipFree or allocating library routine. Class iclass must already hold a
 * projector volume of the same ref_dim and pad_size (an earlier t2x_projector_set allocated it;
 * otherwise T2X_ESTATE, or T2X_EINVAL for another shape). The re/im planes (host or device
 * memory) are staged through a static device array of 2 x 8 Mi floats (64 MiB, held by each GPU
 * the library runs on until the process ends) in chunks, so every volume size is accepted.
 * Every call on one GPU shares that array: calls that overlap on one GPU (several handles or
 * threads) race. The resulting volume equals t2x_projector_set's bit for bit. */
int t2x_projector_set_noalloc(t2x_device* dev, int iclass, int ref_dim, int pad_size, int r_max,
                              double padding_factor, const float* re, const float* im);
/* Diagnostic copy of class iclass's projector volume in the layout t2x_projector_set takes
 * (tests; no reference counterpart). */
int t2x_debug_projector_get(t2x_device* dev, int iclass, float* re, float* im);

/* ---------------------------------------------------------------------------------------
 * Backprojector (AccBackprojector, src/acc/acc_backprojector.h:15-87)
 * ------------------------------------------------------------------------------------- */
/* Replaces AccBackprojector::setMdlDim + initMdl (acc_backprojector_impl.h:8-60): zeroed volumes. */
int t2x_backprojector_init(t2x_device* dev, int iclass, int ref_dim, int pad_size, int r_max,
                           double padding_factor);
/* Replaces AccBackprojector::getMdlData (acc_backprojector_impl.h:84-99): device volumes -> host doubles. */
int t2x_backprojector_get(t2x_device* dev, int iclass, double* re, double* im, double* weight);
/* Replaces AccBackprojector::clear (acc_backprojector_impl.h:101-143): zero the volumes. */
int t2x_backprojector_clear(t2x_device* dev, int iclass);

/* ---------------------------------------------------------------------------------------
 * Sampling helpers (HealpixSampling, src/healpix_sampling.cpp)
 * ------------------------------------------------------------------------------------- */
int t2x_sampling_info_get(const t2x_estep_config* cfg, t2x_sampling_info* info);
/* Coarse grid (HealpixSampling rot_angles/tilt_angles/psi_angles/translations_x/y,
 * healpix_sampling.cpp:290-608): rot/tilt [nr_dir], psi [nr_psi], tx/ty in A [nr_trans]. */
int t2x_sampling_coarse_get(const t2x_estep_config* cfg, double* rot, double* tilt, double* psi,
                            double* tx, double* ty);
/* Fine grid (getOrientations / getTranslationsInPixel with oversampling 1,
 * healpix_sampling.cpp:1800-2050): sub_rot/sub_tilt [nr_dir*4] (NEST children, y outer),
 * sub_psi [nr_psi*2], ftx/fty in px [nr_trans*4]. */
int t2x_sampling_fine_get(const t2x_estep_config* cfg, double* sub_rot, double* sub_tilt,
                          double* sub_psi, double* ftx_px, double* fty_px);

/* ---------------------------------------------------------------------------------------
 * E-step (MlOptimiserCuda::doThreadExpectationSomeParticles, cuda_ml_optimiser.cu:1345-1408,
 * and its CPU twin MlOptimiser::expectationOneParticle, ml_optimiser.cpp:4252-4608).
 *
 * Runs getFourierTransformsAndCtfs, both sampling passes (getAllSquaredDifferences +
 * convertAllSquaredDifferencesToWeights) and storeWeightedSums for n particles.
 * images: n * ori_size^2 floats, host memory (images_on_device = 0) or device memory (1).
 * Backprojected slices are added into the device backprojector of each class; the other
 * weighted sums are added into *wsum; per-particle metadata goes to out[0..n).
 * --------------------------------------------------------------------------------------- */
int t2x_estep(t2x_device* dev, const t2x_estep_config* cfg, const t2x_model* model,
              int n, const float* images, int images_on_device,
              const t2x_particle_in* pin, t2x_particle_out* out,
              t2x_wsum* wsum, t2x_estep_debug* dbg);

/* Accumulated timing / work counters since the last reset. */
int t2x_stats_get(t2x_device* dev, t2x_stats* st);
int t2x_stats_reset(t2x_device* dev);

/* ---------------------------------------------------------------------------------------
 * Device memory helpers for callers that keep inputs resident in HBM.
 * ------------------------------------------------------------------------------------- */
int t2x_device_alloc(t2x_device* dev, size_t bytes, void** dptr);
int t2x_device_free(t2x_device* dev, void* dptr);
int t2x_memcpy_h2d(t2x_device* dev, void* dst, const void* src, size_t bytes);
int t2x_memcpy_d2h(t2x_device* dev, void* dst, const void* src, size_t bytes);
int t2x_device_synchronize(t2x_device* dev);
/* Device-to-device copy by a plain vector-load/store kernel on the handle's stream (bytes a
 * multiple of 16, 16-byte aligned pointers). Its HBM counters read exactly 2*bytes of traffic,
 * which calibrates per-dispatch FETCH_SIZE / WRITE_SIZE figures. No reference counterpart. */
int t2x_device_copy(t2x_device* dev, void* dst, const void* src, size_t bytes);

/* Diagnostic copy-out of internal E-step buffers for parity tests (no reference counterpart):
 * which = 0 coarse slices (re/im), 1 coarse |F|^2, 2 N1, 3 image-side operand, 4 A1, 5 coarse diff2. */
int t2x_debug_copy(t2x_device* dev, int which, size_t offset, size_t count, float* dst);

/* ---------------------------------------------------------------------------------------
 * Synthetic workloads (Gaussian-blob volumes, CTF-modulated noisy projections) synthesised on
 * the GPU from caller-supplied blob lists and per-particle parameters.
 * ------------------------------------------------------------------------------------- */
typedef struct t2x_blob { double x, y, z, sigma, amplitude; } t2x_blob; /* px, centred */

/* Padded Fourier reference of a blob volume (analytic transform on the projector grid), low-passed
 * to lowpass_A (<= 0: none) with a raised-cosine edge of width 2 shells. Output float re/im of
 * pad*pad*(pad/2+1) on the DEVICE (dre/dim device pointers), scaled like Projector::computeFourierTransformMap. */
int t2x_synth_reference_device(t2x_device* dev, int ori_size, double pixel_size, int pad_size, int r_max,
                               double padding_factor, const t2x_blob* blobs, int n_blobs,
                               double lowpass_A, float* dre, float* dim);
/* Particle images on the device: projection of blob set `class_blobs[true_class]` at the true pose,
 * CTF-modulated, shifted by (true_x, true_y) px, white noise with variance var(signal)/snr[i]
 * (var over the whole box); the noise of pixel q of particle i is the counter-based Gaussian
 * sample (noise_seed, i, q). dimages: device n*ori_size^2 floats. */
int t2x_synth_particles_device(t2x_device* dev, int ori_size, double pixel_size, int n,
                               const t2x_blob* blobs, const int* blob_offsets, /* [K+1] */
                               const t2x_particle_in* pin, const double* true_rot,
                               const double* true_tilt, const double* true_psi,
                               const double* true_x, const double* true_y, const int* true_class,
                               const double* snr, uint64_t noise_seed, float* dimages,
                               double* noise_var /* [n] host, may be NULL */);

/* ---------------------------------------------------------------------------------------
 * Learned pose proposal. No counterpart in the reference (it has no learned component): a small
 * CNN (64x64 input, four 3x3 stride-2 convolutions of 32/64/128/128 channels, linear head) maps a
 * particle to a distribution over bins = classes x HEALPix directions one order below the coarse
 * search, plus a 2-D shift regression. It is trained on the GPU on noisy CTF-modulated projections
 * of the current references (the device projectors), with the model's noise spectrum; with
 * cfg->proposal_topk > 0 the E-step's coarse pass scores only the top-k bins of each particle and
 * searches exhaustively when the gate value (cfg->proposal_gate: top softmax probability or
 * top-k probability mass) is below cfg->proposal_min_conf.
 * ------------------------------------------------------------------------------------- */
typedef struct t2x_proposal_train_params {
    int      steps;            /* optimiser steps (2000)                                     */
    int      batch;            /* samples per step (256)                                      */
    int      n_projections;    /* distinct projection parameter sets, cycled (200000)        */
    int      bf16;             /* 1: BF16 operands on MFMA, FP32 accumulate; 0: FP32 operands */
    double   lr;               /* peak Adam learning rate (linear warm-up, cosine decay)     */
    double   shift_weight;     /* weight of the shift regression term                        */
    double   max_shift_px;     /* shifts uniform in +-max_shift_px; targets are shift/max    */
    double   defocus_min, defocus_max;   /* A, defocus U and V uniform in this range          */
    double   voltage, cs, q0;  /* kV, mm, amplitude contrast of the training CTFs            */
    uint64_t seed;
    int      reserved[6];
} t2x_proposal_train_params;

typedef struct t2x_proposal_train_stats {
    double ms_total;           /* wall time of t2x_proposal_train                            */
    double ms_setup;           /* parameter sets, labels, uploads                            */
    double loss_first, loss_last;   /* mean loss over the first / last 20 steps             */
    double top1_last;          /* training top-1 bin accuracy over the last 20 steps         */
    double top16_last;         /* training top-16 bin accuracy over the last 20 steps        */
    int    steps;
    int    n_bins;
    int    reserved[6];
} t2x_proposal_train_stats;

/* (Re)initialise the network (He-uniform, seeded) for n_bins outputs. */
int t2x_proposal_init(t2x_device* dev, int n_bins, uint64_t seed);
int t2x_proposal_num_params(t2x_device* dev, long long* n);
int t2x_proposal_get_params(t2x_device* dev, float* host, long long n);
int t2x_proposal_set_params(t2x_device* dev, const float* host, long long n);
/* Declare the current parameters (e.g. set from a network trained earlier) usable for the
 * geometry of cfg (box, classes, HEALPix order), as a finished t2x_proposal_train does. */
int t2x_proposal_mark_trained(t2x_device* dev, const t2x_estep_config* cfg);
/* One forward + backward pass on host data (tests): x [B][64][64], label [B], shift_target
 * [B][2]; writes the mean loss, the gradient (n_params floats, optional) and the outputs
 * [B][n_bins+2] (optional). bf16 selects the operand precision of the layer GEMMs. */
int t2x_proposal_loss_grad(t2x_device* dev, int B, const float* x, const int* label,
                           const float* shift_target, double shift_weight, int bf16, double* loss,
                           float* grad, float* outputs);
/* Train on the GPU from the projectors of classes 0..K-1 set on this handle; cfg gives the box,
 * pixel size, classes and the coarse HEALPix order (bins are one order below it); model gives
 * the noise spectrum. */
int t2x_proposal_train(t2x_device* dev, const t2x_estep_config* cfg, const t2x_model* model,
                       const t2x_proposal_train_params* tp, t2x_proposal_train_stats* out);
/* Gate values of the particles of the last t2x_estep call on this handle that ran the
 * proposal, in input order: top softmax probability and top-k probability mass (either may be
 * NULL). n must equal that call's particle count. */
int t2x_proposal_gate_values(t2x_device* dev, int n, float* top1, float* topk_mass);

/* ---------------------------------------------------------------------------------------
 * Reconstruction (M-step side of the backprojector), all arithmetic in FP64 on the GPU.
 * ------------------------------------------------------------------------------------- */
typedef struct t2x_recon_params {
    int    ori_size;        /* box size of the output map                                     */
    int    pad_size;        /* BackProjector::pad_size = 2*(ROUND(padding_factor*r_max)+1)+1   */
    int    r_max;           /* BackProjector::r_max (Fourier pixels of the original box)       */
    double padding_factor;
    int    ref_dim;         /* 3                                                              */
    int    data_dim;        /* 2 for particle images (selects the FFT normalisation)          */
    int    do_map;          /* add 1/(pf^3 tau2_fudge tau2) to the weights (MAP regularisation) */
    double tau2_fudge;
    int    minres_map;      /* shells below this one get no tau2 term                         */
    int    skip_gridding;   /* must be 1: relion_refine's default (no --dont_skip_gridding)    */
    int    reserved[6];
} t2x_recon_params;

/* Replaces BackProjector::reconstruct (backprojector.cpp:1344-1770) on its default
 * skip_gridding branch: decenter (projector.h:242-253), MAP tau2 term (:1416-1457), radially
 * averaged weight floor and data/weight division (:1459-1534), windowToOridimRealSpace
 * (:2486-2600: Fourier window to 2*ori, inverse FFT, real-space window, FFT normalisation,
 * softMaskOutsideMap mask.cpp:43-97) and griddingCorrect (projector.cpp Projector::griddingCorrect,
 * trilinear sinc^2). Inputs are BackProjector::data (re, im) and ::weight in the projector
 * layout pad*pad*(pad/2+1) [z][y][x], y and z centred at pad/2; tau2 has ori_size/2+1 shells
 * (used when do_map). vol_out: ori_size^3 doubles, [z][y][x], centre at ori_size/2.
 * The x = 0 and Nyquist planes are Hermitian-symmetrised before the half-complex inverse
 * transform, which is what such a transform does with their imaginary parts. */
int t2x_reconstruct(t2x_device* dev, const t2x_recon_params* p, const double* data_re,
                    const double* data_im, const double* weight, const double* tau2, double* vol_out);

/* The same reconstruction with its inputs taken from class iclass's device backprojector
 * (t2x_backprojector_init / the E-step's accumulation) instead of host arrays: replaces the
 * AccBackprojector::getMdlData copy (acc_backprojector_impl.h:84-98) and its XFLOAT->RFLOAT
 * widening into wsum_model.BPref[j].data/weight (ml_optimiser.cpp:3483-3503, one GPU, BPref
 * starting at zero) followed by BackProjector::reconstruct (backprojector.cpp:1344-1770),
 * without the volumes crossing PCIe. The backprojector's
 * ref_dim, pad_size, r_max and padding_factor must equal p's. Results equal t2x_reconstruct on
 * the volumes t2x_backprojector_get returns. */
int t2x_reconstruct_class(t2x_device* dev, int iclass, const t2x_recon_params* p, const double* tau2,
                          double* vol_out);

/* Replaces getFSC(MultidimArray<RFLOAT>&, MultidimArray<RFLOAT>&, fsc) (fftw.cpp:481-523): Fourier
 * shell correlation of two n^3 real maps ([z][y][x]); fsc has n/2+1 shells, shell = ROUND(|k|). */
int t2x_fsc(t2x_device* dev, int n, const double* map1, const double* map2, double* fsc);

#ifdef __cplusplus
}
#endif
#endif /* T2X_H */
