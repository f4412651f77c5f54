/*
 * oracle.h — FP64 CPU restatement of relion_refine's CPU expectation step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for libt2x: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product
 * (libt2x.so) never links, loads or calls it.
 *
 * Everything is double precision, one particle per OpenMP thread, following
 * src/ml_optimiser.cpp:4252-8992, src/projector.cpp:524-691,
 * src/backprojector.cpp:55-357, src/ctf.{h,cpp}, src/healpix_sampling.cpp.
 */
#ifndef T2X_ORACLE_H
#define T2X_ORACLE_H

#include "../include/t2x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One class: projector (data only) or backprojector (data + weight). */
typedef struct orc_volume {
    int ref_dim;              /* 3 or 2                                          */
    int pad_size;             /* odd; dims pad x pad x (pad/2+1) (3D)            */
    int r_max;                /* Projector::r_max                                */
    double padding_factor;    /* 2                                               */
    double* data;             /* complex interleaved (re,im)                     */
    double* weight;           /* backprojector weight (NULL for a projector)     */
} orc_volume;

/* projection boundary semantics */
enum {
    ORC_PROJ_CLEAN = 0,   /* pixels outside r_max are zero (the comment of projector.cpp:526-528)      */
    ORC_PROJ_STALE = 1    /* one reused output array per pass, as ml_optimiser.cpp reuses Fref: pixels  */
                          /* skipped by the floating-point r_max test keep the previous projection     */
};

int  orc_abi_version(void);

/* --- HEALPix NEST scheme (src/Healpix_2.15a/healpix_base.cc, Gorski et al. 2005) --- */
void orc_healpix_pix2ang_z_phi(int order, long ipix, double* z, double* phi);
void orc_healpix_nest2xyf(int order, long ipix, int* ix, int* iy, int* face);
long orc_healpix_xyf2nest(int order, int ix, int iy, int face);

/* --- sampling (src/healpix_sampling.cpp) --- */
int  orc_sampling_info(const t2x_estep_config* cfg, t2x_sampling_info* info);
/* coarse directions (rot, tilt) [nr_dir], psi [nr_psi], translations in A [nr_trans] */
int  orc_sampling_coarse(const t2x_estep_config* cfg, double* rot, double* tilt, double* psi,
                         double* tx, double* ty);
/* oversampled orientations of coarse (idir, ipsi): rot/tilt/psi [nr_over_rot] */
int  orc_sampling_oversampled_orientations(const t2x_estep_config* cfg, long idir, long ipsi,
                                           int oversampling, double* rot, double* tilt, double* psi);
/* oversampled translations of coarse itrans in px: x/y [nr_over_trans] */
int  orc_sampling_oversampled_translations(const t2x_estep_config* cfg, long itrans,
                                           int oversampling, double* x, double* y);

/* --- geometry and CTF --- */
void   orc_euler_angles2matrix(double rot, double tilt, double psi, double* A9);
double orc_ctf_value(const t2x_ctf_params* p, double X, double Y);
/* FFTW-layout CTF image of size x (size/2+1) for an ori_size box (CTF::getFftwImage, no padding) */
void   orc_ctf_image(const t2x_ctf_params* p, int ori_size, double angpix, int size, double* out);

/* --- FFT (RELION normalisation: forward divided by n*n) --- */
void orc_fft2_r2c(int n, const double* in, double* out_cplx);   /* n x n real -> n x (n/2+1) complex */
void orc_fft2_c2r(int n, const double* in_cplx, double* out);   /* unnormalised inverse               */

/* --- projector / backprojector --- */
/* Projector::project (projector.cpp:524-691) into out (complex, size x (size/2+1), zero-filled first) */
void orc_project(const orc_volume* v, const double* A9, int size, double* out);
/* BackProjector::backproject2Dto3D (backprojector.cpp:55-357) */
void orc_backproject(orc_volume* v, const double* f2d, const double* fweight, int size, const double* A9);

/* --- synthetic data (same formulas as the GPU generator, in double precision) --- */
int orc_synth_reference(int ori_size, double pixel_size, int pad_size, int r_max,
                        double padding_factor, const t2x_blob* blobs, int n_blobs,
                        double lowpass_A, double* data_cplx);
int orc_synth_particles(int ori_size, double pixel_size, int n, const t2x_blob* blobs,
                        const int* blob_offsets, const t2x_particle_in* pin,
                        const double* true_rot, const double* true_tilt, const double* true_psi,
                        const double* true_x, const double* true_y, const int* true_class,
                        const double* snr, uint64_t noise_seed, float* images,
                        double* noise_var /* [n], may be NULL */);
/* The counter-based Gaussian noise sample used by both generators. */
double orc_noise_sample(uint64_t seed, long particle, long pixel);

/* --- the E-step --- */
/* Runs expectationOneParticle for n particles. proj[K] are read; bp[K] are added into.
 * proj_mode: ORC_PROJ_CLEAN or ORC_PROJ_STALE. nthreads <= 0: all cores. */
int orc_estep(const t2x_estep_config* cfg, const t2x_model* model, const orc_volume* proj,
              orc_volume* bp, int n, const float* images, const t2x_particle_in* pin,
              t2x_particle_out* out, t2x_wsum* wsum, t2x_estep_debug* dbg, int proj_mode,
              int nthreads);

/* Preprocessing only (getFourierTransformsAndCtfs, ml_optimiser.cpp:5689-6593) for one image:
 * Fimg / Fimg_nomask (complex, current_size x (current_size/2+1)), Fctf (same, real),
 * power spectrum [ori_size/2+1], highres_Xi2. sigma2_noise [ori_size/2+1] is the particle's
 * optics group's spectrum, read for the noise background (do_zero_mask 0; may be NULL otherwise). */
int orc_preprocess(const t2x_estep_config* cfg, const float* image, const t2x_particle_in* pin,
                   const double* sigma2_noise, double* Fimg, double* Fimg_nomask, double* Fctf,
                   double* power, double* highres_Xi2, double* old_offset_xy);
/* The noise image of the mask background (do_zero_mask 0): ori_size^2 doubles, row-major. */
int orc_background_noise(const t2x_estep_config* cfg, const double* sigma2_noise, long part_id,
                         double* noise);

/* Weight conversion of a dense diff2 row (convertAllSquaredDifferencesToWeights core,
 * ml_optimiser.cpp:7760-8031) with per-sample prior factors: weights out, sum, threshold. */
int orc_weights_from_diff2(long n, const double* diff2, const double* prior, double min_diff2,
                           double adaptive_fraction, long maximum_significants, int count_significant,
                           double* weights, double* sum_weight, double* significant_weight,
                           long* nr_significant);

/* --- reconstruction (recon_oracle.cpp): BackProjector::reconstruct, skip_gridding branch;
 *     getFSC of two real maps. Layouts as t2x_reconstruct / t2x_fsc. --- */
int orc_reconstruct(const t2x_recon_params* p, const double* data_re, const double* data_im,
                    const double* weight, const double* tau2, double* vol_out);
int orc_fsc(int n, const double* m1, const double* m2, double* fsc);

#ifdef __cplusplus
}
#endif
#endif
