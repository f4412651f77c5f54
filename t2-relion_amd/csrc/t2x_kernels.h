// t2x_kernels.h — kernel declarations shared by the E-step driver.
#pragma once
#include "t2x_internal.h"

namespace t2x {

typedef float floatx4 __attribute__((ext_vector_type(4)));
// 8 x 16-bit operand fragment of v_mfma_f32_16x16x32_{bf16,f16}
template <class H> struct Vec8;
template <> struct Vec8<__bf16> { typedef __bf16 type __attribute__((ext_vector_type(8))); };
template <> struct Vec8<_Float16> { typedef _Float16 type __attribute__((ext_vector_type(8))); };

// Per-particle preprocessing scalars (device, double).
struct PartScalars {
    double highres_Xi2;   // sum |F|^2 beyond current_size (ml_optimiser.cpp:6215-6240)
    double old_x, old_y;  // rounded old offset (px)
    double prior_x, prior_y;
    double norm_scale;    // avg_norm / normcorr (1 when norm correction is off)
    double C_coarse;      // highres/2 + 1/2 sum_Mc minv |X|^2
    double C_fine;        // highres/2 + 1/2 sum_Mf minv |X|^2
    double C_low;         // highres/2 + 1/2 sum minv |X|^2 over a low-frequency prefix of Mc (pruning)
    double scale;         // the image group's scale correction (1 without), folded into the CTF
    double ctf_store;     // factor on that CTF in the weighted sums and the backprojection (the
                          // store clamps the scale at 0.001, ml_optimiser.cpp:8333-8343)
    long long part_id;    // key of the mask-background noise
    int og;               // optics group: row of the Minvsigma2 tables and of the noise spectrum
    int grp;              // image group
};

// Per-particle selection results (device).
struct SelResult {
    double sum_w;       // sum of weights
    double sig_w;       // significant weight threshold
    double min_d2;      // min diff2 used for the weights
    long long nsig;     // nr of significant samples counted by the threshold loop
    long long npos;     // nr of positive weights
    double max_w;       // max weight
    long long best_key; // CPU store-order key of the max weight (fine pass)
    long long best_e;   // element index of the max weight
};

enum { SRC_COARSE = 0, SRC_FINE = 1 };

// Most coarse translations of a pass (per-translation LDS tables and significance masks): e.g.
// an offset range of 12 px at a step of 1 px (453 translations).
#define MAX_COARSE_T 512

// Totals of a batch's work lists (k_batch_plan), the first particle with an error status, and
// the work-queue counters of the persistent fine-pass and store kernels.
struct BatchPlan {
    long long npairs, norients, nitems, ntiles;
    int exceeded;       // the lists' capacity was too small: every offset written as zero
    int err_particle;   // first particle (in the batch) with an error status, or -1
    int err_code;       // its status (T2X_EZEROSUM, T2X_ENOSIG)
    int pad;
    unsigned long long queue[2];   // [0] fine tiles, [1] store items
};

// Local angular searches (selectOrientationsWithNonZeroPriorProbability, healpix_sampling.cpp:
// 769-1145): per particle, the normalised prior of every direction and psi angle (0 outside the
// search range) and the mean of their products over the particle's orientations. dir == null:
// global search.
struct LocalPrior {
    const double* dir;    // [B][Nd]
    const double* psi;    // [B][Npsi]
    const double* mean;   // [B]
    int Nd, Npsi;
};

// pdf_orientation / mean of orientation n (all classes) for particle b (ml_optimiser.cpp:
// 7700-7716); prior_o holds pdf_direction / mean (global search) or the class gate 1 / 0 (local).
__device__ __forceinline__ double orient_prior(const double* prior_o, const LocalPrior& lp, int b, long long n)
{
    double p = prior_o[n];
    if (lp.dir && p > 0.) {
        const long long no = (long long)lp.Nd * lp.Npsi;
        const long long o = n % no;
        const int d = (int)(o / lp.Npsi), ps = (int)(o - (long long)d * lp.Npsi);
        const double q = lp.dir[(size_t)b * lp.Nd + d] * lp.psi[(size_t)b * lp.Npsi + ps];
        const double m = lp.mean[b];
        p = (m != 0.) ? q / m : q;
    }
    return p;
}

// Whether orientation n is in particle b's local search range.
__device__ __forceinline__ bool local_in_range(const LocalPrior& lp, int b, long long n)
{
    const long long no = (long long)lp.Nd * lp.Npsi;
    const long long o = n % no;
    const int d = (int)(o / lp.Npsi), ps = (int)(o - (long long)d * lp.Npsi);
    return lp.dir[(size_t)b * lp.Nd + d] > 0. && lp.psi[(size_t)b * lp.Npsi + ps] > 0.;
}

// Description of the weights of one pass, read by the selection kernels.
struct WeightSrc {
    int kind;                  // SRC_COARSE / SRC_FINE
    const float* d2;           // coarse: [B][T][ld]; fine: [total fine]
    long long ld;              // coarse: row stride (orientations, padded)
    int T;                     // coarse translations
    int O_all;                 // coarse orientations (all classes)
    const double* prior_o;     // [O_all]
    const double* log_prior_o; // [O_all] log(prior_o) (-inf where 0), coarse candidates only
    const double* prior_t;     // [B][T], or [B][T][tK] (see trans_prior)
    int tK;                    // 0: one translation prior per particle; > 0: per class
    int tcols;                 // with tK > 0: orientation columns per class
    const float* rowmin;       // coarse: [B*T][ld/64] per-row minima from the GEMM
    int rowmin_ld;
    // fine
    const long long* pair_off; // [B+1] offsets into pair lists (pairs)
    const int2* pairs;         // (n, t) per significant coarse pair
    const unsigned int* fine_min; // [B] ordered-uint minima
    int nover;                 // fine samples per pair (32)
    int over_trans;            // 4
    // coarse, learned proposal: bit per (particle, bin) of the scored orientation blocks; null
    // = all orientations scored. bin of orientation n = n / bin_cols.
    const unsigned int* bin_mask;
    int mask_words;            // 32-bit words per particle
    int bin_cols;              // orientations per bin
    LocalPrior lp;             // local searches (lp.dir null: global)
};

// Translation prior of particle b, coarse translation t, for orientation column n. 2D references
// with class-dependent offset priors (ml_optimiser.cpp:7627-7631, 7678-7682) hold [B][T][tK] (a
// row's classes contiguous), the class of column n being n / tcols; otherwise [B][T].
__device__ __forceinline__ double trans_prior(const double* pt, int T, int tK, int tcols, int b, long long n, int t)
{
    if (tK == 0) return pt[(size_t)b * T + t];
    return pt[((size_t)b * T + t) * tK + (unsigned)n / (unsigned)tcols];
}

__device__ __forceinline__ bool coarse_allowed(const WeightSrc& s, int b, long long n)
{
    if (!s.bin_mask) return true;
    const int bin = (int)(n / s.bin_cols);
    return (s.bin_mask[(size_t)b * s.mask_words + (bin >> 5)] >> (bin & 31)) & 1u;
}

// Epilogue parameters of the coarse cross-term GEMMs.
// Variant of the split-precision GEMM: tile 128 / 256 / 257 / 258 (0: autotuned) and tile-walk
// group in tile rows (0: the variant's default).
struct GemmChoice {
    int tile, group;
};

struct EpiParams {
    float* C;            // output
    long long ldc;
    int M, N;            // valid extent
    // EPI_DIFF2
    int T;               // translations per particle (rows per particle)
    const double* Cb;    // [B] constant term per particle (read as double)
    const float* N1;     // [B][ldn]
    long long ldn;
    const double* prior_o;  // [N] orientation prior (0 = not scored)
    float* rowmin;       // [M][N_pad/64]
    int rowmin_ld;
    // running per-particle minimum (order-preserving uint of the float); when set, the values of a
    // 64-column block whose minimum exceeds it by more than the negligible-weight cut are not
    // stored (the block minimum still is, so every reader skips the block). null: store all.
    unsigned int* pmin;
    // Pruned coarse pass. C null: a bounds-only launch (EPI_DIFF2 writes the 64-column block
    // minima, no values). tile_list: the tiles a list-driven (persistent) launch computes,
    // *tile_count of them (read by the kernels of the LIST form only).
    const unsigned int* tile_list;
    const unsigned int* tile_count;
    // Bounds launch of the pruned coarse pass: N1h (the N1 term over the pixels beyond the
    // low-frequency prefix, ld ldn) non-null turns each N1 read into N1 - c0 sqrt(N1h) -
    // c1 sqrt(N1), with prune_c[2 p], prune_c[2 p + 1] = c0, c1 of particle p.
    const float* N1h;
    const double* prune_c;
};

struct SelParams {
    double fraction;
    long long max_sig;
    int count_sig;
};

}  // namespace t2x
