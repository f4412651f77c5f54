// t2x_estep.hip — the E-step driver and its per-particle kernels.
//
// One t2x_estep call replaces MlOptimiserCuda::doThreadExpectationSomeParticles
// (src/acc/cuda/cuda_ml_optimiser.cu:1345-1408) for a whole set of particles: instead of a
// per-thread task graph interleaving 4 particles, particles are processed in large batches
// with one launch per stage, sized for HBM:
//   1. preprocessing (getFourierTransformsAndCtfs, ml_optimiser.cpp:5689-6593)
//   2. coarse pass: slices of all coarse orientations projected once per call and reused by
//      every particle; diff2 for all (particle, translation) x (class, orientation) as one
//      FP32 MFMA GEMM (getAllSquaredDifferences ipass 0, ml_optimiser.cpp:6922-7507)
//   3. FP64 weights + exact radix-select significance (convertAllSquaredDifferencesToWeights)
//   4. fine pass over the significant coarse pairs (ipass 1)
//   5. FP64 weights + significance of the fine samples, best pose
//   6. weighted sums + backprojection (storeWeightedSums, ml_optimiser.cpp:8043-8992)
#include <algorithm>
#include <numeric>

#include "t2x_kernels.h"
#include "t2x_cnn.h"

#include <chrono>
#include <hipcub/hipcub.hpp>

namespace t2x {

enum { EPI_STORE = 0, EPI_DIFF2 = 1 };
int launch_gemm(int epi, const float* A, long long lda, const float* B, long long ldb, int K,
                int Mrows_pad, int Nrows_pad, const EpiParams& ep, hipStream_t st);
int launch_split(int prec, const float* X, long long ldx, int K, int rows, int valid_rows, void* hi,
                 void* lo, long long ld, float* inv_scale, hipStream_t st);
int launch_gemm_x3_bins(int prec, const void* Ah, const void* Al, const float* As, long long lda, const void* Bh,
                        const void* Bl, const float* Bs, long long ldb, int K, const int* plist, const void* work,
                        int n_work, int T, long long n_clamp, const EpiParams& ep, hipStream_t st);
struct BinWorkHost {
    int pstart, cnt, rt, col0, ncols;
};
namespace cnn {
int proposal_infer(Device* d, const float2* Fimg, int cs, int B, int topk, double min_conf, int gate,
                   std::vector<int>& bins, std::vector<float>& maxp, std::vector<float>& mass,
                   std::vector<int>& fallback);
}
int launch_gemm_x3(int prec, int epi, const void* Ah, const void* Al, const float* As, long long lda,
                   long long a_rows, const void* Bh, const void* Bl, const float* Bs, long long ldb, int K,
                   int Mrows_pad, int Nrows_pad, const EpiParams& ep, hipStream_t st, GemmChoice want,
                   const int* rowmap, GemmChoice* used);
int launch_gemm_h1_bounds(const void* Ah, const float* As, long long lda, long long a_rows, const void* Bh,
                          const float* Bs, long long ldb, int K, int Mrows_pad, int Nrows_pad, const EpiParams& ep,
                          hipStream_t st);
int launch_gemm_pruned(int prec, const void* Ah, const void* Al, const float* As, long long lda, long long a_rows,
                       const void* Bh, const void* Bl, const float* Bs, long long ldb, int K, int Mrows_pad,
                       int Nrows_pad, const EpiParams& ep, const float* lb, const float* seed, unsigned* list,
                       long long list_cap, unsigned* count, int n_cu, hipStream_t st);
int launch_gemm_local_x3(int prec, const void* Sh, const void* Sl, const float* Ss, const void* Wh, const void* Wl,
                         const float* Ws, long long ld, int K, const int2* items, int n_items, int T,
                         const long long* orient_off, const int* rowu, const double* cb, const double* n1,
                         float* d2, unsigned int* dmin, hipStream_t st);

__global__ void k_pass_min(WeightSrc s, int B, SelResult* res);
__global__ void k_select(WeightSrc s, SelParams prm, int B, SelResult* res, const int* only);
__global__ void k_sig_count(WeightSrc s, int B, const SelResult* res, long long* counts, const int* only);
__global__ void k_sig_write(WeightSrc s, int B, const SelResult* res, const long long* pair_off,
                            const long long* orient_off, int2* pairs, int* orients,
                            long long* orient_pair_start, const int* only);
__global__ void k_batch_plan(const long long* counts, const SelResult* selc, int B, int over_rot, int rows_per_tile,
                             int ncb, long long cap_pairs, long long cap_orients, long long* pair_off,
                             long long* orient_off, long long* item_off, long long* tile_off, BatchPlan* plan);
__global__ void k_fine_argmax(WeightSrc s, int B, SelResult* res, long long* nfine_sig);
struct SumPart;
__global__ void k_coarse_sum_rows(WeightSrc s, int B, const SelResult* res, SumPart* part);
__global__ void k_coarse_sum_final(WeightSrc s, int B, const SumPart* part, SelResult* res);
__global__ void k_coarse_candidates(WeightSrc s, int B, SelParams prm, const SelResult* res,
                                    double* cand_w, int* cand_e, int* cand_count);
size_t sum_part_bytes();
__global__ void k_log_prior(const double* p, long long n, double* lp);
__global__ void k_cand_select(WeightSrc s, int B, SelParams prm, SelResult* res, const double* cand_w,
                              const int* cand_e, const int* cand_count, int2* pairs_strided,
                              long long* counts, int* overflow, int force);
__global__ void k_cand_scatter(int B, const int2* pairs_strided, const long long* pair_off,
                               const long long* orient_off, int2* pairs, int* orients,
                               long long* orient_pair_start, const int* skip);
static const int kCandCap = 4096;   // CAND_CAP of t2x_select.hip
// Dynamic LDS of the kernels that hold one operand row while they find its maximum; longer rows
// take the path through HBM (FP32 row, then k_split_rows).
static const size_t kRowLdsMax = 65536;

static const double kPi = 3.14159265358979323846;

static inline long long round_up(long long x, long long m) { return (x + m - 1) / m * m; }
__host__ __device__ static inline int relion_round(double x) { return (x > 0) ? (int)(x + 0.5) : (int)(x - 0.5); }

// ------------------------------------------------------------------------------------------
// Preprocessing kernels
// ------------------------------------------------------------------------------------------

// Norm correction, integer re-centring with zero fill (selfTranslate DONT_WRAP on the rounded
// old offset, ml_optimiser.cpp:6097-6098) and the soft circular mask (softMaskOutsideMap,
// mask.cpp:43-97) whose background is the average value (--zero_mask, noise == null) or the
// particle's noise image noise[b]. Two launches over (particle, band of N / PREP_BANDS rows):
// k_prep_bg sums the background weights and values of its band (pixels at r >= radius), and
// k_prep_mask, with the particle's average from the bands' sums (in band order), writes the
// unmasked image to out[b] and the masked image to out[B + b].
#define PREP_BANDS 8

__device__ __forceinline__ float prep_pixel(const float* __restrict__ src, int N, int i, int j, int sx, int sy,
                                            float scale)
{
    const int si = i - sy, sj = j - sx;
    return (si >= 0 && si < N && sj >= 0 && sj < N) ? src[(size_t)si * N + sj] * scale : 0.f;
}

__global__ __launch_bounds__(256) void k_prep_bg(const float* __restrict__ img, int N,
                                                 const PartScalars* __restrict__ ps, double radius,
                                                 double cosine_width, double* __restrict__ part)
{
    const int b = blockIdx.x, band = blockIdx.y;
    const int r0 = band * N / PREP_BANDS, r1 = (band + 1) * N / PREP_BANDS;
    const float* src = img + (size_t)b * N * N;
    const int sx = (int)ps[b].old_x, sy = (int)ps[b].old_y;
    const float scale = (float)ps[b].norm_scale;
    const double radius_p = radius + cosine_width;
    // squared distances clear of both edges by half a pixel need no square root
    const double in2 = radius > 0.5 ? (radius - 0.5) * (radius - 0.5) : -1., out2 = (radius_p + 0.5) * (radius_p + 0.5);
    double sbg = 0., sw = 0.;
    for (int q = threadIdx.x; q < (r1 - r0) * N; q += blockDim.x) {
        const int i = r0 + q / N, j = q - (q / N) * N;
        {
            const double y = i - N / 2, x = j - N / 2;
            const double n2 = x * x + y * y;
            if (n2 <= in2) continue;
            const float v = prep_pixel(src, N, i, j, sx, sy, scale);
            if (n2 >= out2) { sw += 1.; sbg += v; continue; }
            const double r = sqrt(n2);
            if (r < radius) continue;
            if (r > radius_p) { sw += 1.; sbg += v; }
            else {
                const double rc = 0.5 + 0.5 * cos(kPi * (radius_p - r) / cosine_width);
                sw += rc;
                sbg += rc * v;
            }
        }
    }
    __shared__ double r1s[4], r2s[4];
    for (int o = 32; o > 0; o >>= 1) { sbg += __shfl_xor(sbg, o); sw += __shfl_xor(sw, o); }
    if ((threadIdx.x & 63) == 0) { r1s[threadIdx.x >> 6] = sbg; r2s[threadIdx.x >> 6] = sw; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* o = part + ((size_t)b * PREP_BANDS + band) * 2;
        o[0] = r1s[0] + r1s[1] + r1s[2] + r1s[3];
        o[1] = r2s[0] + r2s[1] + r2s[2] + r2s[3];
    }
}

__global__ __launch_bounds__(256) void k_prep_mask(const float* __restrict__ img, int N, int B,
                                                   const PartScalars* __restrict__ ps, double radius,
                                                   double cosine_width, const float* __restrict__ noise,
                                                   const double* __restrict__ part, float* __restrict__ out)
{
    const int b = blockIdx.x, band = blockIdx.y;
    const int r0 = band * N / PREP_BANDS, r1 = (band + 1) * N / PREP_BANDS;
    const size_t NN = (size_t)N * N;
    const float* src = img + (size_t)b * NN;
    float* um = out + (size_t)b * NN;
    float* mk = out + (size_t)(B + b) * NN;
    const int sx = (int)ps[b].old_x, sy = (int)ps[b].old_y;
    const float scale = (float)ps[b].norm_scale;
    const double radius_p = radius + cosine_width;
    double bg = 0.;
    if (!noise) {
        double s1 = 0., s2 = 0.;
        for (int k = 0; k < PREP_BANDS; k++) {
            s1 += part[((size_t)b * PREP_BANDS + k) * 2];
            s2 += part[((size_t)b * PREP_BANDS + k) * 2 + 1];
        }
        bg = s1 / s2;
    }
    const float* nz = noise ? noise + (size_t)b * NN : nullptr;
    const double in2 = radius > 0.5 ? (radius - 0.5) * (radius - 0.5) : -1., out2 = (radius_p + 0.5) * (radius_p + 0.5);
    for (int t = threadIdx.x; t < (r1 - r0) * N; t += blockDim.x) {
        const int i = r0 + t / N, j = t - (t / N) * N;
        {
            const size_t q = (size_t)i * N + j;
            const double y = i - N / 2, x = j - N / 2;
            const double n2 = x * x + y * y;
            const float v = prep_pixel(src, N, i, j, sx, sy, scale);
            um[q] = v;
            if (n2 <= in2) { mk[q] = v; continue; }
            const double add = nz ? (double)nz[q] : bg;
            if (n2 >= out2) { mk[q] = (float)add; continue; }
            const double r = sqrt(n2);
            if (r < radius) mk[q] = v;
            else if (r > radius_p) mk[q] = (float)add;
            else {
                const double rc = 0.5 + 0.5 * cos(kPi * (radius_p - r) / cosine_width);
                mk[q] = (float)((1 - rc) * v + rc * add);
            }
        }
    }
}

// Fourier components of each particle's mask-background noise (ml_optimiser.cpp:6173-6212) in
// the half layout N x (N/2+1): within shell N/2, complex Gaussians of standard deviation
// sqrt(fudge * sigma2_noise[og][shell]) per part, keyed by (seed, part_id, 2 q + part) with q
// the component's index; the x = 0 and x = N/2 columns Hermitian (row -y the conjugate of row
// y, real at the self-conjugate frequencies) so that the real inverse transform is unambiguous.
__global__ __launch_bounds__(256) void k_noise_spectrum(int N, const PartScalars* __restrict__ ps,
                                                        const double* __restrict__ sigma2, double fudge,
                                                        uint64_t seed, float2* __restrict__ F)
{
    const int b = blockIdx.x, xN = N / 2 + 1;
    const PartScalars s = ps[b];
    const double* s2 = sigma2 + (size_t)s.og * xN;
    float2* out = F + (size_t)b * N * xN;
    for (int q = threadIdx.x; q < N * xN; q += blockDim.x) {
        const int i = q / xN, j = q - i * xN;
        const int ip = (i < xN) ? i : i - N;
        const int ires = relion_round(sqrt((double)(ip * ip + j * j)));
        float2 v = make_float2(0.f, 0.f);
        if (ires < xN) {
            const bool edge = (j == 0 || j == N / 2);
            const bool conj = edge && ip < 0;
            const long src = (long)(conj ? N - i : i) * xN + j;
            const double sd = sqrt(fudge * s2[ires]);
            const double re = sd * noise_sample(seed, (long)s.part_id, 2 * src);
            double im = sd * noise_sample(seed, (long)s.part_id, 2 * src + 1);
            if (edge && (i == 0 || i == N / 2)) im = 0.;
            v = make_float2((float)re, (float)(conj ? -im : im));
        }
        out[q] = v;
    }
}

// From the two full transforms (unnormalised R2C of unmasked and masked images) to the
// current-size windows with the RELION normalisation and CenterFFTbySign, the CTF image,
// and the power spectrum / high-resolution Xi2 of the masked image (ml_optimiser.cpp:6209-6250).
__global__ __launch_bounds__(256) void k_prep_fourier(const float2* __restrict__ F, int N, int B, int cs,
                                                      const CtfConst* __restrict__ ctf, double angpix,
                                                      int do_ctf, float2* __restrict__ Fimg,
                                                      float2* __restrict__ Fimg_nm,
                                                      float* __restrict__ Fctf,
                                                      double* __restrict__ power,
                                                      PartScalars* __restrict__ ps)
{
    extern __shared__ double spec[];   // [N/2+1]
    const int b = blockIdx.x;
    const int xN = N / 2 + 1, xc = cs / 2 + 1;
    const float2* Fu = F + (size_t)b * N * xN;
    const float2* Fm = F + (size_t)(B + b) * N * xN;
    const double inv = 1.0 / ((double)N * N);
    for (int i = threadIdx.x; i < xN; i += blockDim.x) spec[i] = 0.;
    __syncthreads();
    const CtfConst c = ctf[b];
    for (int q = threadIdx.x; q < cs * xc; q += blockDim.x) {
        int i = q / xc, j = q % xc;
        int ip = (i < xc) ? i : i - cs;
        int ii = ip < 0 ? ip + N : ip;
        float sgn = ((i + j) & 1) ? -1.f : 1.f;
        float2 u = Fu[(size_t)ii * xN + j], m = Fm[(size_t)ii * xN + j];
        Fimg_nm[(size_t)b * cs * xc + q] = make_float2((float)(u.x * inv) * sgn, (float)(u.y * inv) * sgn);
        Fimg[(size_t)b * cs * xc + q] = make_float2((float)(m.x * inv) * sgn, (float)(m.y * inv) * sgn);
        double v = 1.0;
        if (do_ctf) v = ctf_value(c, j / (N * angpix), ip / (N * angpix));
        Fctf[(size_t)b * cs * xc + q] = (float)(v * ps[b].scale);   // Frefctf *= myscale (ml_optimiser.cpp:7119-7125)
    }
    double hres = 0.;
    if (cs < N) {
        for (int q = threadIdx.x; q < N * xN; q += blockDim.x) {
            int i = q / xN, j = q % xN;
            int ip = (i < xN) ? i : i - N;
            int ires = relion_round(sqrt((double)(ip * ip + j * j)));
            if (ires > 0 && ires < xN && !(j == 0 && ip < 0)) {
                float2 m = Fm[q];
                double nf = ((double)m.x * m.x + (double)m.y * m.y) * inv * inv;
                atomicAdd(&spec[ires], nf);
                if (ires >= cs / 2 + 1) hres += nf;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) hres += __shfl_xor(hres, o);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = hres;
    __syncthreads();
    if (threadIdx.x == 0) ps[b].highres_Xi2 = red[0] + red[1] + red[2] + red[3];
    for (int i = threadIdx.x; i < xN; i += blockDim.x) power[(size_t)b * xN + i] = spec[i];
}

// Constant terms C_b = highres/2 + 1/2 sum_p minv |X|^2 for one pass table.
__global__ __launch_bounds__(256) void k_const_term(const float2* __restrict__ Fimg, int Pw,
                                                    const int* __restrict__ win,
                                                    const float* __restrict__ minv, int npix,
                                                    PartScalars* __restrict__ ps, int which, int np_low = 0)
{
    const int b = blockIdx.x;
    const float* mv = minv + (size_t)ps[b].og * npix;   // the optics group's Minvsigma2
    double s = 0.;
    // which 2: the first np of the npix pixels (the minv rows keep their stride npix)
    const int np = which == 2 ? np_low : npix;
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
        float2 x = Fimg[(size_t)b * Pw + win[p]];
        s += 0.5 * (double)mv[p] * ((double)x.x * x.x + (double)x.y * x.y);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = ps[b].highres_Xi2 / 2. + red[0] + red[1] + red[2] + red[3];
        if (which == 0) ps[b].C_coarse = v;
        else if (which == 1) ps[b].C_fine = v;
        else ps[b].C_low = v;   // over a low-frequency prefix of the coarse pixels
    }
}

// ------------------------------------------------------------------------------------------
// Coarse pass
// ------------------------------------------------------------------------------------------

// Slices of all coarse orientations (Projector::project, projector.cpp:524-691): one workgroup
// per (class, orientation), threads over the compact coarse pixels. Writes re/im interleaved
// into B[n][2p..2p+1] and |F|^2 into B2[n][p].
__global__ __launch_bounds__(256) void k_project_coarse(const float2* const* __restrict__ vols,
                                                        const int* __restrict__ pads,
                                                        const float* __restrict__ Ai,   // [Norient][9]
                                                        int Norient, const int2* __restrict__ pix,
                                                        const unsigned char* __restrict__ inproj,
                                                        int npix, float* __restrict__ Bm, long long ldb,
                                                        float* __restrict__ B2, long long ldb2, int ref2d)
{
    const int n = blockIdx.x;
    const int c = n / Norient, o = n - c * Norient;
    const float2* vol = vols[c];
    const int pad = pads[c];
    float a[6];
    for (int i = 0; i < 6; i++) a[i] = Ai[o * 9 + (i / 2) * 3 + (i % 2)];  // rows 0..2, cols 0..1
    for (int p = threadIdx.x; p < npix; p += blockDim.x) {
        float2 f = make_float2(0.f, 0.f);
        if (inproj[p]) {
            float x = (float)pix[p].x, y = (float)pix[p].y;
            float xp = a[0] * x + a[1] * y, yp = a[2] * x + a[3] * y, zp = a[4] * x + a[5] * y;
            float2 v;
            if (ref_sample(vol, pad, ref2d, xp, yp, zp, &v)) f = v;
        }
        Bm[(size_t)n * ldb + 2 * p] = f.x;
        Bm[(size_t)n * ldb + 2 * p + 1] = f.y;
        B2[(size_t)n * ldb2 + p] = f.x * f.x + f.y * f.y;
    }
}

// Image side of the coarse GEMM: A[(b,t)][2p..] = minv ctf X_t and A1[b][p] = minv ctf^2 / 2,
// with X_t = X * exp(-2 pi i (x sx + y sy)/N) (shiftImageInFourierTransform, fftw.cpp:842-930).
__global__ __launch_bounds__(256) void k_coarse_imgside(const float2* __restrict__ Fimg,
                                                        const float* __restrict__ Fctf, int Pw,
                                                        const int* __restrict__ win,
                                                        const float* __restrict__ minv, int npix,
                                                        const PartScalars* __restrict__ ps,
                                                        const float2* __restrict__ ab, int T,
                                                        float* __restrict__ A, long long lda,
                                                        float* __restrict__ A1, long long lda1)
{
    const int b = blockIdx.y, t = blockIdx.x;
    const float* mv = minv + (size_t)ps[b].og * npix;
    for (int p = threadIdx.x; p < npix; p += blockDim.x) {
        const size_t w = (size_t)b * Pw + win[p];
        float2 x = Fimg[w];
        float ctf = Fctf[w], m = mv[p];
        float s = m * ctf;
        float2 sh = ab[(size_t)t * npix + p];
        float xr = x.x * sh.x - x.y * sh.y, xi = x.x * sh.y + x.y * sh.x;
        A[((size_t)b * T + t) * lda + 2 * p] = s * xr;
        A[((size_t)b * T + t) * lda + 2 * p + 1] = s * xi;
        if (t == 0) A1[(size_t)b * lda1 + p] = 0.5f * m * ctf * ctf;
    }
}

// pdf_offset / mean of every coarse translation of each particle (ml_optimiser.cpp:7638-7667,
// 7701-7725): one workgroup per particle, the mean summed in translation order. The offsets mix
// pixels (old offset, prior) and the translation grid's Angstrom as the reference does.
__global__ __launch_bounds__(128) void k_prior_t(const PartScalars* __restrict__ ps, int B,
                                                 const double* __restrict__ tx, const double* __restrict__ ty,
                                                 int T, double pix, double sigma2_offset,
                                                 const double* __restrict__ cls_prior, int K,
                                                 double* __restrict__ prior_t)
{
    __shared__ double pt[MAX_COARSE_T];
    __shared__ double s_mean;
    const int b = blockIdx.x;
    if (b >= B) return;
    const PartScalars s = ps[b];
    if (cls_prior) {
        // 2D references with class-dependent offset priors: one table per class, normalised by
        // the mean over every (class, translation) (ml_optimiser.cpp:7607-7660), [b][t][c]
        const int KT = K * T;
        double* out = prior_t + (size_t)b * KT;
        double part = 0.;
        for (int e = threadIdx.x; e < KT; e += blockDim.x) {
            const int t = e / K, c = e - t * K;
            const double ox = s.old_x + tx[t], oy = s.old_y + ty[t];
            const double qx = cls_prior[2 * c], qy = cls_prior[2 * c + 1];
            const double td2 = ((ox - qx) * (ox - qx) + (oy - qy) * (oy - qy)) * pix * pix;
            const double v = (sigma2_offset < 0.0001) ? ((td2 > 0.) ? 0. : 1.)
                                                      : exp(td2 / (-2. * sigma2_offset)) / (2. * kPi * sigma2_offset);
            out[e] = v;
            part += v;
        }
        pt[threadIdx.x] = part;
        __syncthreads();
        for (int h = blockDim.x / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) pt[threadIdx.x] += pt[threadIdx.x + h];
            __syncthreads();
        }
        const double m = pt[0] / (double)KT;
        // each thread rescales the entries it wrote
        for (int e = threadIdx.x; e < KT; e += blockDim.x) out[e] = (m > 0.) ? out[e] / m : out[e];
        return;
    }
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        const double ox = s.old_x + tx[t], oy = s.old_y + ty[t];
        const double td2 = ((ox - s.prior_x) * (ox - s.prior_x) + (oy - s.prior_y) * (oy - s.prior_y)) * pix * pix;
        pt[t] = (sigma2_offset < 0.0001) ? ((td2 > 0.) ? 0. : 1.)
                                         : exp(td2 / (-2. * sigma2_offset)) / (2. * kPi * sigma2_offset);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.;
        for (int t = 0; t < T; t++) m += pt[t];
        s_mean = m / (double)T;
    }
    __syncthreads();
    const double m = s_mean;
    for (int t = threadIdx.x; t < T; t += blockDim.x) prior_t[(size_t)b * T + t] = (m > 0.) ? pt[t] / m : pt[t];
}

// k_coarse_imgside followed by the split of each row into hi + lo halves (k_split_rows), fused:
// one workgroup per GEMM row (particle b, translation t; rows >= M are zero). FP16 (use_scale):
// the row's values kept in LDS between the max-abs reduction and the scaled split; BF16: no
// scale, the halves written as the values are formed (no LDS). The FP32 A matrix is neither
// cleared, written nor read back. Same arithmetic as the two kernels, so the same bits.
template <class H>
__global__ __launch_bounds__(256) void k_coarse_imgside_split(const float2* __restrict__ Fimg,
                                                              const float* __restrict__ Fctf, int Pw,
                                                              const int* __restrict__ win,
                                                              const float* __restrict__ minv, int npix,
                                                              const PartScalars* __restrict__ ps,
                                                              const float2* __restrict__ ab, int T, int M, int Kc,
                                                              H* __restrict__ hi, H* __restrict__ lo,
                                                              float* __restrict__ inv_scale, int use_scale,
                                                              float* __restrict__ A1, long long lda1)
{
    extern __shared__ float row[];   // [Kc]
    __shared__ float red[4];
    const int r = blockIdx.x;
    const bool valid = r < M;
    const int b = valid ? r / T : 0, t = valid ? r - b * T : 0;
    const float* mv = minv + (valid ? (size_t)ps[b].og * npix : 0);
    float mx = 0.f;
    for (int p = threadIdx.x; p < Kc / 2; p += blockDim.x) {
        float vr = 0.f, vi = 0.f;
        if (valid && p < npix) {
            const size_t w = (size_t)b * Pw + win[p];
            const float2 x = Fimg[w];
            const float ctf = Fctf[w], m = mv[p];
            const float s = m * ctf;
            const float2 sh = ab[(size_t)t * npix + p];
            const float xr = x.x * sh.x - x.y * sh.y, xi = x.x * sh.y + x.y * sh.x;
            vr = s * xr;
            vi = s * xi;
            if (t == 0) A1[(size_t)b * lda1 + p] = 0.5f * m * ctf * ctf;
        }
        if (!use_scale) {   // BF16: no row scale, so no row buffer: the halves straight out
            const H hr = (H)vr, hi_ = (H)vi;
            hi[(size_t)r * Kc + 2 * p] = hr;
            hi[(size_t)r * Kc + 2 * p + 1] = hi_;
            lo[(size_t)r * Kc + 2 * p] = (H)(vr - (float)hr);
            lo[(size_t)r * Kc + 2 * p + 1] = (H)(vi - (float)hi_);
            continue;
        }
        row[2 * p] = vr;
        row[2 * p + 1] = vi;
        mx = fmaxf(mx, fmaxf(fabsf(vr), fabsf(vi)));
    }
    if (!use_scale) {
        if (threadIdx.x == 0) inv_scale[r] = 1.f;
        return;
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    int e = 0;
    if (mx > 0.f) (void)frexpf(mx, &e);
    const float sc = ldexpf(1.f, -e);
    if (threadIdx.x == 0) inv_scale[r] = ldexpf(1.f, e);
    for (int k = threadIdx.x; k < Kc; k += blockDim.x) {
        const float v = row[k] * sc;
        const H h = (H)v;
        const H l = (H)(v - (float)h);
        hi[(size_t)r * Kc + k] = h;
        lo[(size_t)r * Kc + k] = l;
    }
}

// ------------------------------------------------------------------------------------------
// Fine pass
// ------------------------------------------------------------------------------------------

struct FineGeom {
    int Norient, Npsi, n_subdir, n_subpsi, over_rot, over_trans, T;
    int ref2d;              // 2D references (bilinear slices, backrotate2D insertion)
    const double* dir_sc;   // [Nd*n_subdir][4] sin/cos of rot and tilt
    const double* psi_sc;   // [Npsi*n_subpsi][2] sin/cos of psi
    double pf;
};

__device__ __forceinline__ int find_particle(const long long* off, int B, long long i)
{
    int lo = 0, hi = B;   // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void fine_matrix(const FineGeom& g, int n, int ir, double* A, int* cls, int* idir)
{
    const int c = n / g.Norient, o = n - c * g.Norient;
    const int d = o / g.Npsi, ps = o - d * g.Npsi;
    const int sd = ir / g.n_subpsi, sp = ir - sd * g.n_subpsi;
    const double* a = g.dir_sc + ((size_t)d * g.n_subdir + sd) * 4;
    const double* p = g.psi_sc + ((size_t)ps * g.n_subpsi + sp) * 2;
    euler_from_sincos(a[0], a[1], a[2], a[3], p[0], p[1], A);
    *cls = c;
    *idir = d;
}

// Fine diff2 of every significant coarse pair as a batched complex GEMM on MFMA
// (getAllSquaredDifferences ipass 1, ml_optimiser.cpp:6922-7507; the reference GPU path runs it as
// a block-sparse split-TF32 GEMM, fine_splitImg_BCF_PMO.cuh:139-649):
//   diff2[(item), (ft)] = C_fine + n1(item) - G(item, ft),  G = sum_p Re(Z_p conj(AB_ft,p)),
// item = (particle, significant orientation, oversampled rotation), Z = F conj(minv ctf X) the
// item's slice times the weighted image, AB_ft the phase row of fine translation ft, n1 = 1/2
// sum minv ctf^2 |F|^2. One workgroup per tile of 16 items of one particle x up to 64 columns
// (the particle's significant coarse translations x their oversampled translations, dense over
// the tile): the pixels stream through LDS in chunks of 256, each chunk's Z rows projected by the
// 256 threads (one pixel each, all 16 rows), then the 4 waves each take a quarter of the chunk's
// K on MFMA (split-K inside the workgroup, summed in LDS at the end). MODE 0: FP32 operands on
// v_mfma_f32_16x16x4f32; MODE 1 / 2: BF16 / FP16 hi+lo operands, Al Bh + Ah Bl + Ah Bh on
// v_mfma_f32_16x16x32, each chunk's FP16 rows scaled by the power of two that maps their
// maximum into [0.5, 1) and the chunk's products scaled back (exactly) before they are summed.
// Only the significant pairs' diff2 are written; there is no limit on the current size.
#define FT_ROWS 16
#define FT_COLS 64
#define FT_KC 256
#define FT_MAX_T MAX_COARSE_T

struct FineTileParams {
    const long long* tile_off;          // [B+1] first tile of each particle
    int B;
    int ncb;                            // column blocks per particle: ceil(T * over_trans / FT_COLS)
    const long long* orient_off;        // [B+1]
    const int* orients;                 // significant (class, orientation) of each row
    const long long* orient_pair_start; // first pair of each significant orientation
    const long long* pair_off;          // [B+1]
    const int2* pairs;                  // (class orientation, coarse translation), CPU order
    const unsigned* tmask;              // [B][TW] significant coarse translations
    int TW;
    FineGeom g;
    const float2* const* vols;
    const int* pads;
    const int2* pix;
    const unsigned char* inproj;
    const float* minv;                  // [og][npix] diff2 Minvsigma2 (0 at the origin)
    const int* win;
    int npix;
    const float2* Fimg;
    const float* Fctf;
    int Pw;
    const float* abf;                   // MODE 0: [T_f][ldk] phase rows (re, im), 0 beyond npix
    const void* abh;                    // MODE 1/2: the same rows as hi / lo halves
    const void* abl;
    int ldk;
    const PartScalars* ps;
    float* fine_d2;
    unsigned int* fine_min;
    unsigned long long* queue;          // work-queue counter (zero at launch)
};

template <int MODE> struct FineHalf { typedef _Float16 type; };
template <> struct FineHalf<1> { typedef __bf16 type; };

template <class H>
__device__ __forceinline__ floatx4 fine_mfma32(typename Vec8<H>::type a, typename Vec8<H>::type b, floatx4 c);
template <>
__device__ __forceinline__ floatx4 fine_mfma32<__bf16>(Vec8<__bf16>::type a, Vec8<__bf16>::type b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ floatx4 fine_mfma32<_Float16>(Vec8<_Float16>::type a, Vec8<_Float16>::type b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// LDS of one tile: the chunk's Z rows (FP32 [16][2 KC + 4], or hi and lo halves [16][2 KC + 8]
// each), reused after the K loop for the 4 waves' partial products [4][16][128].
#define FT_ZLD32 (2 * FT_KC + 4)
#define FT_ZLD16 (2 * FT_KC + 8)
#define FT_SMEM_BYTES (4 * FT_ROWS * FT_COLS * 4 > 2 * FT_ROWS * FT_ZLD16 * 2 ? 4 * FT_ROWS * FT_COLS * 4 \
                                                                               : 2 * FT_ROWS * FT_ZLD16 * 2)

template <int MODE>
__device__ __forceinline__ void fine_tile(const FineTileParams& P, const long long tile)
{
    typedef typename FineHalf<MODE>::type H;
    typedef typename Vec8<H>::type v8;
    __shared__ __attribute__((aligned(16))) unsigned char smem[FT_SMEM_BYTES];
    __shared__ float s_geo[FT_ROWS][6];
    __shared__ const float2* s_vol[FT_ROWS];
    __shared__ int s_pad[FT_ROWS];
    __shared__ unsigned s_cmax[2][FT_ROWS];   // FP16: |Z| maxima of each row in the current / next chunk
    __shared__ double s_n1[FT_ROWS];
    __shared__ int s_rank[FT_MAX_T];
    __shared__ int s_slot_t[FT_MAX_T];
    __shared__ int s_nt;
    __shared__ float s_wmin[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const FineGeom& g = P.g;
    const int b = find_particle(P.tile_off, P.B, tile);
    const long long local = tile - P.tile_off[b];
    const int rt = (int)(local / P.ncb), cb = (int)(local - (long long)rt * P.ncb);

    // the particle's significant coarse translations: rank of each, and the translation of each rank
    const unsigned* tm = P.tmask + (size_t)b * P.TW;
    for (int t = tid; t < g.T; t += 256) {
        const int w = t >> 5, bit = t & 31;
        int r = __popc(tm[w] & ((1u << bit) - 1u));
        for (int q = 0; q < w; q++) r += __popc(tm[q]);
        const bool set = (tm[w] >> bit) & 1u;
        s_rank[t] = set ? r : -1;
        if (set) s_slot_t[r] = t;
    }
    if (tid == 0) {
        int nt = 0;
        for (int q = 0; q < P.TW; q++) nt += __popc(tm[q]);
        s_nt = nt;
    }
    __syncthreads();
    const int ncols = s_nt * g.over_trans;
    const int c0 = cb * FT_COLS;
    if (c0 >= ncols) return;   // uniform over the workgroup: this column block is empty
    const long long o0 = P.orient_off[b];
    const int n_items = (int)(P.orient_off[b + 1] - o0) * g.over_rot;
    const int r0 = rt * FT_ROWS;
    const int nrow = min(FT_ROWS, n_items - r0);
    if (tid < FT_ROWS) {
        s_n1[tid] = 0.;
        s_vol[tid] = nullptr;
        if (tid < nrow) {
            const int li = r0 + tid;
            const long long j = o0 + li / g.over_rot;
            const int ir = li % g.over_rot;
            double A[9], Ai[9];
            int c, idir;
            fine_matrix(g, P.orients[j], ir, A, &c, &idir);
            inv3(A, Ai);   // Ainv = A.inv() * pf (projector.cpp:530-531)
            s_geo[tid][0] = (float)(Ai[0] * g.pf);
            s_geo[tid][1] = (float)(Ai[1] * g.pf);
            s_geo[tid][2] = (float)(Ai[3] * g.pf);
            s_geo[tid][3] = (float)(Ai[4] * g.pf);
            s_geo[tid][4] = (float)(Ai[6] * g.pf);
            s_geo[tid][5] = (float)(Ai[7] * g.pf);
            s_vol[tid] = P.vols[c];
            s_pad[tid] = P.pads[c];
        }
    }
    __syncthreads();
    if (nrow <= 0) return;   // uniform over the workgroup
    const int ncol_tile = min(FT_COLS, ncols - c0);
    const int ntj = (ncol_tile + 15) >> 4;

    // the phase row of each of this lane's columns (column tiles j: column c0 + 16 j + lane % 16;
    // columns past the tile read row 0 and are never stored)
    int colrow[FT_COLS / 16];   // element offsets (T_f ldk < 2^31, checked by the host)
#pragma unroll
    for (int j = 0; j < FT_COLS / 16; j++) {
        const int col = c0 + j * 16 + (lane & 15);
        int ft = 0;
        if (col < ncols) ft = s_slot_t[col / g.over_trans] * g.over_trans + col % g.over_trans;
        colrow[j] = ft * P.ldk;
    }
    // two-level FP32 accumulation: acc holds one chunk's products (128 K values of this wave),
    // tot the running sum of the chunks, so no accumulator adds more than ~128 + npix/256 terms
    floatx4 acc[FT_COLS / 16], tot[FT_COLS / 16];
#pragma unroll
    for (int j = 0; j < FT_COLS / 16; j++) tot[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // per thread (one pixel per chunk) and row: FP32 partial sums of 1/2 minv ctf^2 |F|^2 over
    // the thread's npix / 256 pixels, summed over the workgroup in FP64
    float n1[FT_ROWS];
#pragma unroll
    for (int r = 0; r < FT_ROWS; r++) n1[r] = 0.f;
    const float* mv = P.minv + (size_t)P.ps[b].og * P.npix;
    float* Zs = reinterpret_cast<float*>(smem);
    H* Zh = reinterpret_cast<H*>(smem);
    H* Zl = Zh + FT_ROWS * FT_ZLD16;

    if (tid < 2 * FT_ROWS) s_cmax[tid >> 4][tid & 15] = 0u;
    __syncthreads();
    int chunk = 0;
    for (int k0 = 0; k0 < P.npix; k0 += FT_KC, chunk++) {
        const int p = k0 + tid;
        float yr = 0.f, yi = 0.f, half_mc2 = 0.f;
        float x = 0.f, y = 0.f;
        bool inp = false;
        if (p < P.npix) {
            const int2 xy = P.pix[p];
            x = (float)xy.x;
            y = (float)xy.y;
            inp = P.inproj[p] != 0;
            const size_t w = (size_t)b * P.Pw + P.win[p];
            const float2 X = P.Fimg[w];
            const float ctf = P.Fctf[w], m = mv[p];
            const float sy = m * ctf;
            yr = sy * X.x;
            yi = sy * X.y;
            half_mc2 = 0.5f * m * ctf * ctf;
        }
        float zrr[MODE == 2 ? FT_ROWS : 1], zir[MODE == 2 ? FT_ROWS : 1];
        unsigned* cmax = s_cmax[chunk & 1];
#pragma unroll
        for (int r = 0; r < FT_ROWS; r++) {
            float2 f = make_float2(0.f, 0.f);
            const float2* vol = s_vol[r];
            if (vol && inp) {
                float2 v;
                const float* a = s_geo[r];
                if (ref_sample(vol, s_pad[r], g.ref2d, a[0] * x + a[1] * y, a[2] * x + a[3] * y, a[4] * x + a[5] * y, &v))
                    f = v;
            }
            const float zr = f.x * yr + f.y * yi, zi = f.y * yr - f.x * yi;   // F conj(Y)
            n1[r] += half_mc2 * (f.x * f.x + f.y * f.y);
            if (MODE == 0) {
                *reinterpret_cast<float2*>(&Zs[r * FT_ZLD32 + 2 * tid]) = make_float2(zr, zi);
            } else if (MODE == 1) {   // BF16 keeps FP32's exponent range: no scale
                const H hr = (H)zr, hi_ = (H)zi;
                Zh[r * FT_ZLD16 + 2 * tid] = hr;
                Zh[r * FT_ZLD16 + 2 * tid + 1] = hi_;
                Zl[r * FT_ZLD16 + 2 * tid] = (H)(zr - (float)hr);
                Zl[r * FT_ZLD16 + 2 * tid + 1] = (H)(zi - (float)hi_);
            } else {   // FP16: the chunk's row maximum first (wave maxima, then LDS)
                zrr[r] = zr;
                zir[r] = zi;
                if (vol) {   // uniform: rows past the tile's items stay zero
                    float mx = fmaxf(fabsf(zr), fabsf(zi));
                    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
                    if (lane == 0) atomicMax(&cmax[r], __float_as_uint(mx));
                }
            }
        }
        if (MODE == 2) {
            __syncthreads();
            if (tid < FT_ROWS) s_cmax[(chunk + 1) & 1][tid] = 0u;   // the next chunk's maxima
            // each row of the chunk scaled by 2^-e so that its maximum lies in [0.5, 1)
#pragma unroll
            for (int r = 0; r < FT_ROWS; r++) {
                int e = 0;
                const float m = __uint_as_float(cmax[r]);
                if (m > 0.f) (void)frexpf(m, &e);
                const float sc = ldexpf(1.f, -e);
                const float vr = zrr[r] * sc, vi = zir[r] * sc;
                const H hr = (H)vr, hi_ = (H)vi;
                Zh[r * FT_ZLD16 + 2 * tid] = hr;
                Zh[r * FT_ZLD16 + 2 * tid + 1] = hi_;
                Zl[r * FT_ZLD16 + 2 * tid] = (H)(vr - (float)hr);
                Zl[r * FT_ZLD16 + 2 * tid + 1] = (H)(vi - (float)hi_);
            }
        }
        __syncthreads();
        // MFMA over this wave's quarter of the chunk's 2 KC real K values
#pragma unroll
        for (int j = 0; j < FT_COLS / 16; j++) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
        if (MODE == 0) {
#pragma unroll 1
            for (int s8 = 0; s8 < (2 * FT_KC) / 4 / 16; s8++) {
                const int kk = wave * (2 * FT_KC / 4) + s8 * 16 + 4 * (lane >> 4);
                const float4 a4 = *reinterpret_cast<const float4*>(&Zs[(lane & 15) * FT_ZLD32 + kk]);
#pragma unroll
                for (int j = 0; j < FT_COLS / 16; j++) {
                    if (j < ntj) {
                        const float4 b4 = *reinterpret_cast<const float4*>(P.abf + colrow[j] + 2 * k0 + kk);
                        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.x, b4.x, acc[j], 0, 0, 0);
                        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.y, b4.y, acc[j], 0, 0, 0);
                        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.z, b4.z, acc[j], 0, 0, 0);
                        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.w, b4.w, acc[j], 0, 0, 0);
                    }
                }
            }
        } else {
            const H* bh = reinterpret_cast<const H*>(P.abh);
            const H* bl = reinterpret_cast<const H*>(P.abl);
#pragma unroll 1
            for (int s32 = 0; s32 < (2 * FT_KC) / 4 / 32; s32++) {
                const int kk = wave * (2 * FT_KC / 4) + s32 * 32 + 8 * (lane >> 4);
                const v8 ah = *reinterpret_cast<const v8*>(&Zh[(lane & 15) * FT_ZLD16 + kk]);
                const v8 al = *reinterpret_cast<const v8*>(&Zl[(lane & 15) * FT_ZLD16 + kk]);
#pragma unroll
                for (int j = 0; j < FT_COLS / 16; j++) {
                    if (j < ntj) {
                        const v8 vbh = *reinterpret_cast<const v8*>(bh + colrow[j] + 2 * k0 + kk);
                        const v8 vbl = *reinterpret_cast<const v8*>(bl + colrow[j] + 2 * k0 + kk);
                        acc[j] = fine_mfma32<H>(al, vbh, acc[j]);
                        acc[j] = fine_mfma32<H>(ah, vbl, acc[j]);
                        acc[j] = fine_mfma32<H>(ah, vbh, acc[j]);
                    }
                }
            }
        }
        if (MODE == 2) {   // the chunk's products back to the rows' scale (powers of two: exact)
            float us[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                int e = 0;
                const float m = __uint_as_float(cmax[4 * (lane >> 4) + i]);
                if (m > 0.f) (void)frexpf(m, &e);
                us[i] = ldexpf(1.f, e);
            }
#pragma unroll
            for (int j = 0; j < FT_COLS / 16; j++)
#pragma unroll
                for (int i = 0; i < 4; i++) tot[j][i] += acc[j][i] * us[i];
        } else {
#pragma unroll
            for (int j = 0; j < FT_COLS / 16; j++) tot[j] += acc[j];
        }
        __syncthreads();
    }
    // n1 per row: workgroup sums
#pragma unroll
    for (int r = 0; r < FT_ROWS; r++) {
        double v = (double)n1[r];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) atomicAdd(&s_n1[r], v);
    }
    // the 4 waves' partial products (lane holds rows 4 (lane / 16) + i of column 16 j + lane % 16)
    float* red = reinterpret_cast<float*>(smem);
#pragma unroll
    for (int j = 0; j < FT_COLS / 16; j++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            red[(wave * FT_ROWS + 4 * (lane >> 4) + i) * FT_COLS + j * 16 + (lane & 15)] = tot[j][i];
    __syncthreads();
    for (int e = tid; e < FT_ROWS * FT_COLS; e += 256)
        red[e] = red[e] + red[FT_ROWS * FT_COLS + e] + red[2 * FT_ROWS * FT_COLS + e] + red[3 * FT_ROWS * FT_COLS + e];
    __syncthreads();
    // diff2 of the significant pairs of this tile's rows and columns
    float lmin = 3.0e38f;
    const int nover = g.over_rot * g.over_trans;
    const double Cf = P.ps[b].C_fine;
    for (int r = 0; r < nrow; r++) {
        const int li = r0 + r;
        const long long j = o0 + li / g.over_rot;
        const int ir = li % g.over_rot;
        const long long p0 = P.orient_pair_start[j];
        const long long p1 = (j + 1 < P.orient_off[b + 1]) ? P.orient_pair_start[j + 1] : P.pair_off[b + 1];
        const double base = Cf + s_n1[r];
        const int nq = (int)(p1 - p0) * g.over_trans;
        for (int q = tid; q < nq; q += 256) {
            const long long k = p0 + q / g.over_trans;
            const int iot = q % g.over_trans;
            const int col = s_rank[P.pairs[k].y] * g.over_trans + iot - c0;
            if (col < 0 || col >= ncol_tile) continue;
            const float d = (float)(base - (double)red[r * FT_COLS + col]);
            P.fine_d2[k * nover + ir * g.over_trans + iot] = d;
            lmin = fminf(lmin, d);
        }
    }
    for (int o = 32; o > 0; o >>= 1) lmin = fminf(lmin, __shfl_xor(lmin, o));
    if (lane == 0) s_wmin[wave] = lmin;
    __syncthreads();
    if (tid == 0) {
        const float m = fminf(fminf(s_wmin[0], s_wmin[1]), fminf(s_wmin[2], s_wmin[3]));
        if (m < 3.0e38f) atomicMin(&P.fine_min[b], float_to_ordered(m));
    }
}

// The fine pass of a batch as one persistent launch: each workgroup takes the next tile from
// the batch's work queue (P.queue, cleared by k_batch_plan) until the tiles, whose count the
// plan wrote to tile_off[B], run out. Particles with many significant samples simply own more
// tiles; the grid is sized to the device, not to the work.
template <int MODE>
__global__ __launch_bounds__(256) void k_fine_mfma(FineTileParams P)
{
    __shared__ long long s_tile;
    const long long total = P.tile_off[P.B];
    for (;;) {
        __syncthreads();   // every thread is done with the previous tile's LDS
        if (threadIdx.x == 0) s_tile = (long long)atomicAdd(P.queue, 1ull);
        __syncthreads();
        const long long tile = s_tile;
        if (tile >= total) break;
        fine_tile<MODE>(P, tile);
    }
}

// Significant coarse translations of each particle (bit t of tmask[b]) from its pair list.
__global__ void k_fine_tmask(const long long* __restrict__ pair_off, const int2* __restrict__ pairs, int B, int TW,
                             unsigned* __restrict__ tmask)
{
    __shared__ unsigned m[FT_MAX_T / 32];
    const int b = blockIdx.x;
    for (int w = threadIdx.x; w < TW; w += blockDim.x) m[w] = 0u;
    __syncthreads();
    for (long long k = pair_off[b] + threadIdx.x; k < pair_off[b + 1]; k += blockDim.x) {
        const int t = pairs[k].y;
        atomicOr(&m[t >> 5], 1u << (t & 31));
    }
    __syncthreads();
    for (int w = threadIdx.x; w < TW; w += blockDim.x) tmask[(size_t)b * TW + w] = m[w];
}

// Zero columns [n, ld) of each float2 row.
__global__ void k_zero_row_tails(float2* __restrict__ F, int n, int ld)
{
    for (int p = n + threadIdx.x; p < ld; p += blockDim.x) F[(size_t)blockIdx.x * ld + p] = make_float2(0.f, 0.f);
}

// |F|^2 of float2 rows (the n1 term's operand when the split rows are made from FP32 rows).
__global__ void k_abs2(const float2* __restrict__ F, long long n, float* __restrict__ F2)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float2 f = F[i];
        F2[i] = f.x * f.x + f.y * f.y;
    }
}

// Phase rows of the fine translations with K padded to ldk (multiple of 2 FT_KC): FP32 rows
// (MODE 0) or hi / lo halves (split modes; |values| <= 1, no scale).
template <class H>
__global__ void k_fine_ab_rows(const float2* __restrict__ ab, int npix, int ldk, float* __restrict__ out32,
                               H* __restrict__ hi, H* __restrict__ lo)
{
    const int t = blockIdx.x;
    for (int p = threadIdx.x; p < ldk / 2; p += blockDim.x) {
        const float2 v = p < npix ? ab[(size_t)t * npix + p] : make_float2(0.f, 0.f);
        const size_t o = (size_t)t * ldk + 2 * p;
        if (out32) {
            out32[o] = v.x;
            out32[o + 1] = v.y;
        } else {
            const H hr = (H)v.x, hi_ = (H)v.y;
            hi[o] = hr;
            hi[o + 1] = hi_;
            lo[o] = (H)(v.x - (float)hr);
            lo[o + 1] = (H)(v.y - (float)hi_);
        }
    }
}

// Local searches, coarse pass (FP32) as a GEMM over the batch's (particle, orientation) rows:
// G[(b,o)][t] = sum_p Re(Z_bo,p conj(AB_t,p)) with Z_bo = F_o conj(minv ctf X_b), i.e. the real
// dot product of [Z.re, Z.im] with [AB.re, AB.im] over K = 2 P_c; the phase table AB is the
// shared operand. k_local_zrows forms Z from the row's slice (projected once per batch by
// k_project_list) and n1 = 1/2 sum minv ctf^2 |F|^2; k_local_diff2 forms
// diff2 = C_coarse + n1 - G into the sample list (orientation-major, T per orientation).
// The batch's distinct in-range orientations projected once (exact FP32 slices at coarse size,
// XCD-aware order as k_local_zrows), shared by every particle whose range holds them. SPLIT 0:
// float2 rows F (the FP32 path); SPLIT 1: the rows straight into the split operand (hi, lo and a
// power-of-two row scale for FP16, as k_split_rows) plus |F|^2 for the n1 term.
template <class H, int SPLIT>
__global__ __launch_bounds__(256) void k_project_list(long long nu, const int* __restrict__ ulist, FineGeom g,
                                                      const float2* const* __restrict__ vols,
                                                      const int* __restrict__ pads, const int2* __restrict__ pix,
                                                      const unsigned char* __restrict__ inproj, int npix, int ldf,
                                                      float2* __restrict__ F, H* __restrict__ fh, H* __restrict__ fl,
                                                      float* __restrict__ fs, int use_scale, float* __restrict__ F2)
{
    extern __shared__ float frow[];   // [2 ldf] (split only)
    __shared__ float redf[4];
    const long long per = (long long)gridDim.x / 8;
    const long long u = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (u >= nu) return;
    double A[9], Ai[9];
    int c, idir;
    fine_matrix(g, ulist[u], 0, A, &c, &idir);
    inv3(A, Ai);
    const float a0 = (float)(Ai[0] * g.pf), a1 = (float)(Ai[1] * g.pf);
    const float a2 = (float)(Ai[3] * g.pf), a3 = (float)(Ai[4] * g.pf);
    const float a4 = (float)(Ai[6] * g.pf), a5 = (float)(Ai[7] * g.pf);
    const float2* vol = vols[c];
    const int pad = pads[c];
    float mx = 0.f;
    for (int p = threadIdx.x; p < ldf; p += blockDim.x) {
        float2 f = make_float2(0.f, 0.f);
        if (p < npix && inproj[p]) {
            const float x = (float)pix[p].x, y = (float)pix[p].y;
            float2 v;
            if (ref_sample(vol, pad, g.ref2d, a0 * x + a1 * y, a2 * x + a3 * y, a4 * x + a5 * y, &v)) f = v;
        }
        if (SPLIT == 2) {   // BF16: FP32's exponent range, no row scale: the halves straight out
            const H hx = (H)f.x, hy = (H)f.y;
            fh[u * 2 * ldf + 2 * p] = hx;
            fh[u * 2 * ldf + 2 * p + 1] = hy;
            fl[u * 2 * ldf + 2 * p] = (H)(f.x - (float)hx);
            fl[u * 2 * ldf + 2 * p + 1] = (H)(f.y - (float)hy);
            F2[u * ldf + p] = f.x * f.x + f.y * f.y;
        } else if (SPLIT) {
            frow[2 * p] = f.x;
            frow[2 * p + 1] = f.y;
            mx = fmaxf(mx, fmaxf(fabsf(f.x), fabsf(f.y)));
            F2[u * ldf + p] = f.x * f.x + f.y * f.y;   // |F|^2 for the n1 term
        } else if (p < npix) {
            F[u * ldf + p] = f;
        }
    }
    if (SPLIT == 2) {
        if (threadIdx.x == 0) fs[u] = 1.f;
    } else if (SPLIT) {
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        if ((threadIdx.x & 63) == 0) redf[threadIdx.x >> 6] = mx;
        __syncthreads();
        float sc = 1.f;
        if (use_scale) {
            mx = fmaxf(fmaxf(redf[0], redf[1]), fmaxf(redf[2], redf[3]));
            int e = 0;
            if (mx > 0.f) (void)frexpf(mx, &e);
            sc = ldexpf(1.f, -e);
            if (threadIdx.x == 0) fs[u] = ldexpf(1.f, e);
        } else if (threadIdx.x == 0) {
            fs[u] = 1.f;
        }
        const int Kc = 2 * ldf;
        for (int k = threadIdx.x; k < Kc; k += blockDim.x) {
            const float v = frow[k] * sc;
            const H h = (H)v;
            const H l = (H)(v - (float)h);
            fh[u * Kc + k] = h;
            fl[u * Kc + k] = l;
        }
    }
}

// n1 of every (particle, orientation) row: 1/2 sum_p minv ctf^2 |F|^2 = sum_p A1_b,p |F_u,p|^2 in
// FP64 (one wave per row).
__global__ __launch_bounds__(256) void k_local_n1(int B, long long nrows, const long long* __restrict__ orient_off,
                                                  const int* __restrict__ rowu, const float* __restrict__ F2, int ldf,
                                                  const float* __restrict__ A1, long long lda1, int npix,
                                                  double* __restrict__ n1out)
{
    const long long j = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= nrows) return;
    const int b = find_particle(orient_off, B, j);
    const float* f2 = F2 + (long long)rowu[j] * ldf;
    const float* a1 = A1 + (long long)b * lda1;
    double acc = 0.;
    for (int p = lane; p < npix; p += 64) acc += (double)a1[p] * (double)f2[p];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) n1out[j] = acc;
}

__global__ __launch_bounds__(256) void k_local_zrows(int B, long long nrows, const long long* __restrict__ orient_off,
                                                     const int* __restrict__ rowu, const float2* __restrict__ F,
                                                     int ldf, const float* __restrict__ minv,
                                                     const int* __restrict__ win, int npix,
                                                     const PartScalars* __restrict__ ps,
                                                     const float2* __restrict__ Fimg,
                                                     const float* __restrict__ Fctf, int Pw, int Kc,
                                                     float* __restrict__ zf, double* __restrict__ n1out)
{
    __shared__ double red[4];
    // XCD-aware row order: workgroups are dealt round-robin to the 8 XCDs, so XCD k is given the
    // contiguous rows [k R/8, (k+1) R/8) (the grid is a multiple of 8)
    const long long per = (long long)gridDim.x / 8;
    const long long j = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
    const bool valid = j < nrows;
    double n1 = 0.;
    if (valid) {
        const int b = find_particle(orient_off, B, j);
        const float2* Fu = F + (long long)rowu[j] * ldf;
        const float* mv = minv + (size_t)ps[b].og * npix;
        for (int p = threadIdx.x; p < Kc / 2; p += blockDim.x) {
            float zr = 0.f, zi = 0.f;
            if (p < npix) {
                const float2 f = Fu[p];
                const size_t w = (size_t)b * Pw + win[p];
                const float2 X = Fimg[w];
                const float ctf = Fctf[w], m = mv[p];
                const float s = m * ctf;
                const float yr = s * X.x, yi = s * X.y;
                zr = f.x * yr + f.y * yi;
                zi = f.y * yr - f.x * yi;
                n1 += 0.5 * (double)m * (double)ctf * ctf * ((double)f.x * f.x + (double)f.y * f.y);
            }
            zf[j * Kc + 2 * p] = zr;
            zf[j * Kc + 2 * p + 1] = zi;
        }
    } else {
        for (int k = threadIdx.x; k < Kc; k += blockDim.x) zf[j * Kc + k] = 0.f;
    }
    for (int o = 32; o > 0; o >>= 1) n1 += __shfl_xor(n1, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = n1;
    __syncthreads();
    if (valid && threadIdx.x == 0) n1out[j] = red[0] + red[1] + red[2] + red[3];
}

// The phase table as the GEMM's shared operand: row t = [re, im] of AB_t over the coarse pixels,
// zero beyond 2 P_c and for rows T .. Tpad-1.
__global__ void k_ab_rows(const float2* __restrict__ ab, int T, int npix, int Kc, float* __restrict__ out)
{
    const int t = blockIdx.x;
    for (int p = threadIdx.x; p < Kc / 2; p += blockDim.x) {
        float2 v = make_float2(0.f, 0.f);
        if (t < T && p < npix) v = ab[(size_t)t * npix + p];
        out[(size_t)t * Kc + 2 * p] = v.x;
        out[(size_t)t * Kc + 2 * p + 1] = v.y;
    }
}

// diff2 = C_coarse + n1 - G for every (row, translation); per-particle minimum.
__global__ __launch_bounds__(128) void k_local_diff2(int B, long long nrows, const long long* __restrict__ orient_off,
                                                     const PartScalars* __restrict__ ps,
                                                     const double* __restrict__ n1, const float* __restrict__ G,
                                                     int ldg, int T, float* __restrict__ d2,
                                                     unsigned int* __restrict__ dmin)
{
    __shared__ float wm[2];
    const long long j = blockIdx.x;
    if (j >= nrows) return;
    const int b = find_particle(orient_off, B, j);
    const double base = ps[b].C_coarse + n1[j];
    float m = 3.0e38f;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        const float d = (float)(base - (double)G[j * ldg + t]);
        d2[j * T + t] = d;
        m = fminf(m, d);
    }
    for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float mm = (blockDim.x > 64) ? fminf(wm[0], wm[1]) : wm[0];
        if (mm < 3.0e38f) atomicMin(&dmin[b], float_to_ordered(mm));
    }
}

// Which pixels backproject2Dto3D touches (backprojector.cpp:114-170): the ellipse bounds of
// the row and the r^2 test, evaluated in double with the reference's operation order and no
// contraction so that pixels on the r_max sphere are kept or dropped exactly as there.
__device__ __noinline__ bool bp_pixel_inside(double AtA_xx, double AtA_xy, double AtA_yy,
                                             double AtA_xy2, int max_r2, double a0, double a1,
                                             double a2, double a3, double a4, double a5, int x, int y,
                                             int ref2d)
{
#pragma clang fp contract(off)
    double discr = AtA_xy2 * y * y - AtA_xx * (AtA_yy * y * y - max_r2);
    if (discr < 0.0) return false;
    double dd = sqrt(discr) / AtA_xx;
    double qq = -AtA_xy * y / AtA_xx;
    double lo = qq - dd, hi = qq + dd;
    int first_x = (lo == (int)lo) ? (int)lo : ((lo > 0) ? (int)(lo + 1) : (int)lo);
    int last_x = (hi == (int)hi) ? (int)hi : ((hi > 0) ? (int)hi : (int)(hi - 1));
    if (x < first_x || x > last_x) return false;
    if (ref2d) return true;   // backrotate2D has no further radius test (backprojector.cpp:568-596)
    double xp = a0 * x + a1 * y, yp = a2 * x + a3 * y, zp = a4 * x + a5 * y;
    return !(xp * xp + yp * yp + zp * zp > max_r2);
}

__device__ __noinline__ void bp_ata(double a0, double a1, double a2, double a3, double a4, double a5,
                                    double* xx, double* xy, double* yy, double* xy2)
{
#pragma clang fp contract(off)
    *xx = a0 * a0 + a2 * a2 + a4 * a4;
    *xy = a0 * a1 + a2 * a3 + a4 * a5;
    *yy = a1 * a1 + a3 * a3 + a5 * a5;
    *xy2 = *xy * *xy;
}

// ------------------------------------------------------------------------------------------
// Weighted sums and backprojection (storeWeightedSums, ml_optimiser.cpp:8043-8992;
// BackProjector::backproject2Dto3D, backprojector.cpp:55-357)
// ------------------------------------------------------------------------------------------

struct StoreParams {
    FineGeom g;
    const float2* const* vols;
    const int* pads;
    float2* const* bp_data;
    float* const* bp_weight;
    const int* bp_pads;
    const int* bp_rmax;
    const double* prior_o;
    const double* prior_t;
    int tK;                    // > 0: prior_t per class (trans_prior), cls_prior holds the classes' offsets
    const double* cls_prior;   // [2K] px, 2D references with class-dependent offset priors
    const int2* pix;
    const unsigned char* inproj;
    const float* minv_store;   // [og][npix] Minvsigma2 with the origin term (ml_optimiser.cpp:8099-8104)
    const int* win;
    const int* ires;
    int npix;                  // store pixel set: every pixel with ires >= 0
    const float2* Fimg;
    const float2* Fimg_nm;
    const float* Fctf;
    int Pw;
    const float2* ab_fine;     // [T*over_trans][npix] (same pixel order)
    const double* ftx;
    const double* fty;         // fine translations in px
    const PartScalars* ps;
    const SelResult* fine_res;
    const float* fine_d2;
    double pixel_size;
    int nshell;                // N/2+1
    int cs;
    const double* dvp;         // [K][nshell] data_vs_prior_class (scale correction), or null
    // accumulators
    double* w_sigma2;          // [og][nshell]
    double* w_sumw;            // [og]
    double* w_pdf_class;       // [K]
    double* w_pdf_dir;         // [K*Nd]
    double* w_sigma2_offset;   // [1]
    double* w_prior_offset;    // [2K] 2D references: sum w * pixel * (old offset + translation)
    double* w_norm;            // [B] per particle sum of wdiff2
    double* w_xa;              // [B] per particle scale sums XA and AA (ml_optimiser.cpp:8477-8486)
    double* w_aa;
    int Nd;
    LocalPrior lp;             // local searches (lp.dir null: global)
    // brick backprojection (3D references; bp_val null: every pixel scattered with global
    // atomics): per item its plane (unit normal, class), its frame and its pixels' values
    float4* bp_filt;           // [item_cap] (nx, ny, nz, class as int bits; -1: nothing to insert)
    double* bp_geom;           // [item_cap][6] pf * Ainv (da0..da5)
    float4* bp_val;            // [item_cap][bp_wy][bp_wx] by frequency (x, y + bp_wy0): (data re,
                               // data im, weight, 0); 0 where not inserted (or not a pixel)
    int bp_wx, bp_wy, bp_wy0;
    long long item_cap;        // items from item_cap on take the direct scatter
};

#define STORE_MAX_T 2048

__device__ __forceinline__ void store_item(const long long* __restrict__ item_off, int B,
                                           const long long* __restrict__ orient_off,
                                           const int* __restrict__ orients,
                                           const long long* __restrict__ orient_pair_start,
                                           const long long* __restrict__ pair_off,
                                           const int2* __restrict__ pairs, const StoreParams& P,
                                           const long long item)
{
    __shared__ int s_ft[STORE_MAX_T];
    __shared__ float s_w[STORE_MAX_T];
    __shared__ int s_cnt;
    __shared__ double shell[520];
    __shared__ double red[4];
    const FineGeom& g = P.g;
    const int b = find_particle(item_off, B, item);
    const long long local = item - item_off[b];
    const long long j = orient_off[b] + local / g.over_rot;
    const int ir = (int)(local % g.over_rot);
    const int n = orients[j];
    const long long p0 = orient_pair_start[j];
    const long long p1 = (j + 1 < orient_off[b + 1]) ? orient_pair_start[j + 1] : pair_off[b + 1];
    const double sumw = P.fine_res[b].sum_w, sig = P.fine_res[b].sig_w, mind2 = P.fine_res[b].min_d2;
    const int nover = g.over_rot * g.over_trans;
    const int og = P.ps[b].og;
    const bool brick = P.bp_val && item < P.item_cap;
    if (brick && threadIdx.x == 0) P.bp_filt[item] = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
    if (threadIdx.x == 0) s_cnt = 0;
    for (int i = threadIdx.x; i < P.nshell; i += blockDim.x) shell[i] = 0.;
    __syncthreads();
    // significant oversampled translations of this oversampled rotation (in CPU order)
    const int nt = (int)(p1 - p0) * g.over_trans;
    for (int q = threadIdx.x; q < nt; q += blockDim.x) {
        const long long k = p0 + q / g.over_trans;
        const int iot = q % g.over_trans;
        const int2 nt2 = pairs[k];
        double prior = orient_prior(P.prior_o, P.lp, b, nt2.x) *
                       trans_prior(P.prior_t, g.T, P.tK, g.Norient, b, nt2.x, nt2.y);
        double w = 0.;
        if (prior > 0.) {
            double dd = (double)P.fine_d2[k * nover + ir * g.over_trans + iot] - mind2;
            w = (dd > 700.) ? 0. : prior * exp(-dd);
        }
        if (w >= sig && w > 0.) {
            int slot = atomicAdd(&s_cnt, 1);
            if (slot < STORE_MAX_T) {
                s_ft[slot] = nt2.y * g.over_trans + iot;
                s_w[slot] = (float)(w / sumw);
            }
        }
    }
    __syncthreads();
    const int cnt = min(s_cnt, STORE_MAX_T);
    if (cnt == 0) return;
    double A[9], Ai[9];
    int c, idir;
    fine_matrix(g, n, ir, A, &c, &idir);
    inv3(A, Ai);
    // scalar sums (ml_optimiser.cpp:8505-8552), exact weights recomputed in double
    if (threadIdx.x == 0) {
        double sw = 0., soff = 0., pox = 0., poy = 0.;
        const double px2 = P.pixel_size * P.pixel_size;
        const PartScalars s = P.ps[b];
        const double qx = P.cls_prior ? P.cls_prior[2 * c] : s.prior_x;
        const double qy = P.cls_prior ? P.cls_prior[2 * c + 1] : s.prior_y;
        for (int q = 0; q < nt; q++) {
            const long long k = p0 + q / g.over_trans;
            const int iot = q % g.over_trans;
            const int2 nt2 = pairs[k];
            double prior = orient_prior(P.prior_o, P.lp, b, nt2.x) *
                           trans_prior(P.prior_t, g.T, P.tK, g.Norient, b, nt2.x, nt2.y);
            if (!(prior > 0.)) continue;
            double dd = (double)P.fine_d2[k * nover + ir * g.over_trans + iot] - mind2;
            double w = (dd > 700.) ? 0. : prior * exp(-dd);
            if (!(w >= sig && w > 0.)) continue;
            w /= sumw;
            const int ft = nt2.y * g.over_trans + iot;
            sw += w;
            double dx = qx - s.old_x - P.ftx[ft];
            double dy = qy - s.old_y - P.fty[ft];
            soff += w * px2 * dx * dx + w * px2 * dy * dy;
            pox += w * P.pixel_size * (s.old_x + P.ftx[ft]);   // ml_optimiser.cpp:8504-8508
            poy += w * P.pixel_size * (s.old_y + P.fty[ft]);
        }
        atomicAdd(&P.w_sumw[og], sw);
        atomicAdd(&P.w_pdf_class[c], sw);
        atomicAdd(&P.w_pdf_dir[(size_t)c * P.Nd + idir], sw);
        atomicAdd(P.w_sigma2_offset, soff);
        if (g.ref2d) {
            atomicAdd(&P.w_prior_offset[2 * c], pox);
            atomicAdd(&P.w_prior_offset[2 * c + 1], poy);
        }
    }
    // projection frame: Ainv = A.inv() * pf (float for sampling, double for the scatter geometry)
    const double pf = g.pf;
    const double da0 = Ai[0] * pf, da1 = Ai[1] * pf, da2 = Ai[3] * pf, da3 = Ai[4] * pf, da4 = Ai[6] * pf, da5 = Ai[7] * pf;
    const float fa0 = (float)da0, fa1 = (float)da1, fa2 = (float)da2;
    const float fa3 = (float)da3, fa4 = (float)da4, fa5 = (float)da5;
    if (brick && threadIdx.x == 0) {
        double* gm = P.bp_geom + (size_t)item * 6;
        gm[0] = da0; gm[1] = da1; gm[2] = da2; gm[3] = da3; gm[4] = da4; gm[5] = da5;
        // the plane of the slice: normal (da0, da2, da4) x (da1, da3, da5)
        const double nx = da2 * da5 - da4 * da3, ny = da4 * da1 - da0 * da5, nz = da0 * da3 - da2 * da1;
        const double nn = 1. / sqrt(nx * nx + ny * ny + nz * nz);
        P.bp_filt[item] = make_float4((float)(nx * nn), (float)(ny * nn), (float)(nz * nn), __int_as_float(c));
    }
    const float2* vol = P.vols[c];
    const int pad = P.pads[c];
    float2* bpd = P.bp_data[c];
    float* bpw = P.bp_weight[c];
    const int bpad = P.bp_pads[c];
    const double bpf = pf;
    // backproject2Dto3D rounds r_max * pf, backrotate2D truncates it (`const int r_max_ref`)
    const int rr = g.ref2d ? (int)(P.bp_rmax[c] * bpf) : relion_round(P.bp_rmax[c] * bpf);
    const int max_r2 = rr * rr;
    const int bxdim = bpad / 2 + 1, bh = bpad / 2;
    // ellipse row bounds of backproject2Dto3D (backprojector.cpp:80-135) in double
    double AtA_xx, AtA_xy, AtA_yy, AtA_xy2;
    bp_ata(da0, da1, da2, da3, da4, da5, &AtA_xx, &AtA_xy, &AtA_yy, &AtA_xy2);
    double norm_acc = 0., xa_acc = 0., aa_acc = 0.;
    const float ctf_f = (float)P.ps[b].ctf_store;
    const float* minv_og = P.minv_store + (size_t)og * P.npix;
    const double* dvp_c = P.dvp ? P.dvp + (size_t)c * P.nshell : nullptr;
    for (int p = threadIdx.x; p < P.npix; p += blockDim.x) {
        const int x = P.pix[p].x, y = P.pix[p].y;
        float2 f = make_float2(0.f, 0.f);
        if (P.inproj[p]) {
            float2 v;
            if (ref_sample(vol, pad, g.ref2d, fa0 * x + fa1 * y, fa2 * x + fa3 * y, fa4 * x + fa5 * y, &v)) f = v;
        }
        const size_t wi = (size_t)b * P.Pw + P.win[p];
        const float ctf = P.Fctf[wi] * ctf_f;            // CTF x the group's (clamped) scale
        const float2 X = P.Fimg[wi], Xn = P.Fimg_nm[wi];
        const float fr = f.x * ctf, fi = f.y * ctf;     // Frefctf
        const float ms = minv_og[p];
        const int ir_ = P.ires[p];
        const bool scale_sums = dvp_c && ir_ >= 0 && dvp_c[ir_] > 3.;
        float d2acc = 0.f, accr = 0.f, acci = 0.f, wacc = 0.f, xacc = 0.f, wsum_q = 0.f;
        for (int q = 0; q < cnt; q++) {
            const float w = s_w[q];
            const float2 s = P.ab_fine[(size_t)s_ft[q] * P.npix + p];
            const float xr = X.x * s.x - X.y * s.y, xi = X.x * s.y + X.y * s.x;
            const float dr = fr - xr, di = fi - xi;
            d2acc += w * (dr * dr + di * di);
            if (scale_sums) {
                xacc += w * (fr * xr + fi * xi);
                wsum_q += w;
            }
            const float nr = Xn.x * s.x - Xn.y * s.y, ni = Xn.x * s.y + Xn.y * s.x;
            const float wx = w * ctf * ms;
            accr += nr * wx;
            acci += ni * wx;
            wacc += wx * ctf;
        }
        if (ir_ >= 0) {
            atomicAdd(&shell[ir_], (double)d2acc);
            norm_acc += d2acc;
        }
        if (scale_sums) {
            xa_acc += (double)xacc;
            aa_acc += (double)wsum_q * ((double)fr * fr + (double)fi * fi);
        }
        if (brick) {   // inserted by k_bp_brick
            const bool ins = (wacc > 0.f) &&
                             bp_pixel_inside(AtA_xx, AtA_xy, AtA_yy, AtA_xy2, max_r2, da0, da1, da2, da3, da4, da5, x, y,
                                             g.ref2d);
            P.bp_val[((size_t)item * P.bp_wy + (y + P.bp_wy0)) * P.bp_wx + x] =
                ins ? make_float4(accr, acci, wacc, 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        if (!(wacc > 0.f)) continue;
        if (!bp_pixel_inside(AtA_xx, AtA_xy, AtA_yy, AtA_xy2, max_r2, da0, da1, da2, da3, da4, da5, x, y, g.ref2d))
            continue;
        if (g.ref2d) {
            // BackProjector::backrotate2D (backprojector.cpp:568-643): bilinear insertion
            double xp, yp;
            {
#pragma clang fp contract(off)
                xp = da0 * x + da1 * y;
                yp = da2 * x + da3 * y;
            }
            bool neg = xp < 0;
            if (neg) { xp = -xp; yp = -yp; }
            double fx0 = floor(xp), fy0 = floor(yp);
            int x0 = (int)fx0, y0 = (int)fy0 + bh;
            if (x0 < 0 || x0 + 1 >= bxdim || y0 < 0 || y0 + 1 >= bpad) continue;
            double fx = xp - fx0, fy = yp - fy0;
            double mfx = 1. - fx, mfy = 1. - fy;
            float vr = accr, vi = neg ? -acci : acci;
            const size_t sy = bxdim, base = (size_t)y0 * sy + x0;
            const double wts[4] = {mfy * mfx, mfy * fx, fy * mfx, fy * fx};
            const size_t offs[4] = {0, 1, sy, sy + 1};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float wk = (float)wts[k];
                atomicAdd(&bpd[base + offs[k]].x, wk * vr);
                atomicAdd(&bpd[base + offs[k]].y, wk * vi);
                atomicAdd(&bpw[base + offs[k]], wk * wacc);
            }
            continue;
        }
        double xp, yp, zp;
        {
#pragma clang fp contract(off)
            xp = da0 * x + da1 * y;
            yp = da2 * x + da3 * y;
            zp = da4 * x + da5 * y;
        }
        bool neg = xp < 0;
        if (neg) { xp = -xp; yp = -yp; zp = -zp; }
        double fx0 = floor(xp), fy0 = floor(yp), fz0 = floor(zp);
        int x0 = (int)fx0, y0 = (int)fy0 + bh, z0 = (int)fz0 + bh;
        if (x0 < 0 || x0 + 1 >= bxdim || y0 < 0 || y0 + 1 >= bpad || z0 < 0 || z0 + 1 >= bpad) continue;
        double fx = xp - fx0, fy = yp - fy0, fz = zp - fz0;
        double mfx = 1. - fx, mfy = 1. - fy, mfz = 1. - fz;
        float vr = accr, vi = neg ? -acci : acci;
        const size_t sy = bxdim, sz = (size_t)bpad * bxdim;
        const size_t base = (size_t)z0 * sz + (size_t)y0 * sy + x0;
        const double wts[8] = {mfz * mfy * mfx, mfz * mfy * fx, mfz * fy * mfx, mfz * fy * fx,
                               fz * mfy * mfx, fz * mfy * fx, fz * fy * mfx, fz * fy * fx};
        const size_t offs[8] = {0, 1, sy, sy + 1, sz, sz + 1, sz + sy, sz + sy + 1};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float wk = (float)wts[k];
            atomicAdd(&bpd[base + offs[k]].x, wk * vr);
            atomicAdd(&bpd[base + offs[k]].y, wk * vi);
            atomicAdd(&bpw[base + offs[k]], wk * wacc);
        }
    }
    for (int o = 32; o > 0; o >>= 1) norm_acc += __shfl_xor(norm_acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = norm_acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&P.w_norm[b], red[0] + red[1] + red[2] + red[3]);
    if (dvp_c) {
        __shared__ double red2[8];
        __syncthreads();
        for (int o = 32; o > 0; o >>= 1) { xa_acc += __shfl_xor(xa_acc, o); aa_acc += __shfl_xor(aa_acc, o); }
        if ((threadIdx.x & 63) == 0) { red2[threadIdx.x >> 6] = xa_acc; red2[4 + (threadIdx.x >> 6)] = aa_acc; }
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(&P.w_xa[b], red2[0] + red2[1] + red2[2] + red2[3]);
            atomicAdd(&P.w_aa[b], red2[4] + red2[5] + red2[6] + red2[7]);
        }
    }
    double* w_s2 = P.w_sigma2 + (size_t)og * P.nshell;
    for (int i = threadIdx.x; i < P.nshell; i += blockDim.x)
        if (shell[i] != 0.) atomicAdd(&w_s2[i], shell[i]);
}

// Weighted sums and backprojection of a batch as one persistent launch: each workgroup takes the
// next (significant orientation, oversampled rotation) item from the work queue (cleared by
// k_batch_plan) until the batch's items (item_off[B]) run out.
__global__ __launch_bounds__(256) void k_store(const long long* __restrict__ item_off, int B,
                                               const long long* __restrict__ orient_off,
                                               const int* __restrict__ orients,
                                               const long long* __restrict__ orient_pair_start,
                                               const long long* __restrict__ pair_off,
                                               const int2* __restrict__ pairs, StoreParams P,
                                               unsigned long long* queue)
{
    __shared__ long long s_item;
    const long long total = item_off[B];
    for (;;) {
        __syncthreads();   // every thread is done with the previous item's LDS
        if (threadIdx.x == 0) s_item = (long long)atomicAdd(queue, 1ull);
        __syncthreads();
        const long long item = s_item;
        if (item >= total) break;
        store_item(item_off, B, orient_off, orients, orient_pair_start, pair_off, pairs, P, item);
    }
}

// Backprojection through LDS bricks. In brick mode k_store leaves each item's frame and
// per-pixel values (data re/im and weight: the item's translations summed, before the insertion
// of backproject2Dto3D / backrotate2D, backprojector.cpp:136-357 / 568-643) in bp_geom / bp_val,
// and k_bp_brick inserts them: one workgroup per unit (a brick of one class's volume, BP_S3^3
// voxels for 3D references, a tile of BP_S2^2 for 2D ones, and a share of the batch's items),
// which collects the items that can reach the brick (3D: whose plane crosses it; 2D: every item
// of the class), inserts every pixel of theirs whose interpolation footprint meets the brick into
// an LDS copy of it (LDS atomics), and adds the brick to the volume at the end: a plain
// read-modify-write when one unit covers the brick, global atomics when its items are shared
// among several. The per-pixel arithmetic (positions in double without contraction, the Friedel
// flip, the bounds and the trilinear / bilinear weights) is the direct scatter's; only the order
// of the FP32 sums differs. A wave takes one item at a time and walks the pixels whose positions
// can fall in the brick: the box of their frequencies (x, y) = M^T q / |m0|^2 (M = pf * the
// inverse rotation's first columns, m0 its first column) over the brick's corners, and its mirror
// image for the pixels inserted at -q.
#define BP_S3 16
#define BP_S2 32
#define BP_LIST 1024
#define BP_THREADS 512

template <int D3> struct BpDims {
    static constexpr int S = D3 ? BP_S3 : BP_S2;
    static constexpr int VOX = D3 ? BP_S3 * BP_S3 * BP_S3 : BP_S2 * BP_S2;
};

struct BpBrickParams {
    const long long* item_off;
    int B;
    long long item_cap;
    const float4* filt;
    const double* geom;
    const float4* val;         // [item][pyh][pxw]: the values of frequency (x, y) at row y + py0
    int py0, pxw, pyh;
    float2* const* bp_data;
    float* const* bp_weight;
    const int* bp_pads;
    const int* cls_off;        // [K+1] the batch's items grouped by class (k_bp_bucket)
    const int* cls_items;
};

// A brick of a class's volume within reach of the insertion radius. Its items are split into
// G shares (units of work) by the batch's count of items of that class: G = ceil(reach * count /
// unit_items), reach the expected fraction of the class's items that reach the brick.
struct BpBrick {
    int c, bx, by, bz;   // class, brick origin / brick side in voxel index space (x, y, z; 2D: z 0)
    float reach;
    int pad0, pad1, pad2;
};

template <int D3>
__device__ __forceinline__ void bp_insert_item(const BpBrickParams& P, int item, int bxdim, int bpad, int bh,
                                               const int (&blo)[3], const double (&lo)[3],
                                               const double (&hi)[3], float* s_re, float* s_im, float* s_w,
                                               int lane)
{
    constexpr int S = BpDims<D3>::S;
    const double* gm = P.geom + (size_t)item * 6;
    const double a0 = gm[0], a1 = gm[1], a2 = gm[2], a3 = gm[3], a4 = gm[4], a5 = gm[5];
    double s2, u_lo, u_hi, v_lo, v_hi;
    if constexpr (D3) {
        s2 = a0 * a0 + a2 * a2 + a4 * a4;
        u_lo = fmin(a0 * lo[0], a0 * hi[0]) + fmin(a2 * lo[1], a2 * hi[1]) + fmin(a4 * lo[2], a4 * hi[2]);
        u_hi = fmax(a0 * lo[0], a0 * hi[0]) + fmax(a2 * lo[1], a2 * hi[1]) + fmax(a4 * lo[2], a4 * hi[2]);
        v_lo = fmin(a1 * lo[0], a1 * hi[0]) + fmin(a3 * lo[1], a3 * hi[1]) + fmin(a5 * lo[2], a5 * hi[2]);
        v_hi = fmax(a1 * lo[0], a1 * hi[0]) + fmax(a3 * lo[1], a3 * hi[1]) + fmax(a5 * lo[2], a5 * hi[2]);
    } else {
        s2 = a0 * a0 + a2 * a2;
        u_lo = fmin(a0 * lo[0], a0 * hi[0]) + fmin(a2 * lo[1], a2 * hi[1]);
        u_hi = fmax(a0 * lo[0], a0 * hi[0]) + fmax(a2 * lo[1], a2 * hi[1]);
        v_lo = fmin(a1 * lo[0], a1 * hi[0]) + fmin(a3 * lo[1], a3 * hi[1]);
        v_hi = fmax(a1 * lo[0], a1 * hi[0]) + fmax(a3 * lo[1], a3 * hi[1]);
    }
    u_lo /= s2; u_hi /= s2; v_lo /= s2; v_hi /= s2;
    const double e = 1e-6;
    const int ymin = -P.py0, ymax = P.pyh - 1 - P.py0;
    // pixels inserted at +q (xp >= 0), then at -q (xp < 0)
    const int ax0 = max(0, (int)ceil(u_lo - e)), ax1 = min(P.pxw - 1, (int)floor(u_hi + e));
    const int ay0 = max(ymin, (int)ceil(v_lo - e)), ay1 = min(ymax, (int)floor(v_hi + e));
    const int bx0 = max(0, (int)ceil(-u_hi - e)), bx1 = min(P.pxw - 1, (int)floor(-u_lo + e));
    const int by0 = max(ymin, (int)ceil(-v_hi - e)), by1 = min(ymax, (int)floor(-v_lo + e));
    const int nax = max(0, ax1 - ax0 + 1), nbx = max(0, bx1 - bx0 + 1);
    const int na = nax * max(0, ay1 - ay0 + 1), nb = nbx * max(0, by1 - by0 + 1);
    const float4* val = P.val + (size_t)item * P.pyh * P.pxw;
    for (int k = lane; k < na + nb; k += 64) {
        int x, y;
        bool flip;
        if (k < na) { y = ay0 + k / nax; x = ax0 + k % nax; flip = false; }
        else { const int kk = k - na; y = by0 + kk / nbx; x = bx0 + kk % nbx; flip = true; }
        double xp, yp, zp = 0.;
        {
#pragma clang fp contract(off)
            xp = a0 * x + a1 * y;
            yp = a2 * x + a3 * y;
            if constexpr (D3) zp = a4 * x + a5 * y;
        }
        const bool neg = xp < 0;
        if (neg != flip) continue;   // inserted in the other walk (or by another brick)
        if (neg) { xp = -xp; yp = -yp; zp = -zp; }
        const double fx0 = floor(xp), fy0 = floor(yp);
        const int x0 = (int)fx0, y0 = (int)fy0 + bh;
        if (x0 < 0 || x0 + 1 >= bxdim || y0 < 0 || y0 + 1 >= bpad) continue;
        if (x0 + 1 < blo[0] || x0 >= blo[0] + S || y0 + 1 < blo[1] || y0 >= blo[1] + S) continue;
        double fz0 = 0.;
        int z0 = 0;
        if constexpr (D3) {
            fz0 = floor(zp);
            z0 = (int)fz0 + bh;
            if (z0 < 0 || z0 + 1 >= bpad || z0 + 1 < blo[2] || z0 >= blo[2] + S) continue;
        }
        // the pixel's values (zero for a frequency outside the pixel set or not inserted)
        const float4 v = val[(y + P.py0) * P.pxw + x];
        if (!(v.z > 0.f)) continue;
        const float vr = v.x, vi = neg ? -v.y : v.y, wacc = v.z;
        const double fx = xp - fx0, fy = yp - fy0;
        const double mfx = 1. - fx, mfy = 1. - fy;
        if constexpr (D3) {
            const double fz = zp - fz0, mfz = 1. - fz;
            const double wts[8] = {mfz * mfy * mfx, mfz * mfy * fx, mfz * fy * mfx, mfz * fy * fx,
                                   fz * mfy * mfx, fz * mfy * fx, fz * fy * mfx, fz * fy * fx};
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int lx = x0 + (q & 1) - blo[0], ly = y0 + ((q >> 1) & 1) - blo[1], lz = z0 + (q >> 2) - blo[2];
                if ((unsigned)lx >= S || (unsigned)ly >= S || (unsigned)lz >= S) continue;
                const int li = (lz * S + ly) * S + lx;
                const float wk = (float)wts[q];
                atomicAdd(&s_re[li], wk * vr);
                atomicAdd(&s_im[li], wk * vi);
                atomicAdd(&s_w[li], wk * wacc);
            }
        } else {
            const double wts[4] = {mfy * mfx, mfy * fx, fy * mfx, fy * fx};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int lx = x0 + (q & 1) - blo[0], ly = y0 + (q >> 1) - blo[1];
                if ((unsigned)lx >= S || (unsigned)ly >= S) continue;
                const int li = ly * S + lx;
                const float wk = (float)wts[q];
                atomicAdd(&s_re[li], wk * vr);
                atomicAdd(&s_im[li], wk * vi);
                atomicAdd(&s_w[li], wk * wacc);
            }
        }
    }
}

// The batch's brick items grouped by class (a counting sort in one workgroup: counts, their
// scan, then each item placed at its class's next free slot), so that a unit reads the items of
// its own class only. Items with nothing to insert (class -1) are left out.
// Then the units: each brick's share count G from its class's count, their exclusive scan
// boff[0..nbricks] (boff[nbricks]: the units of the batch), and boff[nbricks + 1], the work-queue
// counter of k_bp_brick, cleared.
#define BP_KMAX 8192
#define BP_GMAX 256
__global__ __launch_bounds__(1024) void k_bp_bucket(const long long* __restrict__ item_off, int B, long long item_cap,
                                                    const float4* __restrict__ filt, int K, int* __restrict__ cls_off,
                                                    int* __restrict__ cls_items, const BpBrick* __restrict__ bricks,
                                                    int nbricks, float unit_items, int* __restrict__ boff)
{
    __shared__ int cnt[BP_KMAX];
    __shared__ int part[1024];
    const long long total = min(item_off[B], item_cap);
    for (int k = threadIdx.x; k < K; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    for (long long i = threadIdx.x; i < total; i += blockDim.x) {
        const int c = __float_as_int(filt[i].w);
        if (c >= 0) atomicAdd(&cnt[c], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = 0; k < K; k++) {
            const int v = cnt[k];
            cnt[k] = run;
            cls_off[k] = run;
            run += v;
        }
        cls_off[K] = run;
    }
    __syncthreads();
    for (long long i = threadIdx.x; i < total; i += blockDim.x) {
        const int c = __float_as_int(filt[i].w);
        if (c >= 0) cls_items[atomicAdd(&cnt[c], 1)] = (int)i;
    }
    __syncthreads();   // cls_off is complete
    // units: thread t takes bricks [t * per, (t + 1) * per)
    const int per = (nbricks + blockDim.x - 1) / blockDim.x;
    const int j0 = min(nbricks, (int)threadIdx.x * per), j1 = min(nbricks, j0 + per);
    int sum = 0;
    for (int j = j0; j < j1; j++) {
        const BpBrick br = bricks[j];
        const int n = cls_off[br.c + 1] - cls_off[br.c];
        const int G = n == 0 ? 0 : max(1, min(BP_GMAX, (int)ceilf(br.reach * (float)n / unit_items)));
        boff[j] = G;   // counts, scanned below by the same thread
        sum += G;
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < (int)blockDim.x; t++) {
            const int v = part[t];
            part[t] = run;
            run += v;
        }
        boff[nbricks] = run;
        boff[nbricks + 1] = 0;
    }
    __syncthreads();
    int run = part[threadIdx.x];
    for (int j = j0; j < j1; j++) {
        const int G = boff[j];
        boff[j] = run;
        run += G;
    }
}

// One unit: share g of G of the items of brick br's class, inserted into an LDS copy of the
// brick, which is then added to the volume.
template <int D3>
__device__ void bp_brick_unit(const BpBrickParams& P, const BpBrick& br, int g, int G)
{
    constexpr int S = BpDims<D3>::S, VOX = BpDims<D3>::VOX;
    __shared__ float s_re[VOX], s_im[VOX], s_w[VOX];
    __shared__ int s_list[BP_LIST];
    __shared__ int s_n;
    const int bpad = P.bp_pads[br.c], bxdim = bpad / 2 + 1, bh = bpad / 2;
    const int blo[3] = {br.bx * S, br.by * S, br.bz * S};
    for (int v = threadIdx.x; v < VOX; v += blockDim.x) { s_re[v] = 0.f; s_im[v] = 0.f; s_w[v] = 0.f; }
    if (threadIdx.x == 0) s_n = 0;
    // positions (centred) whose footprint [floor, floor + 1] meets the brick: [lo, hi) per axis
    const double lo[3] = {(double)(blo[0] - 1), (double)(blo[1] - bh - 1), (double)(blo[2] - bh - 1)};
    const double hi[3] = {(double)(blo[0] + S), (double)(blo[1] - bh + S), (double)(blo[2] - bh + S)};
    const float flo0 = (float)lo[0], flo1 = (float)lo[1], flo2 = (float)lo[2];
    const float fhi0 = (float)hi[0], fhi1 = (float)hi[1], fhi2 = (float)hi[2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    __syncthreads();
    // the items of the unit's class, its share of them
    const int beg = P.cls_off[br.c], end = P.cls_off[br.c + 1];
    for (int base = beg; base < end; base += blockDim.x) {
        const int idx = base + threadIdx.x;
        if (idx < end && (idx - beg) % G == g) {
            const int i = P.cls_items[idx];
            const float4 f = P.filt[i];
            {
                bool take = true;
                if constexpr (D3) {
                    // the plane n.q = 0 meets the box when n.q takes both signs over its corners
                    const float dmin = fminf(f.x * flo0, f.x * fhi0) + fminf(f.y * flo1, f.y * fhi1) +
                                       fminf(f.z * flo2, f.z * fhi2);
                    const float dmax = fmaxf(f.x * flo0, f.x * fhi0) + fmaxf(f.y * flo1, f.y * fhi1) +
                                       fmaxf(f.z * flo2, f.z * fhi2);
                    take = dmin <= 0.01f && dmax >= -0.01f;
                }
                if (take) s_list[atomicAdd(&s_n, 1)] = i;
            }
        }
        __syncthreads();
        const int n = s_n;
        __syncthreads();   // every thread has read the count before the next appends
        if (n > BP_LIST - BP_THREADS || base + (int)blockDim.x >= end) {
            for (int e = wave; e < n; e += nwaves)
                bp_insert_item<D3>(P, s_list[e], bxdim, bpad, bh, blo, lo, hi, s_re, s_im, s_w, lane);
            __syncthreads();
            if (threadIdx.x == 0) s_n = 0;
            __syncthreads();
        }
    }
    float2* bpd = P.bp_data[br.c];
    float* bpw = P.bp_weight[br.c];
    for (int v = threadIdx.x; v < VOX; v += blockDim.x) {
        const float r = s_re[v], im = s_im[v], w = s_w[v];
        if (r == 0.f && im == 0.f && w == 0.f) continue;
        const int lx = v % S, ly = (v / S) % S, lz = D3 ? v / (S * S) : 0;
        const size_t gi = ((size_t)(blo[2] + lz) * bpad + (size_t)(blo[1] + ly)) * bxdim + (size_t)(blo[0] + lx);
        if (G == 1) {
            const float2 o = bpd[gi];
            bpd[gi] = make_float2(o.x + r, o.y + im);
            bpw[gi] += w;
        } else {
            atomicAdd(&bpd[gi].x, r);
            atomicAdd(&bpd[gi].y, im);
            atomicAdd(&bpw[gi], w);
        }
    }
}

// One persistent launch: each workgroup takes the next unit from the queue (boff[nbricks + 1],
// cleared by k_bp_bucket) until the batch's units (boff[nbricks]) run out.
template <int D3>
__global__ __launch_bounds__(BP_THREADS) void k_bp_brick(BpBrickParams P, const BpBrick* __restrict__ bricks,
                                                         int nbricks, int* __restrict__ boff)
{
    __shared__ int s_unit;
    const int nunits = boff[nbricks];
    for (;;) {
        __syncthreads();   // every thread is done with the previous unit's LDS
        if (threadIdx.x == 0) s_unit = atomicAdd(&boff[nbricks + 1], 1);
        __syncthreads();
        const int unit = s_unit;
        if (unit >= nunits) break;
        int lo = 0, hi = nbricks;   // boff[lo] <= unit < boff[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (boff[mid] <= unit) lo = mid;
            else hi = mid;
        }
        const BpBrick br = bricks[lo];
        bp_brick_unit<D3>(P, br, unit - boff[lo], boff[lo + 1] - boff[lo]);
    }
}

__global__ void k_fill_u32(unsigned int* p, unsigned int v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f32(float* p, float v, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Zeroes n 32-bit words. The batch's launches clear buffers with this kernel instead of
// hipMemsetAsync, so that a captured batch graph holds kernel nodes only.
__global__ void k_zero_words(unsigned* __restrict__ p, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

static int zero_async(void* p, size_t bytes, hipStream_t st)
{
    if (bytes % 4) { set_error("zero_async: size not a multiple of 4 bytes"); return T2X_EINVAL; }
    const long long n = (long long)(bytes / 4);
    if (n == 0) return T2X_OK;
    const unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    k_zero_words<<<grid, 256, 0, st>>>(reinterpret_cast<unsigned*>(p), n);
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

// The batch's readback: up to kPackMax device vectors gathered into one device buffer, each at
// an 8-byte aligned offset, so that the host reads them with a single copy into page-locked
// memory (separate copies into pageable vectors each go through a driver blit and a sync).
constexpr int kPackMax = 12;
struct PackSeg {
    const unsigned* src;
    long long off_words;   // offset in dst
    long long n_words;
};
struct PackList {
    PackSeg seg[kPackMax];
};

// grid.y = segment, grid.x strides over the segment's 32-bit words
__global__ void k_pack(PackList L, unsigned* __restrict__ dst)
{
    const PackSeg s = L.seg[blockIdx.y];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < s.n_words;
         i += (long long)gridDim.x * blockDim.x)
        dst[s.off_words + i] = s.src[i];
}

struct Readback {
    struct Item {
        void* dst;
        const void* src;
        size_t bytes, off;
    };
    std::vector<Item> items;
    size_t total = 0;
    void add(void* dst, const void* src, size_t bytes)
    {
        items.push_back({dst, src, bytes, total});
        total += (bytes + 7) & ~size_t(7);
    }
    // packs, copies and waits for the stream; then fills the host destinations
    int run(Device* d, hipStream_t st)
    {
        if (items.size() > (size_t)kPackMax) { set_error("readback: too many vectors"); return T2X_EINVAL; }
        if (items.empty()) return T2X_OK;
        T2X_TRY(d->rb_dev.ensure(total));
        T2X_TRY(d->rb_host.ensure(total));
        PackList L{};
        long long maxw = 0;
        for (size_t i = 0; i < items.size(); i++) {
            if (items[i].bytes % 4) { set_error("readback: size not a multiple of 4 bytes"); return T2X_EINVAL; }
            L.seg[i] = {reinterpret_cast<const unsigned*>(items[i].src), (long long)(items[i].off / 4),
                        (long long)(items[i].bytes / 4)};
            maxw = std::max(maxw, L.seg[i].n_words);
        }
        const dim3 grid((unsigned)std::max<long long>(1, std::min<long long>((maxw + 255) / 256, 256)),
                        (unsigned)items.size());
        k_pack<<<grid, 256, 0, st>>>(L, d->rb_dev.as<unsigned>());
        T2X_HIP(hipGetLastError());
        T2X_HIP(hipMemcpyAsync(d->rb_host.ptr, d->rb_dev.ptr, total, hipMemcpyDeviceToHost, st));
        T2X_HIP(hipStreamSynchronize(st));
        for (const Item& it : items)
            std::memcpy(it.dst, d->rb_host.as<char>() + it.off, it.bytes);
        return T2X_OK;
    }
};

// One pass (adaptive_oversampling 0): the significant pairs' diff2 straight from the coarse
// matrix (rows (particle, translation), columns orientations), and the coarse pass's weight
// normalisation as the stored one.
__global__ void k_one_pass_d2(const long long* __restrict__ pair_off, const int2* __restrict__ pairs, int B,
                              const float* __restrict__ diff2, long long ld, int T, float* __restrict__ d2_out)
{
    const int b = blockIdx.x;
    if (b >= B) return;
    const long long p0 = pair_off[b], p1 = pair_off[b + 1];
    for (long long k = p0 + threadIdx.x; k < p1; k += blockDim.x) {
        const int2 nt = pairs[k];
        d2_out[k] = diff2[((long long)b * T + nt.y) * ld + nt.x];
    }
}

__global__ void k_copy_selresult(const SelResult* __restrict__ src, int B, SelResult* __restrict__ dst)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) dst[b] = src[b];
}

// C_coarse of each particle's scalars into a dense array (the GEMM epilogue's C_b)
__global__ void k_gather_ccoarse(const PartScalars* __restrict__ ps, int B, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < B) out[i] = ps[i].C_coarse;
}

// Pruned coarse pass, the bounds pass's N1 term. The coarse diff2 over the pixels beyond the
// low-frequency prefix (the high set h) is 1/2 sum_h minv |ctf F - X|^2 = Ch + N1h - G_h, with
// Ch = 1/2 sum_h minv |X|^2 (C_coarse - C_low, the constant highres term cancelling), N1h =
// 1/2 sum_h minv ctf^2 |F|^2 and G_h = sum_h minv Re(ctf F conj X); Cauchy-Schwarz bounds
// G_h <= 2 sqrt(N1h Ch). So C_b + N1 - G_low - 2 sqrt(N1h Ch) is a lower bound of the full diff2,
// and the bounds GEMM over the prefix gets it with N1 - 2 sqrt(N1h Ch) as its N1. The factor
// 1 + 2^-8 covers the split operands' rounding (the FP16 reference side's 2^-11 relative) and the
// FP32 sums. A bounds GEMM of one FP16 product per element (A hi . B hi) drops A lo . B hi (and,
// under FP16X3, A hi . B lo): each |A lo| <= 2^-11 |A| (the rows are scaled to a maximum in
// [0.5, 1), so normal FP16 but for negligible elements), so the dropped sum is at most
// 2 . 2^-11 . sum |A||B| <= 2^-10 . 2 sqrt(N1 Cl), Cl = C_low - highres/2 the image energy of the
// prefix (Cauchy-Schwarz again); eps1 = 2^-9 allows twice that, for the FP32 sums of both
// passes. eps1 = 0 for a bounds GEMM in the full split precision. The GEMM's epilogue forms
// N1 - c0 sqrt(N1h) - c1 sqrt(N1) where it reads N1 (N1h from its GEMM over the high set); this
// kernel gives each particle's c0 = 2 (1 + 2^-8) sqrt(Ch) and c1 = 2 eps1 sqrt(Cl).
__global__ void k_prune_c(const PartScalars* __restrict__ ps, int B, double eps1, double* __restrict__ c)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const double ch = fmax(ps[b].C_coarse - ps[b].C_low, 0.);
    const double cl = fmax(ps[b].C_low - ps[b].highres_Xi2 / 2., 0.);
    c[2 * b] = 2. * (1. + 1. / 256.) * sqrt(ch);
    c[2 * b + 1] = 2. * eps1 * sqrt(cl);
}

// Pruned coarse pass: per particle an upper bound of its coarse minimum diff2, the full diff2 of
// the 64 columns of the (row, block) whose low-frequency bound is smallest (where the minimum
// most likely lies), recomputed from the split operands in FP32 (the GEMM's products and scales,
// summed in another order), plus a margin of 1 + 1e-4 |d| for that order; +inf where the
// particle has no scored column. It seeds the pruning test and the lazy stores' running minima.
template <class H>
__global__ __launch_bounds__(256) void k_prune_seed(const float* __restrict__ lb, int lb_ld, int nblk, int T,
                                                    const H* __restrict__ Ah, const H* __restrict__ Al,
                                                    const float* __restrict__ As, const H* __restrict__ Bh,
                                                    const H* __restrict__ Bl, const float* __restrict__ Bs,
                                                    long long ld, int K, const double* __restrict__ Cb,
                                                    const float* __restrict__ N1, long long ldn,
                                                    const double* __restrict__ prior_o, long long N,
                                                    float* __restrict__ seed, unsigned int* __restrict__ pmin)
{
    __shared__ float sv[256];
    __shared__ long long si[256];
    __shared__ float sd[64];
    const int b = blockIdx.x, tid = threadIdx.x;
    float best = 3.0e38f;
    long long bi = -1;
    // the particle's T rows of bounds, four blocks per load (lb_ld is a multiple of 4; blocks at
    // or beyond nblk are padding)
    const int w4 = lb_ld / 4;
    for (long long q4 = tid; q4 < (long long)T * w4; q4 += blockDim.x) {
        const long long t = q4 / w4;
        const int k = (int)(q4 - t * w4) * 4;
        const float4 v = *reinterpret_cast<const float4*>(lb + ((long long)b * T + t) * lb_ld + k);
        const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; e++)
            if (k + e < nblk && vv[e] < best) { best = vv[e]; bi = t * nblk + k + e; }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w && sv[tid + w] < sv[tid]) { sv[tid] = sv[tid + w]; si[tid] = si[tid + w]; }
        __syncthreads();
    }
    const long long q = si[0];
    if (q < 0) {
        if (tid == 0) { seed[b] = __builtin_inff(); pmin[b] = float_to_ordered(__builtin_inff()); }
        return;
    }
    const long long r = (long long)b * T + q / nblk;
    const int c = tid >> 2, part = tid & 3;
    const long long n = (q % nblk) * 64 + c;
    float acc = 0.f;
    if (n < N) {
        const H* ah = Ah + r * ld;
        const H* al = Al + r * ld;
        const H* bh = Bh + n * ld;
        const H* bl = Bl ? Bl + n * ld : nullptr;
        for (int k = part; k < K; k += 4) {
            const float av = (float)ah[k] + (float)al[k];
            const float bv = bl ? (float)bh[k] + (float)bl[k] : (float)bh[k];
            acc += av * bv;
        }
    }
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (part == 0) {
        float d = 3.0e38f;
        if (n < N && prior_o[n] > 0.) {
            const float v = acc * (As ? As[r] : 1.f) * (Bs ? Bs[n] : 1.f);
            d = (float)(Cb[b] + (double)N1[(long long)b * ldn + n] - (double)v);
        }
        sd[c] = d;
    }
    __syncthreads();
    if (tid == 0) {
        float m = 3.0e38f;
        for (int k = 0; k < 64; k++) m = fminf(m, sd[k]);
        const float sdv = m < 3.0e38f ? m + 1.f + 1e-4f * fabsf(m) : __builtin_inff();
        seed[b] = sdv;
        pmin[b] = float_to_ordered(sdv);
    }
}

// ------------------------------------------------------------------------------------------
// Local angular searches: HealpixSampling::selectOrientationsWithNonZeroPriorProbability
// (healpix_sampling.cpp:769-1145) for C1, without bimodal psi or multi-body terms, sigma_cutoff 3,
// and the mean prior of convertAllSquaredDifferencesToWeights (ml_optimiser.cpp:7600-7622).
// One workgroup per particle, FP64 with the reference's operation order and no contraction.
// ------------------------------------------------------------------------------------------
struct LocalSelParams {
    const double* dir_rot;   // [Nd] degrees
    const double* dir_tilt;  // [Nd]
    const double* psi;       // [Npsi]
    int Nd, Npsi, K;
    double sig_dir;          // max(sigma_rot, sigma_tilt); 0: every direction
    double sig_psi;          // 0: every psi angle
};

__device__ __noinline__ double gaussian1d(double x, double sigma)
{
#pragma clang fp contract(off)
    return 1 / sqrt(2 * kPi * sigma * sigma) * exp(-0.5 * ((x / sigma) * (x / sigma)));   // funcs.cpp:356-360
}

__device__ __noinline__ void angles_to_direction(double rot, double tilt, double* v)
{
#pragma clang fp contract(off)
    const double a = rot * kPi / 180, b = tilt * kPi / 180;   // euler.cpp:94-114
    const double ca = cos(a), cb = cos(b), sa = sin(a), sb = sin(b);
    v[0] = sb * ca;
    v[1] = sb * sa;
    v[2] = cb;
}

// angle (degrees) between direction d and the prior direction pv (ACOSD of the dot product)
__device__ __noinline__ double direction_angle(const LocalSelParams& q, int d, const double* pv)
{
#pragma clang fp contract(off)
    double v[3];
    angles_to_direction(q.dir_rot[d], q.dir_tilt[d], v);
    double dot = 0.;
    dot += v[0] * pv[0];
    dot += v[1] * pv[1];
    dot += v[2] * pv[2];
    double a = acos(dot) * 180. / kPi;
    if (a > 180.) a = fabs(a - 360.);
    return a;
}

__device__ __noinline__ double psi_distance(double psi, double prior)
{
#pragma clang fp contract(off)
    double a = fabs(psi - prior);
    if (a > 180.) a = fabs(a - 360.);
    return a;
}

// Lowest-index minimum over the workgroup (the reference keeps the first strict minimum).
__device__ void block_argmin(double& v, int& i, double* sv, int* si)
{
    sv[threadIdx.x] = v;
    si[threadIdx.x] = i;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double ov = sv[threadIdx.x + s];
            const int oi = si[threadIdx.x + s];
            if (oi >= 0 && (si[threadIdx.x] < 0 || ov < sv[threadIdx.x] ||
                            (ov == sv[threadIdx.x] && oi < si[threadIdx.x]))) {
                sv[threadIdx.x] = ov;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    v = sv[0];
    i = si[0];
    __syncthreads();
}

// Indices i < n with v[i] > 0 written to out in increasing order by the whole workgroup (256
// threads, 4 waves; scratch: 5 ints of LDS); returns their number to every thread.
__device__ int block_compact_positive(const double* v, int n, int* out, int* scratch)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) scratch[4] = 0;
    __syncthreads();
    for (int base = 0; base < n; base += blockDim.x) {
        const int i = base + threadIdx.x;
        const bool f = i < n && v[i] > 0.;
        const unsigned long long m = __ballot(f);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) scratch[w] = __popcll(m);
        __syncthreads();
        int off = scratch[4];
        for (int k = 0; k < w; k++) off += scratch[k];
        if (f) out[off + before] = i;
        __syncthreads();
        if (threadIdx.x == 0) scratch[4] += scratch[0] + scratch[1] + scratch[2] + scratch[3];
        __syncthreads();
    }
    const int total = scratch[4];
    __syncthreads();
    return total;
}

// pang: [B][3] prior rot, tilt, psi (already resolved from the current angles where unset).
// Outputs: dir [B][Nd], psi [B][Npsi] normalised priors (0 outside the range), mean [B];
// list [B][Nd + Npsi] scratch (compacted selected indices).
__global__ __launch_bounds__(256) void k_local_prior(const double* __restrict__ pang, int B, LocalSelParams q,
                                                     double* __restrict__ dir, double* __restrict__ psi,
                                                     double* __restrict__ mean, int* __restrict__ list,
                                                     int* __restrict__ nsel)
{
#pragma clang fp contract(off)
    __shared__ double sv[256];
    __shared__ int si[256];
    __shared__ int s_cnt[2];
    const int b = blockIdx.x;
    if (b >= B) return;
    double* db = dir + (size_t)b * q.Nd;
    double* pb = psi + (size_t)b * q.Npsi;
    int* dl = list + (size_t)b * (q.Nd + q.Npsi);
    int* pl = dl + q.Nd;
    const double prot = pang[3 * b], ptilt = pang[3 * b + 1], ppsi = pang[3 * b + 2];
    // directions: Gaussian on the angle to the prior direction within 3 sigma
    double best = 9999.;
    int bi = -1;
    if (q.sig_dir > 0.) {
        double pv[3];
        angles_to_direction(prot, ptilt, pv);
        for (int d = threadIdx.x; d < q.Nd; d += blockDim.x) {
            const double a = direction_angle(q, d, pv);
            db[d] = (a < 3. * q.sig_dir) ? gaussian1d(a, q.sig_dir) : 0.;
            if (a < best) { best = a; bi = d; }
        }
    } else {
        for (int d = threadIdx.x; d < q.Nd; d += blockDim.x) db[d] = 1.;
    }
    block_argmin(best, bi, sv, si);
    // psi angles
    double bestp = 9999.;
    int bp = -1;
    for (int p = threadIdx.x; p < q.Npsi; p += blockDim.x) {
        if (q.sig_psi > 0.) {
            const double a = psi_distance(q.psi[p], ppsi);
            pb[p] = (a < 3. * q.sig_psi) ? gaussian1d(a, q.sig_psi) : 0.;
            if (a < bestp) { bestp = a; bp = p; }
        } else {
            pb[p] = 1.;
        }
    }
    block_argmin(bestp, bp, sv, si);
    __threadfence_block();
    __syncthreads();
    // the selected indices in increasing order (all threads), then the sums in that order
    // (sumprior of the reference) and the normalisation
    int nd = block_compact_positive(db, q.Nd, dl, si);
    int np = block_compact_positive(pb, q.Npsi, pl, si);
    if (threadIdx.x == 0) {
        double sd = 0.;
        for (int k = 0; k < nd; k++) sd += db[dl[k]];
        for (int k = 0; k < nd; k++) db[dl[k]] /= sd;
        if (nd == 0) { const int d0 = bi >= 0 ? bi : 0; db[d0] = 1.; dl[nd++] = d0; }
        double sp = 0.;
        for (int k = 0; k < np; k++) sp += pb[pl[k]];
        for (int k = 0; k < np; k++) pb[pl[k]] /= sp;
        if (np == 0) { const int p0 = bp >= 0 ? bp : 0; pb[p0] = 1.; pl[np++] = p0; }
        // pdf_orientation_mean over classes x local directions x local psi angles
        double m = 0.;
        for (int c = 0; c < q.K; c++)
            for (int k = 0; k < nd; k++) {
                const double x = db[dl[k]];
                for (int j = 0; j < np; j++) m += x * pb[pl[j]];
            }
        mean[b] = m / (double)((long long)q.K * nd * np);
        nsel[2 * b] = nd;
        nsel[2 * b + 1] = np;
    }
}

// The coarse sample list of a local search: for each particle, its orientations (classes with
// non-zero probability x its directions x its psi angles, in the global order) each followed by
// all T translations. orient_off: [B+1] offsets of the particles' orientations; the pairs of
// orientation j are (j * T + t).
__global__ void k_local_pairs(int B, const int* __restrict__ list, int Nd, int Npsi, const int* __restrict__ nsel,
                              const int* __restrict__ cls, int Kact, long long Norient, int T,
                              const long long* __restrict__ orient_off, int* __restrict__ orients,
                              int2* __restrict__ pairs)
{
    const int b = blockIdx.x;
    if (b >= B) return;
    const int* dl = list + (size_t)b * (Nd + Npsi);
    const int* pl = dl + Nd;
    const int nd = nsel[2 * b], np = nsel[2 * b + 1];
    const long long per_c = (long long)nd * np, no = (long long)Kact * per_c, o0 = orient_off[b];
    for (long long j = threadIdx.x; j < no; j += blockDim.x) {
        const int ci = (int)(j / per_c);
        const long long r = j - ci * per_c;
        const int di = (int)(r / np), pi = (int)(r - (long long)di * np);
        const int n = (int)(cls[ci] * Norient + (long long)dl[di] * Npsi + pl[pi]);
        const long long p0 = (o0 + j) * T;
        orients[o0 + j] = n;
        for (int t = 0; t < T; t++) pairs[p0 + t] = make_int2(n, t);
    }
}

// Local searches: N1[b][n] = +inf for orientations outside particle b's range, so that the
// coarse GEMM epilogue (diff2 = C_b + N1 - cross) gives them +inf: never a minimum, never a
// candidate, and their prior is 0.
__global__ void k_local_mask_n1(float* __restrict__ N1, long long ldn, int B, long long O_all, LocalPrior lp)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)B * O_all) return;
    const int b = (int)(i / O_all);
    const long long n = i - (long long)b * O_all;
    if (!local_in_range(lp, b, n)) N1[(long long)b * ldn + n] = __builtin_inff();
}

// The batch's distinct orientations in increasing order: flags over every orientation, an
// exclusive scan (hipCUB), the compacted list and each row's slice index.
__global__ void k_mark_orients(long long nrows, const int* __restrict__ orients, int* __restrict__ flags)
{
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nrows) flags[orients[j]] = 1;
}

__global__ void k_compact_orients(long long n, const int* __restrict__ flags, const int* __restrict__ pos,
                                  int* __restrict__ ulist)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) ulist[pos[i]] = (int)i;
}

__global__ void k_row_slices(long long nrows, const int* __restrict__ orients, const int* __restrict__ pos,
                             int* __restrict__ rowu)
{
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nrows) rowu[j] = pos[orients[j]];
}

// ------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------

// Compact pixel table of one pass: every window pixel with ires >= 0 (Mresol,
// ml_optimiser.cpp:5596-5625), ordered by window index.
// Compact pixel table of one pass with each optics group's Minvsigma2 (ml_optimiser.cpp:6703-6720):
// h_minv holds n_og rows of npix values. The pixel set (every pixel of the Mresol map for the
// store; those with ires > 0 for diff2) does not depend on the noise spectra.
static void build_table(int size, int cur, const double* sigma2, int nsig2, int n_og, double fudge, bool store,
                        PassTables& t)
{
    const int xs = size / 2 + 1, xc = cur / 2 + 1;
    t.size = size;
    t.h_x.clear(); t.h_y.clear(); t.h_win.clear(); t.h_ires.clear(); t.h_minv.clear(); t.h_inproj.clear();
    const int rmo = xs - 1;
    std::vector<int> keep_ires;
    for (int i = 0; i < size; i++) {
        int ip = (i < xs) ? i : i - size;
        for (int j = 0; j < xs; j++) {
            int ires = relion_round(std::sqrt((double)(ip * ip + j * j)));
            if (!(ires < xs && !(j == 0 && ip < 0))) continue;   // Mresol == -1
            if (!store && !(ires > 0 && ires < nsig2)) continue;   // no contribution to diff2
            int ii = ip < 0 ? ip + cur : ip;
            t.h_x.push_back(j);
            t.h_y.push_back(ip);
            t.h_win.push_back(ii * xc + j);
            t.h_ires.push_back(ires);
            // projected pixels: x <= FLOOR(sqrt(r_max_out^2 - y^2)), y = i for i <= r_max_out
            int y_proj = (i <= rmo) ? i : i - size;
            int xmax = (int)std::floor(std::sqrt((double)(rmo * rmo - y_proj * y_proj)));
            bool in = (rmo * rmo - y_proj * y_proj >= 0) && j <= xmax;
            t.h_inproj.push_back(in ? 1 : 0);
            if (y_proj != ip) t.h_y.back() = y_proj;   // row size/2 is projected at +size/2
        }
    }
    t.npix = (int)t.h_x.size();
    if (!store) {
        // the diff2 pixels in order of |frequency| (then window order): every prefix is a disc, so
        // a GEMM over the first K columns gives diff2 over a low-frequency subset, a lower bound
        // of the full diff2 (each pixel adds minv |ctf F - X|^2 / 2 >= 0)
        std::vector<int> ord(t.npix);
        for (int i = 0; i < t.npix; i++) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
            return t.h_x[a] * t.h_x[a] + t.h_y[a] * t.h_y[a] < t.h_x[b] * t.h_x[b] + t.h_y[b] * t.h_y[b];
        });
        auto perm = [&](auto& v) {
            auto w = v;
            for (int i = 0; i < t.npix; i++) w[i] = v[ord[i]];
            v.swap(w);
        };
        perm(t.h_x);
        perm(t.h_y);
        perm(t.h_win);
        perm(t.h_ires);
        perm(t.h_inproj);
    }
    t.n_og = n_og;
    t.h_minv.resize((size_t)n_og * t.npix);
    for (int g = 0; g < n_og; g++) {
        const double* s2 = sigma2 + (size_t)g * nsig2;
        for (int p = 0; p < t.npix; p++) {
            const int ires = t.h_ires[p];
            double mv = 0.;
            if (ires > 0 && ires < nsig2) mv = 1. / (fudge * s2[ires]);
            if (store && ires == 0) mv = 1. / (fudge * s2[0]);   // the origin term of the store
            t.h_minv[(size_t)g * t.npix + p] = (float)mv;
        }
    }
    t.npix_pad = (int)round_up(t.npix, 32);   // K of the N1 GEMM: a multiple of the split GEMM's K step
}

static int upload_table(PassTables& t, hipStream_t st)
{
    T2X_TRY(t.d_pix.ensure(sizeof(int2) * t.npix));
    T2X_TRY(t.d_win.ensure(sizeof(int) * t.npix));
    T2X_TRY(t.d_minv.ensure(sizeof(float) * t.h_minv.size()));
    T2X_TRY(t.d_ires.ensure(sizeof(int) * t.npix));
    T2X_TRY(t.d_inproj.ensure(t.npix));
    std::vector<int2> xy(t.npix);
    for (int i = 0; i < t.npix; i++) xy[i] = make_int2(t.h_x[i], t.h_y[i]);
    T2X_HIP(hipMemcpyAsync(t.d_pix.ptr, xy.data(), sizeof(int2) * t.npix, hipMemcpyHostToDevice, st));
    T2X_HIP(hipMemcpyAsync(t.d_win.ptr, t.h_win.data(), sizeof(int) * t.npix, hipMemcpyHostToDevice, st));
    T2X_HIP(hipMemcpyAsync(t.d_minv.ptr, t.h_minv.data(), sizeof(float) * t.h_minv.size(), hipMemcpyHostToDevice, st));
    T2X_HIP(hipMemcpyAsync(t.d_ires.ptr, t.h_ires.data(), sizeof(int) * t.npix, hipMemcpyHostToDevice, st));
    T2X_HIP(hipMemcpyAsync(t.d_inproj.ptr, t.h_inproj.data(), t.npix, hipMemcpyHostToDevice, st));
    // the frequency window of the pixels: x in [0, size/2], y in [-size/2, size/2]
    t.px_w = t.size / 2 + 1;
    t.py0 = t.size / 2;
    t.py_h = t.size + 1;
    T2X_HIP(hipStreamSynchronize(st));
    return T2X_OK;
}

// exp(-2 pi i (x sx + y sy)/N) for each translation and compact pixel (double -> float2).
static std::vector<float2> phase_table(const PassTables& t, const std::vector<double>& sx,
                                       const std::vector<double>& sy, int N)
{
    std::vector<float2> ab(sx.size() * t.npix);
    for (size_t k = 0; k < sx.size(); k++) {
        double xs = sx[k] / -(double)N, ys = sy[k] / -(double)N;
        bool zero = std::fabs(xs) < 1e-6 && std::fabs(ys) < 1e-6;
        for (int p = 0; p < t.npix; p++) {
            // the phase uses the logical row of the window (i <= size/2 -> i)
            int y = t.h_y[p];
            double dotp = 2 * kPi * (t.h_x[p] * xs + y * ys);
            double a = zero ? 1.0 : std::cos(dotp), bb = zero ? 0.0 : std::sin(dotp);
            ab[k * t.npix + p] = make_float2((float)a, (float)bb);
        }
    }
    return ab;
}

template <class T>
static int upload(DevBuf& buf, const std::vector<T>& v, hipStream_t st)
{
    T2X_TRY(buf.ensure(sizeof(T) * std::max<size_t>(v.size(), 1)));
    if (!v.empty())
        T2X_HIP(hipMemcpyAsync(buf.ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
    return T2X_OK;
}

static int ensure_fft(Device* d, int N, int batch)
{
    if (d->fft_plan && d->fft_n == N && d->fft_batch == batch) return T2X_OK;
    if (d->fft_plan) { (void)hipfftDestroy(d->fft_plan); d->fft_plan = 0; }
    int dims[2] = {N, N};
    if (hipfftPlanMany(&d->fft_plan, 2, dims, nullptr, 1, N * N, nullptr, 1, N * (N / 2 + 1),
                       HIPFFT_R2C, batch) != HIPFFT_SUCCESS) {
        set_error("hipfftPlanMany failed");
        return T2X_EDEVICE;
    }
    if (hipfftSetStream(d->fft_plan, d->stream) != HIPFFT_SUCCESS) {
        set_error("hipfftSetStream failed");
        return T2X_EDEVICE;
    }
    d->fft_n = N;
    d->fft_batch = batch;
    return T2X_OK;
}

// The image side of the split GEMMs: rows (b, t) < M of minv ctf X_t as hi / lo halves with their
// row scales (rows M..rows-1 zero), and A1 = minv ctf^2 / 2 for the N1 term. Rows that fit the
// LDS row buffer go through k_coarse_imgside_split; longer ones (coarse images beyond 8,192
// pixels) through an FP32 matrix in HBM and k_split_rows, with the same bits.
static int imgside_split(Device* d, int prec, const PassTables& tc, const PartScalars* d_ps, int Pw, int T, int B,
                         long long rows, int Kc, int Kn, hipStream_t st)
{
    const long long M = (long long)B * T;
    const size_t lds = sizeof(float) * Kc;
    if (lds <= kRowLdsMax || prec == T2X_PREC_BF16X3) {
        if (prec == T2X_PREC_BF16X3)   // no row scale: the kernel needs no row buffer
            k_coarse_imgside_split<__bf16><<<(unsigned)rows, 256, 0, st>>>(
                d->Fimg_cur.as<float2>(), d->Fctf_cur.as<float>(), Pw, tc.d_win.as<int>(), tc.d_minv.as<float>(),
                tc.npix, d_ps, d->ab_coarse.as<float2>(), T, (int)M, Kc, (__bf16*)d->Ah.ptr, (__bf16*)d->Al.ptr,
                d->As.as<float>(), 0, d->A1.as<float>(), Kn);
        else
            k_coarse_imgside_split<_Float16><<<(unsigned)rows, 256, lds, st>>>(
                d->Fimg_cur.as<float2>(), d->Fctf_cur.as<float>(), Pw, tc.d_win.as<int>(), tc.d_minv.as<float>(),
                tc.npix, d_ps, d->ab_coarse.as<float2>(), T, (int)M, Kc, (_Float16*)d->Ah.ptr, (_Float16*)d->Al.ptr,
                d->As.as<float>(), 1, d->A1.as<float>(), Kn);
        T2X_HIP(hipGetLastError());
        return T2X_OK;
    }
    T2X_TRY(d->A.ensure(sizeof(float) * rows * Kc));
    T2X_TRY(zero_async(d->A.ptr, sizeof(float) * rows * Kc, st));
    if (B > 0) {
        k_coarse_imgside<<<dim3((unsigned)T, (unsigned)B), 256, 0, st>>>(
            d->Fimg_cur.as<float2>(), d->Fctf_cur.as<float>(), Pw, tc.d_win.as<int>(), tc.d_minv.as<float>(), tc.npix,
            d_ps, d->ab_coarse.as<float2>(), T, d->A.as<float>(), Kc, d->A1.as<float>(), Kn);
        T2X_HIP(hipGetLastError());
    }
    return launch_split(prec, d->A.as<float>(), Kc, Kc, (int)rows, (int)M, d->Ah.ptr, d->Al.ptr, Kc, d->As.as<float>(), st);
}

struct EventTimer {
    Device* d;
    bool on;
    int idx = 0;
    void mark() { if (on) (void)hipEventRecord(d->ev[idx++], d->stream); }
};

}  // namespace t2x

using namespace t2x;

extern "C" int t2x_estep(t2x_device* h, const t2x_estep_config* cfg, const t2x_model* model, int n,
                         const float* images, int images_on_device, const t2x_particle_in* pin,
                         t2x_particle_out* out, t2x_wsum* wsum, t2x_estep_debug* dbg)
{
    if (!h || !cfg || !model || (n > 0 && (!images || !pin || !out))) {
        set_error("null argument");
        return T2X_EINVAL;
    }
    if (n < 0) { set_error("negative particle count"); return T2X_EINVAL; }
    if (!model->sigma2_noise || !model->pdf_class || !model->pdf_direction) {
        set_error("model arrays missing");
        return T2X_EINVAL;
    }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    const unsigned long long call_id = ++d->call_counter;   // a batch graph serves one call
    d->eager_B.clear();
    T2X_HIP(hipSetDevice(d->id));
    HostSampling S;
    T2X_TRY(build_sampling(cfg, S));
    const int K = cfg->nr_classes;
    const int N = cfg->ori_size, xN = N / 2 + 1;
    // adaptive_oversampling 0: one pass over the coarse sampling at the current size, whose
    // weights (normalised over every sample) are stored directly (ml_optimiser.cpp:4373-4420,
    // nr_sampling_passes = 1)
    const bool one_pass = cfg->adaptive_oversampling == 0;
    const int cs = cfg->current_size, cc = one_pass ? cs : cfg->coarse_size, xc = cs / 2 + 1;
    const int Pw = cs * xc;
    const long Nd = S.nr_dir, Npsi = S.nr_psi, T = S.nr_trans;
    const long Norient = Nd * Npsi;
    const long O_all = K * Norient;
    const int over_rot = S.over_rot, over_trans = S.over_trans, nover = over_rot * over_trans;
    hipStream_t st = d->stream;
    if (T > MAX_COARSE_T) { set_error("more than 512 coarse translations"); return T2X_EINVAL; }
    if (!(cfg->gemm_tile == 0 || cfg->gemm_tile == 128 || cfg->gemm_tile == 256 || cfg->gemm_tile == 257 ||
          cfg->gemm_tile == 258 || cfg->gemm_tile == 259) || cfg->gemm_group < 0) {
        set_error("gemm_tile must be 0 (autotuned), 128, 256, 257, 258 or 259 and gemm_group >= 0");
        return T2X_EINVAL;
    }
    // learned proposal: the coarse pass scores only the top-k bins (classes x HEALPix directions
    // one order down = blocks of 4 directions x Npsi orientations) of each particle
    const bool use_prop = cfg->proposal_topk > 0;
    const bool local = cfg->orientational_prior_mode == T2X_PRIOR_ROTTILT_PSI;
    if (local && one_pass) {
        set_error("local searches need adaptive_oversampling >= 1 (two passes)");
        return T2X_EINVAL;
    }
    const int n_bins = (int)(K * (Nd / 4)), bin_cols = (int)(4 * Npsi);
    int topk = 0;
    if (use_prop) {
        if (local) {
            set_error("local searches and the learned proposal are exclusive");
            return T2X_EINVAL;
        }
        if (cfg->precision == T2X_PREC_FP32) {
            set_error("the learned proposal runs with a split-precision cross term (BF16X3 or FP16X3)");
            return T2X_EINVAL;
        }
        if (cfg->ref_dim != 3) { set_error("the learned proposal needs 3D references"); return T2X_EINVAL; }
        if (!d->net || !d->net->trained) { set_error("proposal network not trained on this handle"); return T2X_ESTATE; }
        if (Nd % 4 || d->net->n_bins != n_bins) {
            set_error("proposal network bins do not match classes x directions / 4");
            return T2X_EINVAL;
        }
        if (bin_cols % 64) { set_error("learned proposal needs 4 * nr_psi to be a multiple of 64"); return T2X_EINVAL; }
        topk = std::min(cfg->proposal_topk, n_bins);
        if ((size_t)(n_bins + topk) * sizeof(float) > 64 * 1024) {   // the top-k kernel holds them in LDS
            set_error("learned proposal: too many bins");
            return T2X_EINVAL;
        }
        if (cfg->proposal_gate != T2X_GATE_TOP1 && cfg->proposal_gate != T2X_GATE_TOPK_MASS) {
            set_error("learned proposal: unknown proposal_gate");
            return T2X_EINVAL;
        }
        d->prop_gate_top1.clear();
        d->prop_gate_mass.clear();
    }
    if (T * over_trans > STORE_MAX_T) { set_error("too many fine translations"); return T2X_EINVAL; }
    // image groups (scale correction) and optics groups (noise spectra)
    const int n_og = cfg->nr_optics_groups > 0 ? cfg->nr_optics_groups : 1;
    const int n_grp = cfg->nr_groups > 0 ? cfg->nr_groups : 1;
    if (cfg->nr_optics_groups < 0 || cfg->nr_groups < 0) { set_error("negative group count"); return T2X_EINVAL; }
    if (cfg->do_scale_correction) {
        if (!model->scale_correction) { set_error("scale correction without model->scale_correction"); return T2X_EINVAL; }
        for (int g = 0; g < n_grp; g++) {
            const double sc = model->scale_correction[g];
            // relion_refine stops above 10000 (ml_optimiser.cpp:8329-8333); a scale <= 0 has no meaning
            if (!(sc > 0.) || !(sc <= 10000.)) {
                set_error("scale_correction of group " + std::to_string(g) + " outside (0, 10000]");
                return T2X_EINVAL;
            }
        }
    }
    for (int i = 0; i < n; i++)
        if (pin[i].optics_group < 0 || pin[i].optics_group >= n_og || pin[i].group_id < 0 || pin[i].group_id >= n_grp) {
            set_error("particle " + std::to_string(i) + ": group_id or optics_group out of range");
            return T2X_EINVAL;
        }
    std::vector<const float2*> h_vols(K);
    std::vector<int> h_pads(K);
    std::vector<float2*> h_bpd(K);
    std::vector<float*> h_bpw(K);
    std::vector<int> h_bpads(K), h_bprmax(K);
    double pf = 0;
    for (int c = 0; c < K; c++) {
        auto it = d->classes.find(c);
        if (it == d->classes.end() || !it->second.proj || !it->second.bp_data) {
            set_error("projector or backprojector of class " + std::to_string(c) + " not set");
            return T2X_ESTATE;
        }
        ClassVolumes& cv = it->second;
        if (cv.p_ref_dim != cfg->ref_dim || cv.b_ref_dim != cfg->ref_dim) {
            set_error("projector/backprojector of class " + std::to_string(c) + " do not have ref_dim " +
                      std::to_string(cfg->ref_dim));
            return T2X_EINVAL;
        }
        h_vols[c] = cv.proj;
        h_pads[c] = cv.p_pad;
        h_bpd[c] = cv.bp_data;
        h_bpw[c] = cv.bp_weight;
        h_bpads[c] = cv.b_pad;
        h_bprmax[c] = cv.b_rmax;
        if (c == 0) pf = cv.p_pf;
        if (cv.p_pf != pf || cv.b_pf != pf) { set_error("padding factors differ between classes"); return T2X_EINVAL; }
        if (cv.p_rmax * 2 < cs) {
            // the CPU projector only fills r <= r_max; smaller r_max than current_size/2 is allowed
        }
    }
    const int ref2d = cfg->ref_dim == 2 ? 1 : 0;
    // 2D references: the offset prior is the class's (ml_optimiser.cpp:7627-7631, 8283-8287).
    // Equal for every class (the first iteration): one translation prior table per particle;
    // otherwise one per (particle, class), tK = K (k_prior_t, trans_prior).
    double class_prior_x = 0., class_prior_y = 0.;
    int tK = 0;
    if (ref2d && model->prior_offset_class) {
        class_prior_x = model->prior_offset_class[0];
        class_prior_y = model->prior_offset_class[1];
        for (int c = 0; c < 2 * K; c++)
            if (!std::isfinite(model->prior_offset_class[c])) {
                set_error("2D references: non-finite class offset prior");
                return T2X_EINVAL;
            }
        for (int c = 1; c < K; c++)
            if (model->prior_offset_class[2 * c] != class_prior_x || model->prior_offset_class[2 * c + 1] != class_prior_y)
                tK = K;
        if (tK) {
            std::vector<double> cp(model->prior_offset_class, model->prior_offset_class + 2 * K);
            T2X_TRY(upload(d->cls_prior, cp, st));
        }
    }
    // ---- per-configuration tables ----
    const double* sigma2 = model->sigma2_noise;
    build_table(cc, cs, sigma2, xN, n_og, cfg->sigma2_fudge, false, d->coarse_tab);
    build_table(cs, cs, sigma2, xN, n_og, cfg->sigma2_fudge, true, d->fine_tab);
    PassTables& tc = d->coarse_tab;
    PassTables& tf = d->fine_tab;
    T2X_TRY(upload_table(tc, st));
    T2X_TRY(upload_table(tf, st));
    // fine diff2 pixel subset: pixels with minv != 0 come first? (all store pixels are used;
    // the diff2 kernel reads minv of the diff2 definition)
    std::vector<float> minv_diff2((size_t)n_og * tf.npix);
    for (int g = 0; g < n_og; g++)
        for (int p = 0; p < tf.npix; p++)
            minv_diff2[(size_t)g * tf.npix + p] = (tf.h_ires[p] > 0) ? tf.h_minv[(size_t)g * tf.npix + p] : 0.f;
    // noise spectra (mask background) and the scale sums' shell mask
    {
        std::vector<double> s2(sigma2, sigma2 + (size_t)n_og * xN);
        T2X_TRY(upload(d->sigma2_tab, s2, st));
        if (cfg->do_scale_correction && model->data_vs_prior_class) {
            std::vector<double> dv(model->data_vs_prior_class, model->data_vs_prior_class + (size_t)K * xN);
            T2X_TRY(upload(d->dvp_tab, dv, st));
        }
    }
    const bool scale_sums = cfg->do_scale_correction && model->data_vs_prior_class;
    DevBuf& d_minv_fd = d->minv_fd;
    T2X_TRY(upload(d_minv_fd, minv_diff2, st));
    // phase tables (kept from the previous call when its sizes and translations were the same)
    {
        std::vector<double> key = {(double)N, (double)cc, (double)cs, (double)tc.npix, (double)tf.npix,
                                   (double)S.coarse_tx_px.size(), (double)S.fine_tx_px.size()};
        key.insert(key.end(), S.coarse_tx_px.begin(), S.coarse_tx_px.end());
        key.insert(key.end(), S.coarse_ty_px.begin(), S.coarse_ty_px.end());
        key.insert(key.end(), S.fine_tx_px.begin(), S.fine_tx_px.end());
        key.insert(key.end(), S.fine_ty_px.begin(), S.fine_ty_px.end());
        if (key != d->ab_key || !d->ab_coarse.ptr || !d->ab_fine.ptr) {
            d->ab_key.clear();
            std::vector<float2> abc = phase_table(tc, S.coarse_tx_px, S.coarse_ty_px, N);
            std::vector<float2> abf = phase_table(tf, S.fine_tx_px, S.fine_ty_px, N);
            T2X_TRY(upload(d->ab_coarse, abc, st));
            T2X_TRY(upload(d->ab_fine, abf, st));
            d->ab_key = key;
        }
    }
    // the fine pass's GEMM operand: phase rows with K padded to whole pixel chunks
    const int ldk_fine = (int)round_up(tf.npix, FT_KC) * 2;
    {
        const long long Tf = T * over_trans;
        if (cfg->precision == T2X_PREC_FP32) {
            T2X_TRY(d->ab_fk.ensure(sizeof(float) * Tf * ldk_fine));
            k_fine_ab_rows<_Float16><<<(unsigned)Tf, 256, 0, st>>>(d->ab_fine.as<float2>(), tf.npix, ldk_fine,
                                                                  d->ab_fk.as<float>(), nullptr, nullptr);
        } else {
            T2X_TRY(d->ab_fh.ensure(sizeof(uint16_t) * Tf * ldk_fine));
            T2X_TRY(d->ab_fl.ensure(sizeof(uint16_t) * Tf * ldk_fine));
            if (cfg->precision == T2X_PREC_BF16X3)
                k_fine_ab_rows<__bf16><<<(unsigned)Tf, 256, 0, st>>>(d->ab_fine.as<float2>(), tf.npix, ldk_fine, nullptr,
                                                                    (__bf16*)d->ab_fh.ptr, (__bf16*)d->ab_fl.ptr);
            else
                k_fine_ab_rows<_Float16><<<(unsigned)Tf, 256, 0, st>>>(d->ab_fine.as<float2>(), tf.npix, ldk_fine, nullptr,
                                                                      (_Float16*)d->ab_fh.ptr, (_Float16*)d->ab_fl.ptr);
        }
        T2X_HIP(hipGetLastError());
    }
    if (T > FT_MAX_T) { set_error("more than 512 coarse translations"); return T2X_EINVAL; }
    if ((long long)T * over_trans * ldk_fine >= 0x7fffffffLL) {
        set_error("fine pass: phase table beyond 2^31 elements");
        return T2X_EINVAL;
    }
    // local searches in FP32: the phase table as the coarse GEMM's shared operand (Tl rows, Kc columns)
    const bool local_rows = cfg->orientational_prior_mode == T2X_PRIOR_ROTTILT_PSI && !getenv("T2X_LOCAL_DENSE");
    const int Tl = (int)round_up(T, 128);
    if (local_rows && cfg->precision == T2X_PREC_FP32) {   // the split modes take the image side instead
        const int Kcl = tc.npix_pad * 2;
        T2X_TRY(d->lc_ab.ensure(sizeof(float) * Tl * Kcl));
        k_ab_rows<<<Tl, 256, 0, st>>>(d->ab_coarse.as<float2>(), (int)T, tc.npix, Kcl, d->lc_ab.as<float>());
        T2X_HIP(hipGetLastError());
    }
    // orientation prior (NOPRIOR): pdf_direction / mean over all classes and orientations;
    // 0 where the class or the direction has zero probability (never scored). Local searches:
    // the class gate (1, or 0 for a class of zero probability); the per-particle direction and
    // psi priors come from k_local_prior (ml_optimiser.cpp:7031-7045, 7700-7716)
    double mean_o = 0.;
    for (int c = 0; c < K; c++)
        for (long dd = 0; dd < Nd; dd++) mean_o += model->pdf_direction[(size_t)c * Nd + dd] * Npsi;
    mean_o /= (double)(K * Norient);
    std::vector<double> prior_o(O_all);
    for (int c = 0; c < K; c++)
        for (long o = 0; o < Norient; o++) {
            double p = local ? 1. : model->pdf_direction[(size_t)c * Nd + o / Npsi];
            bool scored = model->pdf_class[c] > 0. && p > 0.;
            if (mean_o != 0. && !local) p /= mean_o;
            prior_o[(size_t)c * Norient + o] = scored ? p : 0.;
        }
    T2X_TRY(upload(d->prior_o, prior_o, st));
    T2X_TRY(d->log_prior_o.ensure(sizeof(double) * std::max<long>(1, O_all)));
    k_log_prior<<<(unsigned)std::min<long>((O_all + 255) / 256, 65536), 256, 0, st>>>(d->prior_o.as<double>(), O_all,
                                                                                  d->log_prior_o.as<double>());
    T2X_HIP(hipGetLastError());
    // device copies of angle tables and class pointer arrays
    DevBuf &d_dirsc = d->dirsc, &d_psisc = d->psisc, &d_ftx = d->ftx, &d_fty = d->fty, &d_vols = d->vols_tab,
           &d_pads = d->pads_tab, &d_bpd = d->bpd_tab, &d_bpw = d->bpw_tab, &d_bpads = d->bpads_tab,
           &d_bprmax = d->bprmax_tab;
    T2X_TRY(upload(d_dirsc, S.sub_dir_sc, st));
    T2X_TRY(upload(d_psisc, S.sub_psi_sc, st));
    T2X_TRY(upload(d_ftx, S.fine_tx_px, st));
    T2X_TRY(upload(d_fty, S.fine_ty_px, st));
    T2X_TRY(upload(d_vols, h_vols, st));
    T2X_TRY(upload(d_pads, h_pads, st));
    T2X_TRY(upload(d_bpd, h_bpd, st));
    T2X_TRY(upload(d_bpw, h_bpw, st));
    T2X_TRY(upload(d_bpads, h_bpads, st));
    T2X_TRY(upload(d_bprmax, h_bprmax, st));
    DevBuf &d_ctx = d->ctx, &d_cty = d->cty;   // coarse translations (A) for k_prior_t
    T2X_TRY(upload(d_ctx, S.tx, st));
    T2X_TRY(upload(d_cty, S.ty, st));
    LocalSelParams lsp{};
    // local searches score each particle's own sample list (T2X_LOCAL_DENSE=1: the dense pass over
    // every orientation with out-of-range ones masked, kept for comparison)
    const bool local_pairs = local && !getenv("T2X_LOCAL_DENSE");
    int Kact = 0;
    if (local_pairs) {
        std::vector<int> cls;
        for (int c = 0; c < K; c++) if (model->pdf_class[c] > 0.) cls.push_back(c);
        Kact = (int)cls.size();
        T2X_TRY(upload(d->lc_cls, cls, st));
        std::vector<double> dsc(4 * Nd), psc(2 * Npsi);
        for (long dd = 0; dd < Nd; dd++) {
            const double a = S.rot[dd] * kPi / 180, b = S.tilt[dd] * kPi / 180;
            dsc[4 * dd] = std::sin(a); dsc[4 * dd + 1] = std::cos(a);
            dsc[4 * dd + 2] = std::sin(b); dsc[4 * dd + 3] = std::cos(b);
        }
        for (long p = 0; p < Npsi; p++) {
            const double g = S.psi[p] * kPi / 180;
            psc[2 * p] = std::sin(g); psc[2 * p + 1] = std::cos(g);
        }
        T2X_TRY(upload(d->lc_dirsc, dsc, st));
        T2X_TRY(upload(d->lc_psisc, psc, st));
    }
    if (local) {
        std::vector<double> tab(2 * Nd + Npsi);
        for (long dd = 0; dd < Nd; dd++) { tab[dd] = S.rot[dd]; tab[Nd + dd] = S.tilt[dd]; }
        for (long p = 0; p < Npsi; p++) tab[2 * Nd + p] = S.psi[p];
        T2X_TRY(upload(d->loc_tab, tab, st));
        lsp.dir_rot = d->loc_tab.as<double>();
        lsp.dir_tilt = lsp.dir_rot + Nd;
        lsp.psi = lsp.dir_rot + 2 * Nd;
        lsp.Nd = (int)Nd;
        lsp.Npsi = (int)Npsi;
        lsp.K = K;
        lsp.sig_dir = ref2d ? 0. : std::max(cfg->sigma_rot, cfg->sigma_tilt);
        lsp.sig_psi = cfg->sigma_psi;
    }

    // ---- coarse projections, cached across calls ----
    const int Kc = tc.npix_pad * 2;            // GEMM K (re/im)
    const int Kn = tc.npix_pad;                // N1 GEMM K
    const long long O_pad = round_up(O_all, 256);   // multiple of both GEMM tile widths
    std::string key = std::to_string(cc) + "/" + std::to_string(cs) + "/" + std::to_string(K) + "/" +
                      std::to_string(cfg->healpix_order) + "/" + std::to_string(cfg->psi_step) + "/" +
                      std::to_string(tc.npix) + "/p" + std::to_string(cfg->precision);
    const int prec = cfg->precision;
    for (int c = 0; c < K; c++) key += "/" + std::to_string(d->classes[c].proj_version);
    EventTimer et{d, true};
    (void)hipEventRecord(d->ev[0], st);
    if (!local_pairs && key != d->coarse_key) {
        T2X_TRY(d->coarse_B.ensure(sizeof(float) * O_pad * Kc));
        T2X_TRY(d->coarse_B2.ensure(sizeof(float) * O_pad * Kn));
        T2X_HIP(hipMemsetAsync(d->coarse_B.ptr, 0, sizeof(float) * O_pad * Kc, st));
        T2X_HIP(hipMemsetAsync(d->coarse_B2.ptr, 0, sizeof(float) * O_pad * Kn, st));
        // the projection frames pf * A^-1 of the orientations (kept while the angles are the same)
        std::vector<double> akey = {pf, (double)Nd, (double)Npsi};
        akey.insert(akey.end(), S.rot.begin(), S.rot.begin() + Nd);
        akey.insert(akey.end(), S.tilt.begin(), S.tilt.begin() + Nd);
        akey.insert(akey.end(), S.psi.begin(), S.psi.begin() + Npsi);
        DevBuf& d_Ai = d->coarse_Ai;
        if (akey != d->coarse_Ai_key || !d_Ai.ptr) {
            d->coarse_Ai_key.clear();
            std::vector<float> Ai(Norient * 9);
            for (long dd = 0; dd < Nd; dd++)
                for (long p = 0; p < Npsi; p++) {
                    double A[9], R[9];
                    euler_matrix_host(S.rot[dd], S.tilt[dd], S.psi[p], A);
                    inv3(A, R);
                    float* a = &Ai[(dd * Npsi + p) * 9];
                    for (int k = 0; k < 9; k++) a[k] = (float)(R[k] * pf);
                }
            T2X_TRY(upload(d_Ai, Ai, st));
            d->coarse_Ai_key = akey;
        }
        k_project_coarse<<<(unsigned)O_all, 256, 0, st>>>(
            d_vols.as<const float2*>(), d_pads.as<int>(), d_Ai.as<float>(), (int)Norient,
            tc.d_pix.as<int2>(), tc.d_inproj.as<unsigned char>(), tc.npix, d->coarse_B.as<float>(), Kc,
            d->coarse_B2.as<float>(), Kn, ref2d);
        T2X_HIP(hipGetLastError());
        if (prec != T2X_PREC_FP32) {
            T2X_TRY(d->coarse_Bh.ensure(sizeof(uint16_t) * O_pad * Kc));
            T2X_TRY(d->coarse_Bl.ensure(sizeof(uint16_t) * O_pad * Kc));
            T2X_TRY(d->coarse_Bs.ensure(sizeof(float) * O_pad));
            T2X_TRY(launch_split(prec, d->coarse_B.as<float>(), Kc, Kc, (int)O_pad, (int)O_all, d->coarse_Bh.ptr,
                                 d->coarse_Bl.ptr, Kc, d->coarse_Bs.as<float>(), st));
            // |F|^2 of the N1 term, split the same way
            T2X_TRY(d->coarse_B2h.ensure(sizeof(uint16_t) * O_pad * Kn));
            T2X_TRY(d->coarse_B2l.ensure(sizeof(uint16_t) * O_pad * Kn));
            T2X_TRY(d->coarse_B2s.ensure(sizeof(float) * O_pad));
            T2X_TRY(launch_split(prec, d->coarse_B2.as<float>(), Kn, Kn, (int)O_pad, (int)O_all, d->coarse_B2h.ptr,
                                 d->coarse_B2l.ptr, Kn, d->coarse_B2s.as<float>(), st));
        }
        T2X_HIP(hipStreamSynchronize(st));
        d->coarse_key = key;
    }
    (void)hipEventRecord(d->ev[1], st);
    {
        float ms = 0;
        (void)hipEventSynchronize(d->ev[1]);
        (void)hipEventElapsedTime(&ms, d->ev[0], d->ev[1]);
        d->stats.s.ms_coarse_project += ms;
    }

    // ---- device accumulators ----
    const size_t nws = (size_t)n_og * xN + n_og + K + K * Nd + 1 + 2 * K;
    T2X_TRY(d->wsum_dev.ensure(sizeof(double) * nws));
    T2X_HIP(hipMemsetAsync(d->wsum_dev.ptr, 0, sizeof(double) * nws, st));
    double* w_sigma2 = d->wsum_dev.as<double>();   // [og][xN]
    double* w_sumw = w_sigma2 + (size_t)n_og * xN;   // [og]
    double* w_pdfc = w_sumw + n_og;
    double* w_pdfd = w_pdfc + K;
    double* w_off = w_pdfd + K * Nd;
    double* w_poff = w_off + 1;   // [2K] per-class prior offsets, 2D references

    // ---- batch size: coarse diff2 dominates memory (B*T*O_pad floats) ----
    const size_t noise_bytes = cfg->do_zero_mask ? 0 : (size_t)N * N * 4 + (size_t)N * xN * 8;
    size_t per_particle = (size_t)T * O_pad * 4 + (size_t)T * Kc * (prec == T2X_PREC_FP32 ? 4 : 8) +
                          (size_t)4 * N * N * 4 + (size_t)2 * N * xN * 8 * 2 + noise_bytes;
    if (local_pairs) {
        // expected sample list: directions within 3 sigma (+ one step) of the prior, psi angles
        // within 3 sigma_psi (+ one step) on either side; 16 bytes per sample (pair, diff2)
        const double step_d = ref2d ? 0. : 360. / (6 << cfg->healpix_order), step_p = 360. / (double)Npsi;
        const double rd = std::min(180., 3. * std::max(cfg->sigma_rot, cfg->sigma_tilt) + step_d);
        const double nd_est = (ref2d || !(cfg->sigma_rot > 0.)) ? (double)Nd
                              : std::min((double)Nd, Nd * (1. - std::cos(rd * kPi / 180.)) / 2. + 16.);
        const double np_est = (cfg->sigma_psi > 0.) ? std::min((double)Npsi, Npsi * (6. * cfg->sigma_psi + 2. * step_p) / 360. + 2.)
                                                    : (double)Npsi;
        // per sample list entry 16 bytes; per (particle, orientation) row its Z row (4 bytes per K
        // element), G row (Tl floats) and n1
        const double rows_est = Kact * nd_est * np_est;
        per_particle = (size_t)(16. * rows_est * T + rows_est * (4. * Kc + 4. * Tl + 8.)) + (size_t)4 * N * N * 4 +
                       (size_t)2 * N * xN * 8 * 2 + noise_bytes;
    }
    // scratch for the batch: 24 GiB; 64 GiB for local searches, whose per-particle estimate is
    // a bound (every in-range slice distinct) and whose batches share more slices the larger they
    // are (local360: 16,700 -> 18,600 particles/s at 64 GiB); both far below HBM
    size_t budget = (size_t)(local_pairs ? 64 : 24) << 30;
    const char* env_gb = getenv("T2X_SCRATCH_GB");
    if (env_gb && atoi(env_gb) > 0) budget = (size_t)atoi(env_gb) << 30;
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr / 3 < budget) budget = fr / 3;
    }
    // up to 4,096 particles per batch (small references, e.g. 2D classification, amortise the
    // per-batch launches and synchronisations over more particles)
    int Bmax = (int)std::max<size_t>(1, std::min<size_t>(4096, budget / per_particle));
    const char* env_b = getenv("T2X_BATCH");
    if (env_b && atoi(env_b) > 0) Bmax = std::min(Bmax, atoi(env_b));

    const double pix = cfg->pixel_size;
    // logsigma2 of dLL per optics group (ml_optimiser.cpp:8888-8896)
    std::vector<double> logsigma2(n_og, 0.);
    for (int g = 0; g < n_og; g++)
        for (int i = 0; i < cs; i++) {
            int ip = (i < xc) ? i : i - cs;
            for (int j = 0; j < xc; j++) {
                int ires = relion_round(std::sqrt((double)(ip * ip + j * j)));
                if (!(ires < xc && !(j == 0 && ip < 0))) continue;
                if (ires > 0 && ires < xN) logsigma2[g] += std::log(2. * kPi * sigma2[(size_t)g * xN + ires]);
            }
        }
    std::vector<double> h_power_sum((size_t)n_og * xN, 0.);
    std::vector<double> h_xa(n_grp, 0.), h_aa(n_grp, 0.);
    double LL = 0., Pmax = 0., avg_norm = 0.;
    t2x_stats& stt = d->stats.s;
    stt.coarse_size = cc;
    stt.current_size = cs;
    stt.P_coarse = tc.npix;
    stt.P_fine = tf.npix;

    // ---- pruned coarse pass: the low-frequency prefix of the (|frequency|-ordered) coarse pixels ----
    // (global searches in a split precision, without the proposal or diagnostics). The prefix holds
    // the pixels within 0.82 of the coarse radius (two thirds of them), rounded up to 32 pixels.
    int K_low = 0;
    bool prune_call = !cfg->coarse_exhaustive && prec != T2X_PREC_FP32 && !local && !use_prop &&
                      !(dbg && dbg->n_debug > 0);
    if (prune_call) {
        double frac = 0.82;
        const char* env_pr = getenv("T2X_PRUNE_RADIUS");   // fraction of the coarse radius
        if (env_pr && atof(env_pr) > 0.) frac = atof(env_pr);
        const double rl = frac * (cc / 2.0);
        int n_low = 0;
        while (n_low < tc.npix && tc.h_x[n_low] * tc.h_x[n_low] + tc.h_y[n_low] * tc.h_y[n_low] <= rl * rl) n_low++;
        K_low = (int)round_up(2LL * n_low, 64);
        if (n_low < 16 || K_low >= Kc - 64) prune_call = false;
    }

    // ---- brick backprojection: the call's work units and buffers ----
    const bool bp_bricks = !cfg->bp_direct && K <= BP_KMAX;
    long long bp_item_cap = 0;
    int bp_nbricks = 0;
    float bp_unit_items = 0.f;
    if (bp_bricks) {
        // a batch's items held for the bricks: up to 256 per particle and 2 GiB of pixel values
        // (the items beyond are scattered directly)
        const int Bc = std::min(Bmax, n);
        const long long win = (long long)tf.px_w * tf.py_h;   // frequencies of the value layout
        bp_item_cap = std::min<long long>((long long)Bc * 256, std::max<long long>(1, (2ll << 30) / (win * 16)));
        const char* env_cap = getenv("T2X_BP_ITEM_CAP");   // tests: a small capacity sends items direct
        if (env_cap && atoll(env_cap) > 0) bp_item_cap = atoll(env_cap);
        // the bricks within reach of the insertion radius, each with the expected share of its
        // class's items that reach it: for a 3D brick at distance r from the origin the share of
        // planes through the origin that cross it, about (half-diagonal) / r; every item of a
        // 2D class reaches all its tiles. Per batch a unit of work takes about unit_items of the
        // items that reach its brick.
        double unit_items = 32.;
        const char* env_ui = getenv("T2X_BP_UNIT_ITEMS");
        if (env_ui && atof(env_ui) > 0.) unit_items = atof(env_ui);
        const int S = ref2d ? BP_S2 : BP_S3;
        const double hd = 0.5 * std::sqrt(3.) * (S + 1);
        std::vector<BpBrick> bricks;
        for (int c = 0; c < K; c++) {
            const int bpad = h_bpads[c], bxdim = bpad / 2 + 1, bh = bpad / 2;
            // the insertion radius (backrotate2D truncates r_max * pf, backproject2Dto3D rounds it)
            const int rr = ref2d ? (int)(h_bprmax[c] * pf) : relion_round(h_bprmax[c] * pf);
            const int nbx = (bxdim + S - 1) / S, nby = (bpad + S - 1) / S, nbz = ref2d ? 1 : nby;
            for (int bz = 0; bz < nbz; bz++)
                for (int by = 0; by < nby; by++)
                    for (int bx = 0; bx < nbx; bx++) {
                        // positions (centred) inserted into the brick: [lo, hi) per axis
                        const int nd = ref2d ? 2 : 3;
                        const double lo[3] = {bx * S - 1., by * S - bh - 1., bz * S - bh - 1.};
                        const double hi[3] = {bx * S + (double)S, by * S - bh + (double)S, bz * S - bh + (double)S};
                        double dn2 = 0., dc2 = 0.;
                        for (int k = 0; k < nd; k++) {
                            const double q = std::max(lo[k], std::min(0., hi[k]));   // nearest to the origin
                            const double m = 0.5 * (lo[k] + hi[k]);
                            dn2 += q * q;
                            dc2 += m * m;
                        }
                        if (dn2 > (double)(rr + 1) * (rr + 1)) continue;   // beyond r_max: nothing lands
                        const double f = ref2d ? 1. : std::min(1., hd / std::max(1., std::sqrt(dc2)));
                        bricks.push_back(BpBrick{c, bx, by, bz, (float)f, 0, 0, 0});
                    }
        }
        bp_nbricks = (int)bricks.size();
        bp_unit_items = (float)unit_items;
        T2X_TRY(upload(d->bp_units, bricks, st));
        T2X_TRY(d->bp_boff.ensure(sizeof(int) * (bp_nbricks + 2)));
        T2X_TRY(d->bp_cls.ensure(sizeof(int) * (K + 1 + bp_item_cap)));
        T2X_TRY(d->bp_filt.ensure(sizeof(float4) * bp_item_cap));
        T2X_TRY(d->bp_geom.ensure(sizeof(double) * 6 * bp_item_cap));
        // frequencies outside the pixel set are never written: zero for the call's layout
        T2X_TRY(d->bp_val.ensure(sizeof(float4) * bp_item_cap * win));
        T2X_TRY(zero_async(d->bp_val.ptr, sizeof(float4) * bp_item_cap * win, st));
    }

    for (int b0 = 0; b0 < n; b0 += Bmax) {
        const int B = std::min(Bmax, n - b0);
        const bool prune_batch = prune_call;   // the pruned coarse pass for this batch
        // its full GEMM's tiles (the 256x256 tiles of k_gemm_x3p), known here for graph replays too
        const long long prune_ntiles = prune_batch ? (round_up((long long)B * T, 256) / 256) * (O_pad / 256) : 0;
        const size_t NN = (size_t)N * N;
        int ev = 0;
        (void)hipEventRecord(d->ev[ev++], st);
        // ---- inputs ----
        const float* d_img;
        if (images_on_device) d_img = images + (size_t)b0 * NN;
        else {
            T2X_TRY(d->img.ensure(sizeof(float) * B * NN));
            T2X_HIP(hipMemcpyAsync(d->img.ptr, images + (size_t)b0 * NN, sizeof(float) * B * NN,
                                   hipMemcpyHostToDevice, st));
            d_img = d->img.as<float>();
        }
        std::vector<PartScalars> hps(B);
        std::vector<CtfConst> hctf(B);
        for (int i = 0; i < B; i++) {
            const t2x_particle_in& p = pin[b0 + i];
            PartScalars& s = hps[i];
            std::memset(&s, 0, sizeof(s));
            s.old_x = relion_round(p.old_offset_x);
            s.old_y = relion_round(p.old_offset_y);
            double px_ = p.prior_offset_x, py_ = p.prior_offset_y;
            if (px_ > 998.99 && px_ < 999.01) px_ = 0.;
            if (py_ > 998.99 && py_ < 999.01) py_ = 0.;
            s.prior_x = ref2d ? class_prior_x : px_;
            s.prior_y = ref2d ? class_prior_y : py_;
            s.norm_scale = cfg->do_norm_correction ? cfg->avg_norm_correction / p.norm_correction : 1.0;
            s.og = p.optics_group;
            s.grp = p.group_id;
            s.part_id = p.part_id;
            s.scale = cfg->do_scale_correction ? model->scale_correction[p.group_id] : 1.0;
            s.ctf_store = cfg->do_scale_correction ? std::max(0.001, s.scale) / s.scale : 1.0;
            hctf[i] = ctf_const(p.ctf);
        }
        T2X_TRY(upload(d->pscal, hps, st));
        DevBuf& d_ctf = d->ctfc;
        T2X_TRY(upload(d_ctf, hctf, st));
        PartScalars* d_ps = d->pscal.as<PartScalars>();
        T2X_TRY(d->prior_t.ensure(sizeof(double) * B * T * (tK ? tK : 1)));
        k_prior_t<<<B, 128, 0, st>>>(d_ps, B, d_ctx.as<double>(), d_cty.as<double>(), (int)T, pix,
                                     cfg->sigma2_offset, tK ? d->cls_prior.as<double>() : nullptr, tK,
                                     d->prior_t.as<double>());
        T2X_HIP(hipGetLastError());
        // local searches: each particle's direction / psi priors around its prior angles (the
        // current angles where the priors are unset, ml_optimiser.cpp:5851-5857)
        LocalPrior lp{};
        if (local) {
            std::vector<double> pang((size_t)3 * B);
            auto unset = [](double v) { return v > 998.99 && v < 999.01; };
            for (int i = 0; i < B; i++) {
                const t2x_particle_in& p = pin[b0 + i];
                pang[3 * i] = unset(p.prior_rot) ? p.rot : p.prior_rot;
                pang[3 * i + 1] = unset(p.prior_tilt) ? p.tilt : p.prior_tilt;
                pang[3 * i + 2] = unset(p.prior_psi) ? p.psi : p.prior_psi;
                if (!std::isfinite(pang[3 * i]) || !std::isfinite(pang[3 * i + 1]) || !std::isfinite(pang[3 * i + 2])) {
                    set_error("local searches: non-finite prior angles for particle " + std::to_string(b0 + i));
                    return T2X_EINVAL;
                }
            }
            T2X_TRY(upload(d->loc_ang, pang, st));
            T2X_TRY(d->loc_dir.ensure(sizeof(double) * B * Nd));
            T2X_TRY(d->loc_psi.ensure(sizeof(double) * B * Npsi));
            T2X_TRY(d->loc_mean.ensure(sizeof(double) * B));
            T2X_TRY(d->loc_list.ensure(sizeof(int) * B * (Nd + Npsi)));
            T2X_TRY(d->loc_nsel.ensure(sizeof(int) * 2 * B));
            k_local_prior<<<B, 256, 0, st>>>(d->loc_ang.as<double>(), B, lsp, d->loc_dir.as<double>(),
                                            d->loc_psi.as<double>(), d->loc_mean.as<double>(),
                                            d->loc_list.as<int>(), d->loc_nsel.as<int>());
            T2X_HIP(hipGetLastError());
            lp.dir = d->loc_dir.as<double>();
            lp.psi = d->loc_psi.as<double>();
            lp.mean = d->loc_mean.as<double>();
            lp.Nd = (int)Nd;
            lp.Npsi = (int)Npsi;
        }
        // ---- preprocessing ----
        // the batched FFTs run over the call's largest batch whatever B (the plans are made once
        // per call shape, not again for a smaller last batch; transforms beyond B are not read)
        const int Bfft = std::min(Bmax, n);
        T2X_TRY(d->img2.ensure(sizeof(float) * 2 * Bfft * NN));
        T2X_TRY(d->fft_out.ensure(sizeof(float2) * 2 * Bfft * N * xN));
        T2X_TRY(d->Fimg_cur.ensure(sizeof(float2) * B * Pw));
        T2X_TRY(d->Fimg_nm_cur.ensure(sizeof(float2) * B * Pw));
        T2X_TRY(d->Fctf_cur.ensure(sizeof(float) * B * Pw));
        T2X_TRY(d->power.ensure(sizeof(double) * B * xN));
        const float* d_noise = nullptr;
        if (!cfg->do_zero_mask) {
            // the mask's background: per particle a noise image with its optics group's spectrum
            T2X_TRY(d->noise_spec.ensure(sizeof(float2) * Bfft * N * xN));
            T2X_TRY(d->noise_img.ensure(sizeof(float) * Bfft * NN));
            k_noise_spectrum<<<B, 256, 0, st>>>(N, d_ps, d->sigma2_tab.as<double>(), cfg->sigma2_fudge,
                                                cfg->random_seed ^ 0x6d61736b6e6f6973ULL, d->noise_spec.as<float2>());
            T2X_HIP(hipGetLastError());
            if (!d->noise_plan || d->noise_n != N || d->noise_batch != Bfft) {
                if (d->noise_plan) { (void)hipfftDestroy(d->noise_plan); d->noise_plan = 0; }
                int dims[2] = {N, N};
                if (hipfftPlanMany(&d->noise_plan, 2, dims, nullptr, 1, N * xN, nullptr, 1, N * N, HIPFFT_C2R, Bfft) !=
                        HIPFFT_SUCCESS ||
                    hipfftSetStream(d->noise_plan, st) != HIPFFT_SUCCESS) {
                    set_error("hipfft C2R plan for the mask background failed");
                    return T2X_EDEVICE;
                }
                d->noise_n = N;
                d->noise_batch = Bfft;
            }
            if (hipfftExecC2R(d->noise_plan, d->noise_spec.as<hipfftComplex>(), d->noise_img.as<hipfftReal>()) !=
                HIPFFT_SUCCESS) {
                set_error("hipfftExecC2R failed");
                return T2X_EDEVICE;
            }
            d_noise = d->noise_img.as<float>();
        }
        T2X_TRY(d->prep_part.ensure(sizeof(double) * 2 * PREP_BANDS * B));
        if (!d_noise)
            k_prep_bg<<<dim3(B, PREP_BANDS), 256, 0, st>>>(d_img, N, d_ps, cfg->particle_diameter / (2. * pix),
                                                          (double)cfg->width_mask_edge, d->prep_part.as<double>());
        k_prep_mask<<<dim3(B, PREP_BANDS), 256, 0, st>>>(d_img, N, B, d_ps, cfg->particle_diameter / (2. * pix),
                                                        (double)cfg->width_mask_edge, d_noise,
                                                        d->prep_part.as<double>(), d->img2.as<float>());
        T2X_HIP(hipGetLastError());
        T2X_TRY(ensure_fft(d, N, 2 * Bfft));
        if (hipfftExecR2C(d->fft_plan, d->img2.as<hipfftReal>(), d->fft_out.as<hipfftComplex>()) != HIPFFT_SUCCESS) {
            set_error("hipfftExecR2C failed");
            return T2X_EDEVICE;
        }
        k_prep_fourier<<<B, 256, sizeof(double) * xN, st>>>(
            d->fft_out.as<float2>(), N, B, cs, d_ctf.as<CtfConst>(), pix, cfg->do_ctf_correction,
            d->Fimg_cur.as<float2>(), d->Fimg_nm_cur.as<float2>(), d->Fctf_cur.as<float>(),
            d->power.as<double>(), d_ps);
        T2X_HIP(hipGetLastError());
        k_const_term<<<B, 256, 0, st>>>(d->Fimg_cur.as<float2>(), Pw, tc.d_win.as<int>(),
                                        tc.d_minv.as<float>(), tc.npix, d_ps, 0);
        k_const_term<<<B, 256, 0, st>>>(d->Fimg_cur.as<float2>(), Pw, tf.d_win.as<int>(),
                                        d_minv_fd.as<float>(), tf.npix, d_ps, 1);
        T2X_HIP(hipGetLastError());
        (void)hipEventRecord(d->ev[ev++], st);   // 1: preprocess done

        // ---- coarse pass ----
        const long long M = (long long)B * T;
        T2X_TRY(d->sel_out.ensure(sizeof(SelResult) * 2 * B));
        SelResult* d_selc = d->sel_out.as<SelResult>();
        SelResult* d_self = d_selc + B;
        T2X_TRY(d->sig_count.ensure(sizeof(long long) * 2 * B));
        std::vector<SelResult> hselc(B);
        // particles sent to the radix select (candidate path: those whose candidates overflowed;
        // null: every particle, as in local searches)
        const int* d_radix = nullptr;
        bool cand_path = false;
        WeightSrc wc{};
        long long npairs_c = 0;
        std::vector<unsigned> prop_mask_host;   // learned proposal: per particle, the bins scored
        // Launches of the batch after its preprocessing, in two parts: the coarse pass with its
        // selection, then (from the batch plan on) the fine pass, weights and store. With
        // launch_graph they are captured once per batch size as a hipGraph and replayed.
        bool capturing = false;
        auto mark = [&]() {   // stage events (not recorded into a graph)
            if (!capturing) (void)hipEventRecord(d->ev[ev], st);
            ev++;
        };
        auto coarse_part = [&]() -> int {
        if (local_pairs) {
            // local search on per-particle sample lists: each particle's orientations in range x all
            // translations, scored by a GEMM of the batch's (particle, orientation) rows against
            // the phase table (no dense pass over every orientation); weights, significance and the
            // significant pairs as in the dense pass, read from the lists
            std::vector<int> hnsel(2 * B);
            T2X_HIP(hipMemcpyAsync(hnsel.data(), d->loc_nsel.ptr, sizeof(int) * 2 * B, hipMemcpyDeviceToHost, st));
            T2X_HIP(hipStreamSynchronize(st));
            std::vector<long long> coff(B + 1, 0), cpoff(B + 1, 0);
            for (int i = 0; i < B; i++) coff[i + 1] = coff[i] + (long long)Kact * hnsel[2 * i] * hnsel[2 * i + 1];
            for (int i = 0; i <= B; i++) cpoff[i] = coff[i] * T;
            const long long norients_c = coff[B];
            npairs_c = cpoff[B];
            if (norients_c > 0x7fffffffLL) {
                set_error("local search: more than 2^31 orientations in one batch");
                return T2X_EINVAL;
            }
            T2X_TRY(upload(d->lc_orient_off, coff, st));
            T2X_TRY(upload(d->lc_pair_off, cpoff, st));
            T2X_TRY(d->lc_orients.ensure(sizeof(int) * std::max<long long>(norients_c, 1)));
            T2X_TRY(d->lc_pairs.ensure(sizeof(int2) * std::max<long long>(npairs_c, 1)));
            T2X_TRY(d->lc_d2.ensure(sizeof(float) * std::max<long long>(npairs_c, 1)));
            T2X_TRY(d->lc_min.ensure(sizeof(unsigned) * B));
            k_local_pairs<<<B, 256, 0, st>>>(B, d->loc_list.as<int>(), (int)Nd, (int)Npsi, d->loc_nsel.as<int>(),
                                            d->lc_cls.as<int>(), Kact, (long long)Norient, (int)T,
                                            d->lc_orient_off.as<long long>(), d->lc_orients.as<int>(),
                                            d->lc_pairs.as<int2>());
            k_fill_u32<<<(B + 255) / 256, 256, 0, st>>>(d->lc_min.as<unsigned int>(), 0xffffffffu, B);
            T2X_HIP(hipGetLastError());
            FineGeom gc{};
            gc.Norient = (int)Norient;
            gc.Npsi = (int)Npsi;
            gc.n_subdir = 1;
            gc.n_subpsi = 1;
            gc.over_rot = 1;
            gc.over_trans = 1;
            gc.T = (int)T;
            gc.dir_sc = d->lc_dirsc.as<double>();
            gc.psi_sc = d->lc_psisc.as<double>();
            gc.pf = pf;
            gc.ref2d = ref2d;
            // the batch's distinct orientations (particles sorted by prior orientation share many)
            // distinct orientations of the batch (increasing order: a direction's psi rotations,
            // which share the volume's central plane, are adjacent) and each row's slice
            long long nu = 0;
            if (norients_c > 0) {
                T2X_TRY(d->lc_flags.ensure(sizeof(int) * (O_all + 1)));
                T2X_TRY(d->lc_pos.ensure(sizeof(int) * (O_all + 1)));
                k_fill_u32<<<(unsigned)((O_all + 1 + 255) / 256), 256, 0, st>>>(d->lc_flags.as<unsigned>(), 0u, (int)(O_all + 1));
                k_mark_orients<<<(unsigned)((norients_c + 255) / 256), 256, 0, st>>>(norients_c, d->lc_orients.as<int>(),
                                                                                   d->lc_flags.as<int>());
                size_t tmp = 0;
                if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, d->lc_flags.as<int>(), d->lc_pos.as<int>(),
                                                     (int)(O_all + 1), st) != hipSuccess) {
                    set_error("hipcub scan size query failed");
                    return T2X_EDEVICE;
                }
                T2X_TRY(d->lc_scan_tmp.ensure(std::max<size_t>(tmp, 1)));
                if (hipcub::DeviceScan::ExclusiveSum(d->lc_scan_tmp.ptr, tmp, d->lc_flags.as<int>(), d->lc_pos.as<int>(),
                                                     (int)(O_all + 1), st) != hipSuccess) {
                    set_error("hipcub scan failed");
                    return T2X_EDEVICE;
                }
                int hnu = 0;   // pos[O_all]: the number of distinct orientations (flags[O_all] = 0)
                T2X_HIP(hipMemcpyAsync(&hnu, d->lc_pos.as<int>() + O_all, sizeof(int), hipMemcpyDeviceToHost, st));
                T2X_HIP(hipStreamSynchronize(st));
                nu = hnu;
                T2X_TRY(d->lc_ulist.ensure(sizeof(int) * std::max<long long>(nu, 1)));
                T2X_TRY(d->lc_rowu.ensure(sizeof(int) * norients_c));
                k_compact_orients<<<(unsigned)((O_all + 255) / 256), 256, 0, st>>>(O_all, d->lc_flags.as<int>(),
                                                                                 d->lc_pos.as<int>(), d->lc_ulist.as<int>());
                k_row_slices<<<(unsigned)((norients_c + 255) / 256), 256, 0, st>>>(norients_c, d->lc_orients.as<int>(),
                                                                                 d->lc_pos.as<int>(), d->lc_rowu.as<int>());
                T2X_HIP(hipGetLastError());
            }
            stt.local_slices += (double)nu;
            stt.local_rows += (double)norients_c;
            const int ldf = tc.npix_pad;
            if (prec == T2X_PREC_FP32) T2X_TRY(d->lc_F.ensure(sizeof(float2) * std::max<long long>(nu, 1) * ldf));
            else T2X_TRY(d->lc_F2.ensure(sizeof(float) * std::max<long long>(nu, 1) * ldf));
            mark();   // 2: coarse diff2 start
            if (norients_c > 0) {
                const unsigned ugrid = (unsigned)round_up(nu, 8);
                if (prec == T2X_PREC_FP32) {
                    k_project_list<_Float16, 0><<<ugrid, 256, 0, st>>>(
                        nu, d->lc_ulist.as<int>(), gc, d_vols.as<const float2*>(), d_pads.as<int>(), tc.d_pix.as<int2>(),
                        tc.d_inproj.as<unsigned char>(), tc.npix, ldf, d->lc_F.as<float2>(), nullptr, nullptr, nullptr,
                        0, nullptr);
                } else {
                    T2X_TRY(d->lc_fh.ensure(sizeof(uint16_t) * nu * Kc));
                    T2X_TRY(d->lc_fl.ensure(sizeof(uint16_t) * nu * Kc));
                    T2X_TRY(d->lc_fs.ensure(sizeof(float) * nu));
                    const size_t lds = sizeof(float) * Kc;
                    if (prec == T2X_PREC_BF16X3) {
                        k_project_list<__bf16, 2><<<ugrid, 256, 0, st>>>(
                            nu, d->lc_ulist.as<int>(), gc, d_vols.as<const float2*>(), d_pads.as<int>(),
                            tc.d_pix.as<int2>(), tc.d_inproj.as<unsigned char>(), tc.npix, ldf, nullptr,
                            (__bf16*)d->lc_fh.ptr, (__bf16*)d->lc_fl.ptr, d->lc_fs.as<float>(), 0, d->lc_F2.as<float>());
                    } else if (lds > kRowLdsMax) {
                        // rows beyond the LDS row buffer: FP32 slices in HBM, then the same split
                        // (row maximum, power-of-two scale, hi / lo) and |F|^2 from them: the same bits
                        T2X_TRY(d->lc_F.ensure(sizeof(float2) * nu * ldf));
                        k_project_list<_Float16, 0><<<ugrid, 256, 0, st>>>(
                            nu, d->lc_ulist.as<int>(), gc, d_vols.as<const float2*>(), d_pads.as<int>(),
                            tc.d_pix.as<int2>(), tc.d_inproj.as<unsigned char>(), tc.npix, ldf, d->lc_F.as<float2>(),
                            nullptr, nullptr, nullptr, 0, nullptr);
                        // k_project_list leaves the padding columns npix..ldf-1 unwritten in FP32 rows
                        k_zero_row_tails<<<(unsigned)nu, 64, 0, st>>>(d->lc_F.as<float2>(), tc.npix, ldf);
                        T2X_HIP(hipGetLastError());
                        T2X_TRY(launch_split(prec, d->lc_F.as<float>(), Kc, Kc, (int)nu, (int)nu, d->lc_fh.ptr,
                                             d->lc_fl.ptr, Kc, d->lc_fs.as<float>(), st));
                        k_abs2<<<(unsigned)std::min<long long>((nu * ldf + 255) / 256, 65536), 256, 0, st>>>(
                            d->lc_F.as<float2>(), nu * ldf, d->lc_F2.as<float>());
                    } else
                        k_project_list<_Float16, 1><<<ugrid, 256, lds, st>>>(
                            nu, d->lc_ulist.as<int>(), gc, d_vols.as<const float2*>(), d_pads.as<int>(),
                            tc.d_pix.as<int2>(), tc.d_inproj.as<unsigned char>(), tc.npix, ldf, nullptr,
                            (_Float16*)d->lc_fh.ptr, (_Float16*)d->lc_fl.ptr, d->lc_fs.as<float>(), 1, d->lc_F2.as<float>());
                }
                T2X_HIP(hipGetLastError());
                // rows = the batch's (particle, orientation) pairs, columns = translations
                const long long rows = norients_c, rows_pad = round_up(rows, 128);
                T2X_TRY(d->lc_n1.ensure(sizeof(double) * rows));
                if (prec == T2X_PREC_FP32) {
                    // FP32: (particle, orientation) operand rows against the phase table
                    EpiParams eg{};
                    T2X_TRY(d->lc_G.ensure(sizeof(float) * rows_pad * Tl));
                    eg.C = d->lc_G.as<float>();
                    eg.ldc = Tl;
                    eg.M = (int)rows;
                    eg.N = (int)T;
                    T2X_TRY(d->lc_z.ensure(sizeof(float) * rows_pad * Kc));
                    k_local_zrows<<<(unsigned)rows_pad, 256, 0, st>>>(
                        B, rows, d->lc_orient_off.as<long long>(), d->lc_rowu.as<int>(), d->lc_F.as<float2>(), ldf,
                        tc.d_minv.as<float>(), tc.d_win.as<int>(), tc.npix, d_ps,
                        d->Fimg_cur.as<float2>(), d->Fctf_cur.as<float>(), Pw, Kc, d->lc_z.as<float>(),
                        d->lc_n1.as<double>());
                    T2X_HIP(hipGetLastError());
                    T2X_TRY(launch_gemm(EPI_STORE, d->lc_z.as<float>(), Kc, d->lc_ab.as<float>(), Kc, Kc,
                                        (int)rows_pad, Tl, eg, st));
                    k_local_diff2<<<(unsigned)rows, 128, 0, st>>>(B, rows, d->lc_orient_off.as<long long>(), d_ps,
                                                                 d->lc_n1.as<double>(), d->lc_G.as<float>(), Tl, (int)T,
                                                                 d->lc_d2.as<float>(), d->lc_min.as<unsigned int>());
                    T2X_HIP(hipGetLastError());
                } else {
                    // grouped split GEMM: each particle's slices (gathered) x its translation rows
                    // the image side, 128 readable rows past the last particle
                    const long long wrows = round_up(M + 128, 128);
                    T2X_TRY(d->Ah.ensure(sizeof(uint16_t) * wrows * Kc));
                    T2X_TRY(d->Al.ensure(sizeof(uint16_t) * wrows * Kc));
                    T2X_TRY(d->As.ensure(sizeof(float) * wrows));
                    T2X_TRY(d->A1.ensure(sizeof(float) * B * Kn));
                    T2X_TRY(imgside_split(d, prec, tc, d_ps, Pw, (int)T, B, wrows, Kc, Kn, st));
                    k_local_n1<<<(unsigned)((rows + 3) / 4), 256, 0, st>>>(
                        B, rows, d->lc_orient_off.as<long long>(), d->lc_rowu.as<int>(), d->lc_F2.as<float>(), ldf,
                        d->A1.as<float>(), Kn, tc.npix, d->lc_n1.as<double>());
                    T2X_HIP(hipGetLastError());
                    std::vector<int2> items;
                    for (int i = 0; i < B; i++)
                        for (long long s0 = 0; s0 < coff[i + 1] - coff[i]; s0 += 128) items.push_back(make_int2(i, (int)s0));
                    T2X_TRY(upload(d->lc_items, items, st));
                    T2X_TRY(d->tables.ensure(sizeof(double) * B));
                    T2X_HIP(hipMemcpy2DAsync(d->tables.ptr, sizeof(double), (char*)d_ps + offsetof(PartScalars, C_coarse),
                                             sizeof(PartScalars), sizeof(double), B, hipMemcpyDeviceToDevice, st));
                    T2X_TRY(launch_gemm_local_x3(prec, d->lc_fh.ptr, d->lc_fl.ptr, d->lc_fs.as<float>(), d->Ah.ptr,
                                                 d->Al.ptr, d->As.as<float>(), Kc, Kc, d->lc_items.as<int2>(),
                                                 (int)items.size(), (int)T, d->lc_orient_off.as<long long>(),
                                                 d->lc_rowu.as<int>(), d->tables.as<double>(), d->lc_n1.as<double>(),
                                                 d->lc_d2.as<float>(), d->lc_min.as<unsigned int>(), st));
                }
            }
            mark();   // 3: coarse diff2 end
            wc.kind = SRC_FINE;
            wc.d2 = d->lc_d2.as<float>();
            wc.pair_off = d->lc_pair_off.as<long long>();
            wc.pairs = d->lc_pairs.as<int2>();
            wc.fine_min = d->lc_min.as<unsigned int>();
            wc.nover = 1;
            wc.over_trans = 1;
            wc.T = (int)T;
            wc.O_all = (int)O_all;
            wc.prior_o = d->prior_o.as<double>();
            wc.prior_t = d->prior_t.as<double>();
            wc.tK = tK;
            wc.tcols = (int)Norient;
            wc.lp = lp;
            k_pass_min<<<B, 1024, 0, st>>>(wc, B, d_selc);
            SelParams sp{cfg->adaptive_fraction, (long long)cfg->maximum_significants, 1};
            k_select<<<B, 1024, 0, st>>>(wc, sp, B, d_selc, nullptr);
            k_sig_count<<<B, 1024, 0, st>>>(wc, B, d_selc, d->sig_count.as<long long>(), nullptr);
            T2X_HIP(hipGetLastError());
            mark();   // 4: coarse weights done
        } else {
            const long long M_pad = round_up(M, 256);
            const long long B_pad = round_up(B, prec == T2X_PREC_FP32 ? 128 : 256);   // N1 GEMM rows
            if (prec == T2X_PREC_FP32) T2X_TRY(d->A.ensure(sizeof(float) * M_pad * Kc));
            T2X_TRY(d->A1.ensure(sizeof(float) * B_pad * Kn));
            T2X_TRY(d->N1.ensure(sizeof(float) * B_pad * O_pad));
            T2X_TRY(d->diff2.ensure(sizeof(float) * M_pad * O_pad));
            const int rowmin_ld = (int)(O_pad / 64);
            DevBuf& d_rowmin = d->minb;
            T2X_TRY(d_rowmin.ensure(sizeof(float) * M_pad * rowmin_ld));
            T2X_TRY(zero_async(d->A1.ptr, sizeof(float) * B_pad * Kn, st));
            if (prec == T2X_PREC_FP32) {
                T2X_TRY(zero_async(d->A.ptr, sizeof(float) * M_pad * Kc, st));
                dim3 grid((unsigned)T, (unsigned)B);
                k_coarse_imgside<<<grid, 256, 0, st>>>(d->Fimg_cur.as<float2>(), d->Fctf_cur.as<float>(), Pw,
                                                       tc.d_win.as<int>(), tc.d_minv.as<float>(), tc.npix, d_ps,
                                                       d->ab_coarse.as<float2>(), (int)T, d->A.as<float>(), Kc,
                                                       d->A1.as<float>(), Kn);
                T2X_HIP(hipGetLastError());
            } else {
                // the image side straight into the split operand (hi, lo, per-row scale)
                T2X_TRY(d->Ah.ensure(sizeof(uint16_t) * M_pad * Kc));
                T2X_TRY(d->Al.ensure(sizeof(uint16_t) * M_pad * Kc));
                T2X_TRY(d->As.ensure(sizeof(float) * M_pad));
                T2X_TRY(imgside_split(d, prec, tc, d_ps, Pw, (int)T, B, M_pad, Kc, Kn, st));
            }
            {
                EpiParams ep{};
                ep.C = d->N1.as<float>();
                ep.ldc = O_pad;
                ep.M = B;
                ep.N = (int)O_all;
                if (prec == T2X_PREC_FP32) {
                    T2X_TRY(launch_gemm(EPI_STORE, d->A1.as<float>(), Kn, d->coarse_B2.as<float>(), Kn, Kn,
                                        (int)B_pad, (int)O_pad, ep, st));
                } else {   // N1 = sum_p (minv ctf^2 / 2) |F|^2 in the same split precision as the cross term
                    T2X_TRY(d->A1h.ensure(sizeof(uint16_t) * B_pad * Kn));
                    T2X_TRY(d->A1l.ensure(sizeof(uint16_t) * B_pad * Kn));
                    T2X_TRY(d->A1s.ensure(sizeof(float) * B_pad));
                    T2X_TRY(launch_split(prec, d->A1.as<float>(), Kn, Kn, (int)B_pad, (int)B, d->A1h.ptr, d->A1l.ptr, Kn,
                                         d->A1s.as<float>(), st));
                    // (3 products also under FP16X2: N1 is a sum of positive terms, large beside diff2)
                    T2X_TRY(launch_gemm_x3(prec == T2X_PREC_FP16X2 ? T2X_PREC_FP16X3 : prec, EPI_STORE, d->A1h.ptr,
                                           d->A1l.ptr, d->A1s.as<float>(), Kn, B_pad,
                                           d->coarse_B2h.ptr, d->coarse_B2l.ptr, d->coarse_B2s.as<float>(), Kn, Kn,
                                           (int)B_pad, (int)O_pad, ep, st,
                                           GemmChoice{cfg->gemm_tile, cfg->gemm_group}, nullptr, nullptr));
                    if (prune_batch) {   // N1h: the same over the pixels beyond the low-frequency prefix
                        T2X_TRY(d->prune_n1.ensure(sizeof(float) * B_pad * O_pad));
                        EpiParams el = ep;
                        el.C = d->prune_n1.as<float>();
                        const int k0 = K_low / 2;   // a multiple of 32 elements: 64-byte aligned rows
                        T2X_TRY(launch_gemm_x3(prec == T2X_PREC_FP16X2 ? T2X_PREC_FP16X3 : prec, EPI_STORE,
                                               (const uint16_t*)d->A1h.ptr + k0, (const uint16_t*)d->A1l.ptr + k0,
                                               d->A1s.as<float>(), Kn, B_pad, (const uint16_t*)d->coarse_B2h.ptr + k0,
                                               (const uint16_t*)d->coarse_B2l.ptr + k0, d->coarse_B2s.as<float>(),
                                               Kn, Kn - k0, (int)B_pad, (int)O_pad, el, st,
                                               GemmChoice{cfg->gemm_tile, cfg->gemm_group}, nullptr, nullptr));
                    }
                }
                if (local) {
                    const long long ne = (long long)B * O_all;
                    k_local_mask_n1<<<(unsigned)((ne + 255) / 256), 256, 0, st>>>(d->N1.as<float>(), O_pad, B, O_all, lp);
                    T2X_HIP(hipGetLastError());
                }
            }
            // ---- learned proposal: bins per particle, work lists of the restricted coarse GEMM ----
            int n_work = 0, n_fb_rows = 0;
            if (use_prop) {
                const auto tp0 = std::chrono::steady_clock::now();
                std::vector<int> pbins, pfb;
                std::vector<float> pmax, pmass;
                T2X_TRY(cnn::proposal_infer(d, d->Fimg_cur.as<float2>(), cs, B, topk, cfg->proposal_min_conf,
                                            cfg->proposal_gate, pbins, pmax, pmass, pfb));
                d->prop_gate_top1.insert(d->prop_gate_top1.end(), pmax.begin(), pmax.end());
                d->prop_gate_mass.insert(d->prop_gate_mass.end(), pmass.begin(), pmass.end());
                const int mw = (n_bins + 31) / 32;
                std::vector<unsigned> mask((size_t)B * mw, 0u);
                std::vector<std::vector<int>> members(n_bins);
                std::vector<int> fb_rows;   // exhaustive particles: all their rows through the dense kernel
                long long scored = 0;
                for (int i = 0; i < B; i++) {
                    if (pfb[i]) {
                        for (int j = 0; j < mw; j++) mask[(size_t)i * mw + j] = ~0u;
                        for (long t = 0; t < T; t++) fb_rows.push_back((int)(i * T + t));
                        scored += n_bins;
                        d->stats.s.n_proposal_fallback++;
                    } else {
                        for (int k = 0; k < topk; k++) {
                            const int j = pbins[(size_t)i * topk + k];
                            mask[(size_t)i * mw + (j >> 5)] |= 1u << (j & 31);
                            members[j].push_back(i);
                        }
                        scored += topk;
                    }
                }
                std::vector<int> plist;
                std::vector<BinWorkHost> work;
                for (int j = 0; j < n_bins; j++) {
                    if (members[j].empty()) continue;
                    const int pstart = (int)plist.size(), cnt = (int)members[j].size();
                    plist.insert(plist.end(), members[j].begin(), members[j].end());
                    const long long rows = (long long)cnt * T;
                    for (int rt = 0; (long long)rt * 128 < rows; rt++)
                        for (int c0 = 0; c0 < bin_cols; c0 += 128)
                            work.push_back(BinWorkHost{pstart, cnt, rt, j * bin_cols + c0, std::min(128, bin_cols - c0)});
                }
                T2X_TRY(upload(d->prop_mask, mask, st));
                prop_mask_host = mask;
                T2X_TRY(upload(d->prop_plist, plist, st));
                T2X_TRY(upload(d->prop_work, work, st));
                T2X_TRY(upload(d->prop_rows, fb_rows, st));
                n_work = (int)work.size();
                n_fb_rows = (int)fb_rows.size();
                d->stats.s.n_proposal_particles += B;
                d->stats.s.proposal_bins_scored += (double)scored;
                d->stats.s.ms_proposal +=
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp0).count();
            }
            mark();   // 2: coarse gemm start
            {
                EpiParams ep{};
                ep.C = d->diff2.as<float>();
                ep.ldc = O_pad;
                ep.M = (int)M;
                ep.N = (int)O_all;
                ep.T = (int)T;
                // C_b is read from the PartScalars array: pass a strided view via a packed copy
                ep.Cb = nullptr;
                DevBuf& d_cb = d->tables;
                T2X_TRY(d_cb.ensure(sizeof(double) * B));
                k_gather_ccoarse<<<(unsigned)((B + 255) / 256), 256, 0, st>>>(d_ps, B, d_cb.as<double>());
                ep.Cb = d_cb.as<double>();
                ep.N1 = d->N1.as<float>();
                ep.ldn = O_pad;
                ep.prior_o = d->prior_o.as<double>();
                ep.rowmin = d_rowmin.as<float>();
                ep.rowmin_ld = rowmin_ld;
                // running per-particle minima for the lazy stores of negligible blocks (all values are
                // stored when diagnostic copies of the coarse diff2 are requested)
                T2X_TRY(d->pmin.ensure(sizeof(unsigned) * B));
                k_fill_u32<<<(unsigned)((B + 255) / 256), 256, 0, st>>>(d->pmin.as<unsigned>(), 0xFF800000u, B);  // +inf
                ep.pmin = (dbg && dbg->n_debug > 0) ? nullptr : d->pmin.as<unsigned>();
                // rows beyond the valid columns of the last tile keep +inf
                k_fill_f32<<<(unsigned)((M_pad * rowmin_ld + 255) / 256), 256, 0, st>>>(d_rowmin.as<float>(), 3.0e38f,
                                                                                        M_pad * rowmin_ld);
                if (prune_batch) {
                    // 1. bounds: block minima of C_b + N1 - G_low - 2 sqrt(N1h Ch), the cross term over
                    //    the first K_low columns and the rest bounded by Cauchy-Schwarz, no values stored
                    k_const_term<<<B, 256, 0, st>>>(d->Fimg_cur.as<float2>(), Pw, tc.d_win.as<int>(),
                                                    tc.d_minv.as<float>(), tc.npix, d_ps, 2,
                                                    std::min(K_low / 2, tc.npix));
                    // one FP16 product per element unless the precision is BF16X3 (whose dropped
                    // terms would cost 2^-7 of the bound) or the full split is asked for
                    const char* env_fb = getenv("T2X_PRUNE_FULL_BOUNDS");
                    const bool h1 = prec != T2X_PREC_BF16X3 && !(env_fb && atoi(env_fb) != 0);
                    T2X_TRY(d->prune_c.ensure(sizeof(double) * 2 * B));
                    k_prune_c<<<(unsigned)((B + 255) / 256), 256, 0, st>>>(d_ps, B, h1 ? 1. / 512. : 0.,
                                                                           d->prune_c.as<double>());
                    T2X_HIP(hipGetLastError());
                    // the bounds go into the block minima themselves: a pruned tile keeps them, a
                    // kept tile's full GEMM overwrites them
                    EpiParams eb = ep;
                    eb.C = nullptr;
                    eb.N1h = d->prune_n1.as<float>();
                    eb.prune_c = d->prune_c.as<double>();
                    eb.pmin = nullptr;
                    if (h1)
                        T2X_TRY(launch_gemm_h1_bounds(d->Ah.ptr, d->As.as<float>(), Kc, M_pad, d->coarse_Bh.ptr,
                                                      d->coarse_Bs.as<float>(), Kc, K_low, (int)M_pad, (int)O_pad,
                                                      eb, st));
                    else
                        T2X_TRY(launch_gemm_x3(prec, EPI_DIFF2, d->Ah.ptr, d->Al.ptr, d->As.as<float>(), Kc, M_pad,
                                               d->coarse_Bh.ptr, d->coarse_Bl.ptr, d->coarse_Bs.as<float>(), Kc, K_low,
                                               (int)M_pad, (int)O_pad, eb, st,
                                               GemmChoice{cfg->gemm_tile, cfg->gemm_group}, nullptr, nullptr));
                    // 2. seeds: per particle the full diff2 of its most promising block (also the lazy
                    //    stores' starting minima)
                    T2X_TRY(d->prune_seed.ensure(sizeof(float) * B));
                    const int nblk = (int)((O_all + 63) / 64);
                    if (prec == T2X_PREC_BF16X3)
                        k_prune_seed<__bf16><<<B, 256, 0, st>>>(
                            d_rowmin.as<float>(), rowmin_ld, nblk, (int)T, (const __bf16*)d->Ah.ptr,
                            (const __bf16*)d->Al.ptr, d->As.as<float>(), (const __bf16*)d->coarse_Bh.ptr,
                            (const __bf16*)d->coarse_Bl.ptr, d->coarse_Bs.as<float>(), Kc, Kc, ep.Cb, ep.N1, O_pad,
                            ep.prior_o, O_all, d->prune_seed.as<float>(), d->pmin.as<unsigned>());
                    else
                        k_prune_seed<_Float16><<<B, 256, 0, st>>>(
                            d_rowmin.as<float>(), rowmin_ld, nblk, (int)T, (const _Float16*)d->Ah.ptr,
                            (const _Float16*)d->Al.ptr, d->As.as<float>(), (const _Float16*)d->coarse_Bh.ptr,
                            prec == T2X_PREC_FP16X2 ? nullptr : (const _Float16*)d->coarse_Bl.ptr,
                            d->coarse_Bs.as<float>(), Kc, Kc, ep.Cb, ep.N1, O_pad, ep.prior_o, O_all,
                            d->prune_seed.as<float>(), d->pmin.as<unsigned>());
                    T2X_HIP(hipGetLastError());
                    // 3. the full pass computes the tiles with a block not beyond the cut (below)
                    T2X_TRY(d->prune_cnt.ensure(sizeof(unsigned)));
                    T2X_TRY(zero_async(d->prune_cnt.ptr, sizeof(unsigned), st));
                    T2X_TRY(d->prune_list.ensure(sizeof(unsigned) * prune_ntiles));
                }
                if (prec == T2X_PREC_FP32) {
                    T2X_TRY(launch_gemm(EPI_DIFF2, d->A.as<float>(), Kc, d->coarse_B.as<float>(), Kc, Kc,
                                        (int)M_pad, (int)O_pad, ep, st));
                } else {
                    const GemmChoice gwant{cfg->gemm_tile, cfg->gemm_group};
                    GemmChoice gused{0, 0};
                    if (use_prop) {
                        if (n_fb_rows > 0) {
                            EpiParams ef = ep;
                            ef.M = n_fb_rows;
                            T2X_TRY(launch_gemm_x3(prec, EPI_DIFF2, d->Ah.ptr, d->Al.ptr, d->As.as<float>(), Kc,
                                                   M_pad, d->coarse_Bh.ptr, d->coarse_Bl.ptr, d->coarse_Bs.as<float>(),
                                                   Kc, Kc, (int)round_up(n_fb_rows, 256), (int)O_pad, ef, st, gwant,
                                                   d->prop_rows.as<int>(), &gused));
                        }
                        T2X_TRY(launch_gemm_x3_bins(prec, d->Ah.ptr, d->Al.ptr, d->As.as<float>(), Kc, d->coarse_Bh.ptr,
                                                    d->coarse_Bl.ptr, d->coarse_Bs.as<float>(), Kc, Kc,
                                                    d->prop_plist.as<int>(), d->prop_work.ptr, n_work, (int)T,
                                                    O_pad - 1, ep, st));
                    } else if (prune_batch) {
                        T2X_TRY(launch_gemm_pruned(prec, d->Ah.ptr, d->Al.ptr, d->As.as<float>(), Kc, M_pad,
                                                   d->coarse_Bh.ptr, d->coarse_Bl.ptr, d->coarse_Bs.as<float>(), Kc, Kc,
                                                   (int)M_pad, (int)O_pad, ep, d_rowmin.as<float>(),
                                                   d->prune_seed.as<float>(), d->prune_list.as<unsigned>(),
                                                   prune_ntiles, d->prune_cnt.as<unsigned>(), d->n_cu, st));
                        gused = GemmChoice{258, 0};   // group 0: the tile list
                    } else
                        T2X_TRY(launch_gemm_x3(prec, EPI_DIFF2, d->Ah.ptr, d->Al.ptr, d->As.as<float>(), Kc,
                                               M_pad, d->coarse_Bh.ptr, d->coarse_Bl.ptr, d->coarse_Bs.as<float>(), Kc,
                                               Kc, (int)M_pad, (int)O_pad, ep, st, gwant, nullptr, &gused));
                    if (gused.tile) {
                        d->stats.s.gemm_tile_used = gused.tile;
                        d->stats.s.gemm_group_used = gused.group;
                    }
                }
            }
            mark();   // 3: coarse gemm end
            // ---- coarse weights and significance ----
            wc.kind = SRC_COARSE;
            if (use_prop) {
                wc.bin_mask = d->prop_mask.as<unsigned>();
                wc.mask_words = (n_bins + 31) / 32;
                wc.bin_cols = bin_cols;
            }
            wc.d2 = d->diff2.as<float>();
            wc.ld = O_pad;
            wc.T = (int)T;
            wc.O_all = (int)O_all;
            wc.prior_o = d->prior_o.as<double>();
            wc.log_prior_o = d->log_prior_o.as<double>();
            wc.prior_t = d->prior_t.as<double>();
            wc.tK = tK;
            wc.tcols = (int)Norient;
            wc.rowmin = d_rowmin.as<float>();
            wc.rowmin_ld = rowmin_ld;
            wc.nover = nover;
            wc.over_trans = over_trans;
            wc.lp = lp;
            k_pass_min<<<B, 1024, 0, st>>>(wc, B, d_selc);
            SelParams sp{cfg->adaptive_fraction, (long long)cfg->maximum_significants, 1};
            // candidate path: sum pass, compaction of the weights that can be significant, exact
            // sorted walk per particle; the radix-select path handles batches that overflow
            T2X_TRY(d->cand_w.ensure(sizeof(double) * B * kCandCap));
            T2X_TRY(d->cand_e.ensure(sizeof(int) * B * kCandCap));
            T2X_TRY(d->cand_n.ensure(sizeof(int) * 2 * B));
            T2X_TRY(d->pairs_strided.ensure(sizeof(int2) * B * kCandCap));
            int* d_cand_count = d->cand_n.as<int>();
            int* d_overflow = d_cand_count + B;
            T2X_TRY(zero_async(d_cand_count, sizeof(int) * 2 * B, st));   // counts and overflow flags
            T2X_TRY(d->sum_part.ensure(sum_part_bytes() * B * T));
            // local searches take the radix-select path directly (the candidate filter works from the
            // global log prior)
            if (!local) {
                {
                    const dim3 grid((unsigned)T, (unsigned)B);
                    k_coarse_sum_rows<<<grid, 256, 0, st>>>(wc, B, d_selc, d->sum_part.as<SumPart>());
                    k_coarse_sum_final<<<B, 64, 0, st>>>(wc, B, d->sum_part.as<SumPart>(), d_selc);
                    k_coarse_candidates<<<grid, 256, 0, st>>>(wc, B, sp, d_selc, d->cand_w.as<double>(),
                                                              d->cand_e.as<int>(), d_cand_count);
                }
                // T2X_FORCE_RADIX_SELECT (tests): "1" sends every particle to the radix select,
                // "odd" every odd-numbered one
                const char* force_env = getenv("T2X_FORCE_RADIX_SELECT");
                const int force = !force_env ? 0 : (std::strcmp(force_env, "odd") == 0 ? 2 : 1);
                k_cand_select<<<B, 1024, 0, st>>>(wc, B, sp, d_selc, d->cand_w.as<double>(), d->cand_e.as<int>(),
                                                  d_cand_count, d->pairs_strided.as<int2>(),
                                                  d->sig_count.as<long long>(), d_overflow, force);
                T2X_HIP(hipGetLastError());
                cand_path = true;
                d_radix = d_overflow;
            }
            // the particles whose candidates overflowed (all of them without the candidate path)
            // take the exact radix select over their whole row, on the device: no host decision
            k_select<<<B, 1024, 0, st>>>(wc, sp, B, d_selc, d_radix);
            k_sig_count<<<B, 1024, 0, st>>>(wc, B, d_selc, d->sig_count.as<long long>(), d_radix);
            T2X_HIP(hipGetLastError());
            mark();   // 4: coarse weights done
        }
        return T2X_OK;
        };
        // ---- the batch's work lists, planned on the device ----
        // Offsets of every particle's significant pairs, orientations, fine items and fine tiles
        // come from k_batch_plan, and the fine pass and the store run as persistent work-queue
        // launches over them: nothing is read back between the coarse selection and the end of
        // the batch. The pair lists hold kCandCap pairs per particle (the candidate path's
        // bound; a local search's scored pairs bound its significant ones); radix-selected
        // particles can exceed that, and then the plan reports its totals, no kernel after it
        // does any work, and the batch is planned again with lists of that size.
        const int TW = (int)((T + 31) / 32);
        const int ncb = (int)((T * over_trans + FT_COLS - 1) / FT_COLS);
        DevBuf &d_pair_off = d->pair_off_d, &d_orient_off = d->orient_off_d, &d_item_off = d->item_off_d;
        T2X_TRY(d_pair_off.ensure(sizeof(long long) * (B + 1)));
        T2X_TRY(d_orient_off.ensure(sizeof(long long) * (B + 1)));
        T2X_TRY(d_item_off.ensure(sizeof(long long) * (B + 1)));
        T2X_TRY(d->fine_tile_off.ensure(sizeof(long long) * (B + 1)));
        T2X_TRY(d->plan.ensure(sizeof(BatchPlan)));
        BatchPlan* d_plan = d->plan.as<BatchPlan>();
        long long cap = local_pairs ? std::max<long long>(npairs_c, 1) : (long long)B * kCandCap;
        const char* cap_env = getenv("T2X_PLAN_CAP");   // tests: a small capacity forces a second plan
        if (cap_env && atoll(cap_env) > 0) cap = atoll(cap_env);
        else cap = std::max(cap, d->plan_cap);
        BatchPlan hplan{};
        std::vector<long long> pair_off(B + 1, 0);
        long long npairs = 0, nitems = 0;
        int ev_plan = ev;
        // per-particle results, read back at the end of the batch
        std::vector<SelResult> hself(B);
        std::vector<long long> hnsig(B);
        std::vector<double> hnorm(3 * B), hpower((size_t)B * xN);
        std::vector<PartScalars> hps2(B);
        std::vector<int> hflag(B);   // candidate path: particles sent to the radix select
        unsigned hprune = 0u;   // pruned coarse pass: the tiles its full GEMM computed
        auto post_part = [&]() -> int {
        T2X_TRY(d->sig_pairs.ensure(sizeof(int2) * cap));
        T2X_TRY(d->sig_orient.ensure(sizeof(int) * cap));
        T2X_TRY(d->orient_pairs_off.ensure(sizeof(long long) * cap));
        T2X_TRY(d->fine_diff2.ensure(sizeof(float) * cap * nover));
        k_batch_plan<<<1, 1024, 0, st>>>(d->sig_count.as<long long>(), d_selc, B, over_rot, FT_ROWS, ncb, cap, cap,
                                         d_pair_off.as<long long>(), d_orient_off.as<long long>(),
                                         d_item_off.as<long long>(), d->fine_tile_off.as<long long>(), d_plan);
        if (cand_path)
            k_cand_scatter<<<B, 256, 0, st>>>(B, d->pairs_strided.as<int2>(), d_pair_off.as<long long>(),
                                              d_orient_off.as<long long>(), d->sig_pairs.as<int2>(),
                                              d->sig_orient.as<int>(), d->orient_pairs_off.as<long long>(), d_radix);
        k_sig_write<<<B, 1024, 0, st>>>(wc, B, d_selc, d_pair_off.as<long long>(), d_orient_off.as<long long>(),
                                        d->sig_pairs.as<int2>(), d->sig_orient.as<int>(),
                                        d->orient_pairs_off.as<long long>(), d_radix);
        T2X_HIP(hipGetLastError());
        const bool want_dbg = dbg && dbg->n_debug > 0 && b0 < dbg->n_debug;
        if (want_dbg) {   // diagnostics read the plan and the coarse results on the host
            T2X_HIP(hipMemcpyAsync(&hplan, d_plan, sizeof(BatchPlan), hipMemcpyDeviceToHost, st));
            T2X_HIP(hipMemcpyAsync(pair_off.data(), d_pair_off.ptr, sizeof(long long) * (B + 1), hipMemcpyDeviceToHost, st));
            T2X_HIP(hipMemcpyAsync(hselc.data(), d_selc, sizeof(SelResult) * B, hipMemcpyDeviceToHost, st));
            T2X_HIP(hipStreamSynchronize(st));
            npairs = hplan.exceeded ? 0 : hplan.npairs;
            if (hplan.err_code) {
                set_error(std::string(hplan.err_code == T2X_EZEROSUM ? "zero sum of weights in the coarse pass"
                                                                      : "no significant coarse sample") +
                          " for particle " + std::to_string(b0 + hplan.err_particle));
                return hplan.err_code;
            }
        }
        // debug: coarse diff2 / significance in ihidden order (class, dir, psi, trans)
        if (dbg && dbg->n_debug > 0 && b0 < dbg->n_debug && local_pairs) {
            const int nd = std::min(B, dbg->n_debug - b0);
            const long Nh = O_all * T;
            std::vector<float> cd(std::max<long long>(npairs_c, 1));
            std::vector<int2> cp(std::max<long long>(npairs_c, 1));
            std::vector<long long> cpo(B + 1);
            T2X_HIP(hipStreamSynchronize(st));
            T2X_HIP(hipMemcpy(cd.data(), d->lc_d2.ptr, sizeof(float) * cd.size(), hipMemcpyDeviceToHost));
            T2X_HIP(hipMemcpy(cp.data(), d->lc_pairs.ptr, sizeof(int2) * cp.size(), hipMemcpyDeviceToHost));
            T2X_HIP(hipMemcpy(cpo.data(), d->lc_pair_off.ptr, sizeof(long long) * (B + 1), hipMemcpyDeviceToHost));
            for (int i = 0; i < nd; i++) {
                if (dbg->coarse_diff2) {
                    float* row = dbg->coarse_diff2 + (size_t)(b0 + i) * Nh;
                    std::fill(row, row + Nh, -999.f);
                    for (long long k = cpo[i]; k < cpo[i + 1]; k++) row[(long)cp[k].x * T + cp[k].y] = cd[k];
                }
                if (dbg->coarse_min_diff2) dbg->coarse_min_diff2[b0 + i] = hselc[i].min_d2;
                if (dbg->coarse_sum_weight) dbg->coarse_sum_weight[b0 + i] = hselc[i].sum_w;
                if (dbg->coarse_sig_weight) dbg->coarse_sig_weight[b0 + i] = hselc[i].sig_w;
            }
            if (dbg->coarse_significant) {
                std::vector<int2> hp(std::max<long long>(npairs, 1));
                T2X_HIP(hipMemcpy(hp.data(), d->sig_pairs.ptr, sizeof(int2) * hp.size(), hipMemcpyDeviceToHost));
                for (int i = 0; i < nd; i++) {
                    unsigned char* m = dbg->coarse_significant + (size_t)(b0 + i) * Nh;
                    std::memset(m, 0, Nh);
                    for (long long k = pair_off[i]; k < pair_off[i + 1]; k++) m[(long)hp[k].x * T + hp[k].y] = 1;
                }
            }
        } else if (dbg && dbg->n_debug > 0 && b0 < dbg->n_debug) {
            int nd = std::min(B, dbg->n_debug - b0);
            const long Nh = O_all * T;
            std::vector<float> row((size_t)T * O_pad);
            T2X_HIP(hipStreamSynchronize(st));
            for (int i = 0; i < nd; i++) {
                const int mw = (n_bins + 31) / 32;
                if (dbg->coarse_diff2)
                    T2X_HIP(hipMemcpy(row.data(), d->diff2.as<float>() + (size_t)i * T * O_pad,
                                      sizeof(float) * T * O_pad, hipMemcpyDeviceToHost));
                for (long o = 0; o < O_all && dbg->coarse_diff2; o++)
                    for (long t = 0; t < T; t++) {
                        float v = row[(size_t)t * O_pad + o];
                        if (!(prior_o[o] > 0.) || std::isinf(v)) v = -999.f;   // not scored / out of range
                        if (use_prop) {   // a bin outside the particle's scored set keeps stale values
                            const long bin = o / bin_cols;
                            if (!((prop_mask_host[(size_t)i * mw + bin / 32] >> (bin % 32)) & 1u)) v = -999.f;
                        }
                        if (dbg->coarse_diff2) dbg->coarse_diff2[(size_t)(b0 + i) * Nh + o * T + t] = v;
                    }
                if (dbg->coarse_min_diff2) dbg->coarse_min_diff2[b0 + i] = hselc[i].min_d2;
                if (dbg->coarse_sum_weight) dbg->coarse_sum_weight[b0 + i] = hselc[i].sum_w;
                if (dbg->coarse_sig_weight) dbg->coarse_sig_weight[b0 + i] = hselc[i].sig_w;
            }
            if (dbg->coarse_significant) {
                std::vector<int2> hp(std::max<long long>(npairs, 1));
                T2X_HIP(hipMemcpy(hp.data(), d->sig_pairs.ptr, sizeof(int2) * hp.size(), hipMemcpyDeviceToHost));
                for (int i = 0; i < nd; i++) {
                    unsigned char* m = dbg->coarse_significant + (size_t)(b0 + i) * Nh;
                    std::memset(m, 0, Nh);
                    for (long long k = pair_off[i]; k < pair_off[i + 1]; k++) m[(long)hp[k].x * T + hp[k].y] = 1;
                }
            }
        }
        // ---- fine pass ----
        T2X_TRY(d->fine_min.ensure(sizeof(unsigned int) * B));
        k_fill_u32<<<(B + 255) / 256, 256, 0, st>>>(d->fine_min.as<unsigned int>(), 0xffffffffu, B);
        FineGeom g{};
        g.Norient = (int)Norient;
        g.Npsi = (int)Npsi;
        g.n_subdir = S.n_subdir;
        g.n_subpsi = S.n_subpsi;
        g.over_rot = over_rot;
        g.over_trans = over_trans;
        g.T = (int)T;
        g.dir_sc = d_dirsc.as<double>();
        g.psi_sc = d_psisc.as<double>();
        g.pf = pf;
        g.ref2d = ref2d;
        mark();   // 5: fine start
        if (one_pass) {   // no second pass: the significant pairs keep their coarse diff2
            k_one_pass_d2<<<B, 256, 0, st>>>(d_pair_off.as<long long>(), d->sig_pairs.as<int2>(), B,
                                             d->diff2.as<float>(), O_pad, (int)T, d->fine_diff2.as<float>());
            T2X_HIP(hipGetLastError());
        } else {
            // tiles of 16 items x 64 columns per particle (items: its significant orientations x
            // oversampled rotations; columns: its significant coarse translations x oversampled
            // ones), planned by k_batch_plan, taken from a work queue by a persistent grid
            T2X_TRY(d->fine_tmask.ensure(sizeof(unsigned) * B * TW));
            k_fine_tmask<<<B, 256, 0, st>>>(d_pair_off.as<long long>(), d->sig_pairs.as<int2>(), B, TW,
                                            d->fine_tmask.as<unsigned>());
            T2X_HIP(hipGetLastError());
            FineTileParams fp{};
            fp.tile_off = d->fine_tile_off.as<long long>();
            fp.B = B;
            fp.ncb = ncb;
            fp.orient_off = d_orient_off.as<long long>();
            fp.orients = d->sig_orient.as<int>();
            fp.orient_pair_start = d->orient_pairs_off.as<long long>();
            fp.pair_off = d_pair_off.as<long long>();
            fp.pairs = d->sig_pairs.as<int2>();
            fp.tmask = d->fine_tmask.as<unsigned>();
            fp.TW = TW;
            fp.g = g;
            fp.vols = d_vols.as<const float2*>();
            fp.pads = d_pads.as<int>();
            fp.pix = tf.d_pix.as<int2>();
            fp.inproj = tf.d_inproj.as<unsigned char>();
            fp.minv = d_minv_fd.as<float>();
            fp.win = tf.d_win.as<int>();
            fp.npix = tf.npix;
            fp.Fimg = d->Fimg_cur.as<float2>();
            fp.Fctf = d->Fctf_cur.as<float>();
            fp.Pw = Pw;
            fp.abf = d->ab_fk.as<float>();
            fp.abh = d->ab_fh.ptr;
            fp.abl = d->ab_fl.ptr;
            fp.ldk = ldk_fine;
            fp.ps = d_ps;
            fp.fine_d2 = d->fine_diff2.as<float>();
            fp.fine_min = d->fine_min.as<unsigned int>();
            fp.queue = &d_plan->queue[0];
            // 4 workgroups per compute unit (the tile's LDS allows 4); those beyond the batch's
            // tiles leave at their first queue read
            const unsigned grid = (unsigned)(4 * d->n_cu);
            if (prec == T2X_PREC_FP32) k_fine_mfma<0><<<grid, 256, 0, st>>>(fp);
            else if (prec == T2X_PREC_BF16X3) k_fine_mfma<1><<<grid, 256, 0, st>>>(fp);
            else k_fine_mfma<2><<<grid, 256, 0, st>>>(fp);
            T2X_HIP(hipGetLastError());
        }
        mark();   // 6: fine diff2 end
        WeightSrc wf = wc;
        wf.kind = SRC_FINE;
        wf.nover = nover;
        wf.over_trans = over_trans;
        wf.d2 = d->fine_diff2.as<float>();
        wf.pair_off = d_pair_off.as<long long>();
        wf.pairs = d->sig_pairs.as<int2>();
        wf.fine_min = d->fine_min.as<unsigned int>();
        if (one_pass) {   // the coarse pass's minimum, sum and threshold over every sample
            k_copy_selresult<<<(B + 255) / 256, 256, 0, st>>>(d_selc, B, d_self);
        } else {
            k_pass_min<<<B, 1024, 0, st>>>(wf, B, d_self);
            SelParams spf{cfg->adaptive_fraction, (long long)cfg->maximum_significants, 0};
            k_select<<<B, 1024, 0, st>>>(wf, spf, B, d_self, nullptr);
        }
        T2X_TRY(d->fine_sel.ensure(sizeof(long long) * B));
        k_fine_argmax<<<B, 1024, 0, st>>>(wf, B, d_self, d->fine_sel.as<long long>());
        T2X_HIP(hipGetLastError());
        mark();   // 7: fine weights done
        if (dbg && dbg->n_debug > 0 && b0 < dbg->n_debug && dbg->fine_count) {
            int nd = std::min(B, dbg->n_debug - b0);
            std::vector<float> fd(std::max<long long>(npairs * nover, 1));
            std::vector<int2> hp(std::max<long long>(npairs, 1));
            T2X_HIP(hipStreamSynchronize(st));
            T2X_HIP(hipMemcpy(fd.data(), d->fine_diff2.ptr, sizeof(float) * fd.size(), hipMemcpyDeviceToHost));
            T2X_HIP(hipMemcpy(hp.data(), d->sig_pairs.ptr, sizeof(int2) * hp.size(), hipMemcpyDeviceToHost));
            for (int i = 0; i < nd; i++) {
                long cnt = 0;
                // CPU order: orientation, oversampled rotation, translation, oversampled translation
                long long k0 = pair_off[i], k1 = pair_off[i + 1];
                long long k = k0;
                while (k < k1) {
                    long long ke = k;
                    while (ke < k1 && hp[ke].x == hp[k].x) ke++;
                    for (int ir = 0; ir < over_rot; ir++)
                        for (long long kk = k; kk < ke; kk++)
                            for (int iot = 0; iot < over_trans; iot++) {
                                if (cnt < dbg->fine_capacity) {
                                    size_t o = (size_t)(b0 + i) * dbg->fine_capacity + cnt;
                                    if (dbg->fine_diff2) dbg->fine_diff2[o] = fd[kk * nover + ir * over_trans + iot];
                                    if (dbg->fine_hidden)
                                        dbg->fine_hidden[o] = ((long)hp[kk].x * T + hp[kk].y) * nover + ir * over_trans + iot;
                                }
                                cnt++;
                            }
                    k = ke;
                }
                dbg->fine_count[b0 + i] = cnt;
            }
        }
        // ---- weighted sums and backprojection ----
        T2X_TRY(d->best.ensure(sizeof(double) * 3 * B));   // per particle: wdiff2 sum, XA, AA
        T2X_TRY(zero_async(d->best.ptr, sizeof(double) * 3 * B, st));
        StoreParams P{};
        P.g = g;
        P.vols = d_vols.as<const float2*>();
        P.pads = d_pads.as<int>();
        P.bp_data = d_bpd.as<float2*>();
        P.bp_weight = d_bpw.as<float*>();
        P.bp_pads = d_bpads.as<int>();
        P.bp_rmax = d_bprmax.as<int>();
        P.prior_o = d->prior_o.as<double>();
        P.prior_t = d->prior_t.as<double>();
        P.tK = tK;
        P.cls_prior = tK ? d->cls_prior.as<double>() : nullptr;
        P.pix = tf.d_pix.as<int2>();
        P.inproj = tf.d_inproj.as<unsigned char>();
        P.minv_store = tf.d_minv.as<float>();
        P.win = tf.d_win.as<int>();
        P.ires = tf.d_ires.as<int>();
        P.npix = tf.npix;
        P.Fimg = d->Fimg_cur.as<float2>();
        P.Fimg_nm = d->Fimg_nm_cur.as<float2>();
        P.Fctf = d->Fctf_cur.as<float>();
        P.Pw = Pw;
        P.ab_fine = d->ab_fine.as<float2>();
        P.ftx = d_ftx.as<double>();
        P.fty = d_fty.as<double>();
        P.ps = d_ps;
        P.fine_res = d_self;
        P.fine_d2 = d->fine_diff2.as<float>();
        P.pixel_size = pix;
        P.nshell = xN;
        P.cs = cs;
        P.w_sigma2 = w_sigma2;
        P.w_sumw = w_sumw;
        P.w_pdf_class = w_pdfc;
        P.w_pdf_dir = w_pdfd;
        P.w_sigma2_offset = w_off;
        P.w_prior_offset = w_poff;
        P.w_norm = d->best.as<double>();
        P.w_xa = P.w_norm + B;
        P.w_aa = P.w_norm + 2 * B;
        P.dvp = scale_sums ? d->dvp_tab.as<double>() : nullptr;
        P.Nd = (int)Nd;
        P.lp = lp;
        if (bp_bricks) {
            P.bp_filt = d->bp_filt.as<float4>();
            P.bp_geom = d->bp_geom.as<double>();
            P.bp_val = d->bp_val.as<float4>();
            P.bp_wx = tf.px_w;
            P.bp_wy = tf.py_h;
            P.bp_wy0 = tf.py0;
            P.item_cap = bp_item_cap;
        }
        if (xN > 520) { set_error("box too large for the shell accumulator"); return T2X_EINVAL; }
        // one persistent launch: 8 workgroups per compute unit take the items from the queue
        k_store<<<(unsigned)(8 * d->n_cu), 256, 0, st>>>(d_item_off.as<long long>(), B, d_orient_off.as<long long>(),
                                                           d->sig_orient.as<int>(), d->orient_pairs_off.as<long long>(),
                                                           d_pair_off.as<long long>(), d->sig_pairs.as<int2>(), P,
                                                           &d_plan->queue[1]);
        T2X_HIP(hipGetLastError());
        if (bp_bricks && bp_nbricks > 0) {
            BpBrickParams Q{};
            Q.item_off = d_item_off.as<long long>();
            Q.B = B;
            Q.item_cap = bp_item_cap;
            Q.filt = d->bp_filt.as<float4>();
            Q.geom = d->bp_geom.as<double>();
            Q.val = d->bp_val.as<float4>();
            Q.py0 = tf.py0;
            Q.pxw = tf.px_w;
            Q.pyh = tf.py_h;
            Q.bp_data = d_bpd.as<float2*>();
            Q.bp_weight = d_bpw.as<float*>();
            Q.bp_pads = d_bpads.as<int>();
            Q.cls_off = d->bp_cls.as<int>();
            Q.cls_items = Q.cls_off + K + 1;
            const BpBrick* d_bricks = d->bp_units.as<BpBrick>();
            int* d_boff = d->bp_boff.as<int>();
            k_bp_bucket<<<1, 1024, 0, st>>>(Q.item_off, B, bp_item_cap, Q.filt, K, d->bp_cls.as<int>(),
                                            d->bp_cls.as<int>() + K + 1, d_bricks, bp_nbricks, bp_unit_items,
                                            d_boff);
            // workgroups that fit on a compute unit at once (3D: 52 KB of LDS, 2D: 16 KB) take the
            // units from the queue
            if (ref2d) k_bp_brick<0><<<(unsigned)(4 * d->n_cu), BP_THREADS, 0, st>>>(Q, d_bricks, bp_nbricks, d_boff);
            else k_bp_brick<1><<<(unsigned)(2 * d->n_cu), BP_THREADS, 0, st>>>(Q, d_bricks, bp_nbricks, d_boff);
            T2X_HIP(hipGetLastError());
        }
        mark();   // 8: store done
        return T2X_OK;
        };
        // graph mode: global searches without the proposal and without diagnostics; a batch
        // size is captured only after it ran once eagerly in this call (the GEMM autotuner
        // has timed its shapes), and a graph is replayed only while no device buffer has
        // been reallocated since its capture
        const bool gmode = cfg->launch_graph && !local && !use_prop && !(dbg && dbg->n_debug > 0);
        bool graph_used = false;
        for (int attempt = 0;; attempt++) {
        if (attempt == 0) {
            const bool replay = gmode && d->graph_exec && d->graph_call == call_id && d->graph_B == B &&
                                d->graph_prune == prune_batch && d->graph_gen == alloc_generation();
            if (replay || (gmode && d->eager_B.count(B))) {
                if (!replay) {
                    destroy_graph(d);
                    T2X_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                    capturing = true;
                    int rc = coarse_part();
                    ev_plan = ev;
                    if (rc == T2X_OK) rc = post_part();
                    hipGraph_t gr = nullptr;
                    const hipError_t ec = hipStreamEndCapture(st, &gr);
                    capturing = false;
                    if (rc != T2X_OK || ec != hipSuccess) {
                        if (gr) (void)hipGraphDestroy(gr);
                        if (rc == T2X_OK) set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
                        return rc != T2X_OK ? rc : T2X_EDEVICE;
                    }
                    hipGraphExec_t ex = nullptr;
                    if (hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess) {
                        (void)hipGraphDestroy(gr);
                        set_error("hipGraphInstantiate of the batch graph failed");
                        return T2X_EDEVICE;
                    }
                    d->graph = gr;
                    d->graph_exec = ex;
                    d->graph_B = B;
                    d->graph_prune = prune_batch;
                    d->graph_gen = alloc_generation();
                    d->graph_call = call_id;
                }
                (void)hipEventRecord(d->ev[14], st);
                T2X_HIP(hipGraphLaunch(d->graph_exec, st));
                (void)hipEventRecord(d->ev[15], st);
                graph_used = true;
            } else {
                T2X_TRY(coarse_part());
                ev_plan = ev;
                T2X_TRY(post_part());
                if (gmode) d->eager_B.insert(B);
            }
        } else {
            ev = ev_plan;
            T2X_TRY(post_part());
        }
        // ---- per-particle results: the batch's one readback ----
        {
            Readback rb;
            rb.add(&hplan, d_plan, sizeof(BatchPlan));
            rb.add(pair_off.data(), d_pair_off.ptr, sizeof(long long) * (B + 1));
            rb.add(hselc.data(), d_selc, sizeof(SelResult) * B);
            rb.add(hself.data(), d_self, sizeof(SelResult) * B);
            rb.add(hnsig.data(), d->fine_sel.ptr, sizeof(long long) * B);
            rb.add(hnorm.data(), d->best.ptr, sizeof(double) * 3 * B);
            rb.add(hpower.data(), d->power.ptr, sizeof(double) * B * xN);
            rb.add(hps2.data(), d_ps, sizeof(PartScalars) * B);
            if (cand_path) rb.add(hflag.data(), d_radix, sizeof(int) * B);
            if (prune_batch) rb.add(&hprune, d->prune_cnt.ptr, sizeof(unsigned));
            T2X_TRY(rb.run(d, st));
        }
        if (hplan.err_code) {
            set_error(std::string(hplan.err_code == T2X_EZEROSUM ? "zero sum of weights in the coarse pass"
                                                                  : "no significant coarse sample") +
                      " for particle " + std::to_string(b0 + hplan.err_particle));
            return hplan.err_code;
        }
        if (!hplan.exceeded) break;
        // lists too small for the radix-selected particles: grow them to the plan's totals and
        // run the batch's fine pass and store again (the first run did no work)
        if (attempt > 0) { set_error("batch plan: work lists exceeded twice"); return T2X_EDEVICE; }
        cap = std::max(hplan.npairs, hplan.norients);
        d->plan_cap = std::max(d->plan_cap, cap);
        }   // attempt
        npairs = hplan.npairs;
        nitems = hplan.nitems;
        if (bp_bricks) stt.bp_brick_items += (double)std::min(nitems, bp_item_cap);
        if (prune_batch) {
            const double skipped = (double)prune_ntiles - (double)hprune;
            stt.coarse_tiles_pruned += skipped;
            stt.coarse_tiles_checked += (double)prune_ntiles;
            stt.coarse_bound_flop += 4.0 * (double)T * (double)O_all * (double)(K_low / 2) * B;
            // a pass that prunes little costs more than the exhaustive one: off for the call's rest
            if (prune_ntiles > 0 && skipped < 0.4 * (double)prune_ntiles) prune_call = false;
        }
        if (!cand_path) {
            stt.n_select_fallbacks += 1;
        } else {   // did any particle of the batch take the radix select?
            for (int i = 0; i < B; i++)
                if (hflag[i]) { stt.n_select_fallbacks += 1; break; }
        }
        {
            float ms[8] = {0};
            // a graph's stages are timed together (its launch between events 14 and 15)
            const int ev_timed = graph_used ? 2 : ev;
            for (int e = 1; e < ev_timed; e++) (void)hipEventElapsedTime(&ms[e - 1], d->ev[e - 1], d->ev[e]);
            if (graph_used) {
                float mg = 0.f;
                (void)hipEventElapsedTime(&mg, d->ev[14], d->ev[15]);
                stt.ms_graph += mg;
                stt.n_graph_launches += 1;
            }
            stt.ms_preprocess += ms[0];
            stt.ms_coarse_diff2 += ms[1] + ms[2];
            stt.ms_coarse_gemm += ms[2];
            stt.ms_coarse_weights += ms[3];
            stt.ms_fine_diff2 += ms[4] + ms[5];
            stt.ms_fine_weights += ms[6];
            stt.ms_store += ms[7];
        }
        stt.n_coarse_diff2_launches += 1;
        stt.coarse_diff2_flop += local_pairs ? 4.0 * (double)npairs_c * tc.npix
                                             : 4.0 * (double)T * (double)O_all * (double)tc.npix * B;
        stt.coarse_diff2_bytes += 4.0 * ((double)M * Kc + (double)O_pad * Kc + (double)M * O_pad);
        stt.fine_samples += (double)npairs * nover;
        stt.fine_flop += 4.0 * (double)npairs * nover * tf.npix;
        stt.bp_orientations += (double)nitems;
        for (int i = 0; i < B; i++) {
            const int gi = b0 + i;
            const SelResult& r = hself[i];
            t2x_particle_out& o = out[gi];
            std::memset(&o, 0, sizeof(o));
            if (!(r.sum_w > 0.) || std::isnan(r.sum_w)) {
                set_error("zero sum of weights in the fine pass for particle " + std::to_string(gi));
                return T2X_EZEROSUM;
            }
            // decode the best key: (((n*over_rot + ir)*T + t)*over_trans + iot)
            long long kkey = r.best_key;
            const int iot = (int)(kkey % over_trans); kkey /= over_trans;
            const long t = (long)(kkey % T); kkey /= T;
            const int ir = (int)(kkey % over_rot); kkey /= over_rot;
            const long nn = (long)kkey;
            const int c = (int)(nn / Norient);
            const long oo = nn % Norient, dd = oo / Npsi, pp = oo % Npsi;
            o.rot = S.sub_rot[dd * S.n_subdir + ir / S.n_subpsi];
            o.tilt = S.sub_tilt[dd * S.n_subdir + ir / S.n_subpsi];
            o.psi = S.sub_psi[pp * S.n_subpsi + ir % S.n_subpsi];
            o.xoff = hps[i].old_x + S.fine_tx_px[t * over_trans + iot];
            o.yoff = hps[i].old_y + S.fine_ty_px[t * over_trans + iot];
            o.iclass = c + 1;
            o.pmax = r.max_w / r.sum_w;
            o.best_hidden = ((nn * T + t) * over_rot + ir) * over_trans + iot;
            o.sum_weight = r.sum_w;
            o.min_diff2 = r.min_d2;
            o.significant_weight = r.sig_w;
            o.nr_significant_coarse = hselc[i].nsig;
            o.nr_fine_samples = (pair_off[i + 1] - pair_off[i]) * nover;
            o.nr_fine_significant = hnsig[i];
            const int og = pin[gi].optics_group;
            double wnorm = hnorm[i];
            for (int ires = cs / 2 + 1; ires < xN; ires++) {
                h_power_sum[(size_t)og * xN + ires] += hpower[(size_t)i * xN + ires];
                wnorm += hpower[(size_t)i * xN + ires];
            }
            if (cfg->do_scale_correction) {
                // XA / scale and AA / scale^2 with the group's scale (ml_optimiser.cpp:8855-8862)
                const double sc = hps[i].scale;
                h_xa[pin[gi].group_id] += hnorm[B + i] / sc;
                h_aa[pin[gi].group_id] += hnorm[2 * B + i] / (sc * sc);
            }
            double normcorr = pin[gi].norm_correction;
            if (cfg->do_norm_correction) {
                double old = pin[gi].norm_correction / cfg->avg_norm_correction;
                normcorr = old * std::sqrt(wnorm * 2.);
                avg_norm += normcorr;
            }
            o.norm_correction = normcorr;
            o.dLL = std::log(r.sum_w) - r.min_d2 - logsigma2[og];
            LL += o.dLL;
            Pmax += o.pmax;
        }
        stt.n_particles += B;
        for (int i = 0; i < B; i++) stt.sig_coarse += (double)hselc[i].nsig;
    }
    stt.n_calls += 1;
    // ---- weighted sums back to the caller ----
    std::vector<double> hw(nws);
    {
        Readback rb;
        rb.add(hw.data(), d->wsum_dev.ptr, sizeof(double) * nws);
        T2X_TRY(rb.run(d, st));
    }
    if (wsum) {
        const size_t o_sumw = (size_t)n_og * xN, o_pdfc = o_sumw + n_og, o_pdfd = o_pdfc + K,
                     o_off = o_pdfd + (size_t)K * Nd, o_poff = o_off + 1;
        if (wsum->sigma2_noise)
            for (size_t i = 0; i < (size_t)n_og * xN; i++) wsum->sigma2_noise[i] += hw[i] + h_power_sum[i];
        if (wsum->sumw_group)
            for (int g = 0; g < n_og; g++) wsum->sumw_group[g] += hw[o_sumw + g];
        if (wsum->pdf_class)
            for (int c = 0; c < K; c++) wsum->pdf_class[c] += hw[o_pdfc + c];
        if (wsum->pdf_direction)
            for (long i = 0; i < K * Nd; i++) wsum->pdf_direction[i] += hw[o_pdfd + i];
        if (wsum->sigma2_offset) wsum->sigma2_offset[0] += hw[o_off];
        if (wsum->prior_offset_class && ref2d)
            for (int c = 0; c < 2 * K; c++) wsum->prior_offset_class[c] += hw[o_poff + c];
        if (cfg->do_scale_correction) {
            if (wsum->wsum_signal_product)
                for (int g = 0; g < n_grp; g++) wsum->wsum_signal_product[g] += h_xa[g];
            if (wsum->wsum_reference_power)
                for (int g = 0; g < n_grp; g++) wsum->wsum_reference_power[g] += h_aa[g];
        }
        if (wsum->avg_norm_correction && cfg->do_norm_correction) wsum->avg_norm_correction[0] += avg_norm;
        if (wsum->LL) wsum->LL[0] += LL;
        if (wsum->ave_Pmax) wsum->ave_Pmax[0] += Pmax;
    }
    return T2X_OK;
}
