/*
 * recon_oracle.cpp — FP64 CPU restatement of BackProjector::reconstruct (default skip_gridding
 * branch) and getFSC. TEST INFRASTRUCTURE ONLY: the checker for t2x_reconstruct / t2x_fsc.
 *
 * Follows src/backprojector.cpp:1344-1770 (reconstruct), :2486-2600 (windowToOridimRealSpace),
 * src/projector.h:242-253 (decenter), src/projector.cpp Projector::griddingCorrect,
 * src/mask.cpp:43-97 (softMaskOutsideMap), src/fftw.h:98-119 (FFTW-layout loops), :389-402
 * (CenterFFTbySign), :737-787 (windowFourierTransform), src/fftw.cpp:481-523 (getFSC),
 * src/macros.h:396-405 (Xmipp origins). Transforms are separable mixed-radix DFTs on the full
 * complex grid (any size), normalised like RELION's FourierTransformer (forward / size).
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

typedef std::complex<double> cplx;
static const double kPi = 3.14159265358979323846;

/* mixed-radix recursive DFT of length n (radix = smallest prime factor) */
static void dft1(const cplx* in, long stride, long n, cplx* out, int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    long p = 2;
    while (p * p <= n && n % p) p++;
    if (n % p) p = n;
    const long m = n / p;
    std::vector<cplx> sub(n);
    for (long r = 0; r < p; r++) dft1(in + r * stride, stride * p, m, &sub[r * m], sign);
    for (long k = 0; k < m; k++)
        for (long q = 0; q < p; q++) {
            const long kk = k + m * q;
            cplx acc = 0;
            for (long r = 0; r < p; r++) {
                const double ang = sign * 2.0 * kPi * (double)((r * kk) % n) / (double)n;
                acc += sub[r * m + k] * cplx(std::cos(ang), std::sin(ang));
            }
            out[kk] = acc;
        }
}

/* in-place 3D DFT of a full complex n^3 array [z][y][x] */
static void dft3(std::vector<cplx>& a, long n, int sign)
{
    std::vector<cplx> line(n), res(n);
    for (int axis = 0; axis < 3; axis++) {
        const long st = axis == 0 ? 1 : (axis == 1 ? n : n * n);
        for (long u = 0; u < n; u++)
            for (long v = 0; v < n; v++) {
                long base;
                if (axis == 0) base = (u * n + v) * n;
                else if (axis == 1) base = u * n * n + v;
                else base = u * n + v;
                for (long t = 0; t < n; t++) line[t] = a[base + t * st];
                dft1(line.data(), 1, n, res.data(), sign);
                for (long t = 0; t < n; t++) a[base + t * st] = res[t];
            }
    }
}

static inline long wrap(long k, long n) { return k < 0 ? k + n : k; }

/* FFTW-layout logical index of direct index k along a dimension of size zdim with xsize half */
static inline long fftw_logical(long k, long xsize, long zdim) { return k < xsize ? k : k - zdim; }

/* softMaskOutsideMap(vol) with radius = XSIZE/2, cosine_width = 3, average background */
static void soft_mask_outside(std::vector<double>& vol, long n)
{
    const double radius = (double)n / 2., cw = 3., radius_p = radius + cw;
    const long o = n / 2;   /* FIRST_XMIPP_INDEX(n) = -n/2 for even n */
    double sum = 0., sum_bg = 0.;
    for (long z = 0; z < n; z++)
        for (long y = 0; y < n; y++)
            for (long x = 0; x < n; x++) {
                const long k = z - o, i = y - o, j = x - o;
                const double r = std::sqrt((double)(k * k + i * i + j * j));
                const double v = vol[(z * n + y) * n + x];
                if (r < radius) continue;
                if (r > radius_p) { sum += 1.; sum_bg += v; }
                else {
                    const double rc = 0.5 + 0.5 * std::cos(kPi * (radius_p - r) / cw);
                    sum += rc;
                    sum_bg += rc * v;
                }
            }
    sum_bg /= sum;
    for (long z = 0; z < n; z++)
        for (long y = 0; y < n; y++)
            for (long x = 0; x < n; x++) {
                const long k = z - o, i = y - o, j = x - o;
                const double r = std::sqrt((double)(k * k + i * i + j * j));
                double& v = vol[(z * n + y) * n + x];
                if (r < radius) continue;
                if (r > radius_p) v = sum_bg;
                else {
                    const double rc = 0.5 + 0.5 * std::cos(kPi * (radius_p - r) / cw);
                    v = (1 - rc) * v + rc * sum_bg;
                }
            }
}

extern "C" int orc_reconstruct(const t2x_recon_params* p, const double* data_re, const double* data_im,
                               const double* weight, const double* tau2, double* vol_out)
{
    if (!p || p->ref_dim != 3 || !p->skip_gridding) return -1;
    const long pad = p->pad_size, xh = pad / 2 + 1, ori = p->ori_size;
    const double pf = p->padding_factor;
    const long rr = std::lround(p->r_max * pf);                  /* ROUND(r_max * padding_factor) */
    const long max_r2 = rr * rr;
    const double oversampling_correction = pf * pf * pf;
    const long c0 = pad / 2;                                     /* -FIRST_XMIPP_INDEX(pad)      */
    const size_t nv = (size_t)pad * pad * xh;

    /* decenter: projector layout (logical kp, ip, jp) -> FFTW layout of the pad^3 transform */
    std::vector<double> Fw(nv, 0.);
    std::vector<cplx> Fc(nv, 0.);
    for (long k = 0; k < pad; k++)
        for (long i = 0; i < pad; i++)
            for (long j = 0; j < xh; j++) {
                const long kp = fftw_logical(k, xh, pad), ip = fftw_logical(i, xh, pad), jp = j;
                if (kp * kp + ip * ip + jp * jp <= max_r2) {
                    const size_t src = ((size_t)(kp + c0) * pad + (ip + c0)) * xh + jp;
                    const size_t dst = ((size_t)k * pad + i) * xh + j;
                    Fw[dst] = weight[src];
                    Fc[dst] = cplx(data_re[src], data_im[src]);
                }
            }
    /* MAP term */
    if (p->do_map) {
        for (long k = 0; k < pad; k++)
            for (long i = 0; i < pad; i++)
                for (long j = 0; j < xh; j++) {
                    const long kp = fftw_logical(k, xh, pad), ip = fftw_logical(i, xh, pad), jp = j;
                    const long r2 = kp * kp + ip * ip + jp * jp;
                    if (r2 >= max_r2) continue;
                    const long ires = std::lround(std::sqrt((double)r2) / pf);
                    const size_t e = ((size_t)k * pad + i) * xh + j;
                    double invw = Fw[e], invtau2;
                    if (tau2[ires] > 0.) invtau2 = 1. / (oversampling_correction * p->tau2_fudge * tau2[ires]);
                    else if (tau2[ires] < 1e-20) invtau2 = (invw > 1e-20) ? 1. / (0.001 * invw) : 0.;
                    else return -2;
                    if (ires >= p->minres_map) Fw[e] = invw + invtau2;
                }
    }
    /* radially averaged weight floor (1/1000 of the shell average) and division */
    const long round_max_r2 = std::lround(p->r_max * pf * p->r_max * pf);
    std::vector<double> radavg(p->r_max, 0.), counter(p->r_max, 0.);
    for (long k = 0; k < pad; k++)
        for (long i = 0; i < pad; i++)
            for (long j = 0; j < xh; j++) {
                const long kp = fftw_logical(k, xh, pad), ip = fftw_logical(i, xh, pad), jp = j;
                const long r2 = kp * kp + ip * ip + jp * jp;
                if (r2 < round_max_r2) {
                    const long ires = (long)std::floor(std::sqrt((double)r2) / pf);
                    if (ires >= p->r_max) return -3;
                    radavg[ires] += Fw[((size_t)k * pad + i) * xh + j];
                    counter[ires] += 1.;
                }
            }
    for (long s = 0; s < p->r_max; s++) {
        if (counter[s] > 0. || radavg[s] > 0.) radavg[s] /= 1000. * counter[s];
        else return -4;
    }
    for (long k = 0; k < pad; k++)
        for (long i = 0; i < pad; i++)
            for (long j = 0; j < xh; j++) {
                const long kp = fftw_logical(k, xh, pad), ip = fftw_logical(i, xh, pad), jp = j;
                const long r2 = kp * kp + ip * ip + jp * jp;
                const long ires = (long)std::floor(std::sqrt((double)r2) / pf);
                const size_t e = ((size_t)k * pad + i) * xh + j;
                const double w = std::max(Fw[e], radavg[(ires < p->r_max) ? ires : (p->r_max - 1)]);
                if (w != 0.) Fc[e] /= w;
            }

    /* windowToOridimRealSpace: Fourier window to P = padoridim */
    long P = std::lround(pf * ori);
    P += P % 2;
    const long Ph = P / 2 + 1;
    std::vector<cplx> W((size_t)P * P * Ph, 0.);
    if (Ph > xh) {
        const long mr2 = (xh - 1) * (xh - 1);
        for (long k = 0; k < pad; k++)
            for (long i = 0; i < pad; i++)
                for (long j = 0; j < xh; j++) {
                    const long kp = fftw_logical(k, xh, pad), ip = fftw_logical(i, xh, pad), jp = j;
                    if (kp * kp + ip * ip + jp * jp <= mr2)
                        W[((size_t)wrap(kp, P) * P + wrap(ip, P)) * Ph + jp] = Fc[((size_t)k * pad + i) * xh + j];
                }
    } else {
        for (long k = 0; k < P; k++)
            for (long i = 0; i < P; i++)
                for (long j = 0; j < Ph; j++) {
                    const long kp = fftw_logical(k, Ph, P), ip = fftw_logical(i, Ph, P), jp = j;
                    W[((size_t)k * P + i) * Ph + j] = Fc[((size_t)wrap(kp, pad) * pad + wrap(ip, pad)) * xh + jp];
                }
    }
    /* CenterFFTbySign */
    for (long k = 0; k < P; k++)
        for (long i = 0; i < P; i++)
            for (long j = 0; j < Ph; j++)
                if (((k ^ i ^ j) & 1) != 0) W[((size_t)k * P + i) * Ph + j] *= -1.;
    /* half-complex inverse transform: full Hermitian grid from the half (x = 0 and x = P/2
     * planes enter through the real part of their transform, i.e. Hermitian-symmetrised) */
    std::vector<cplx> full((size_t)P * P * P);
    for (long k = 0; k < P; k++)
        for (long i = 0; i < P; i++)
            for (long j = 0; j < P; j++) {
                cplx v;
                if (j < Ph && j > 0 && j < P / 2) v = W[((size_t)k * P + i) * Ph + j];
                else if (j == 0 || j == P / 2) {
                    const cplx a = W[((size_t)k * P + i) * Ph + j];
                    const cplx b = std::conj(W[((size_t)((P - k) % P) * P + (P - i) % P) * Ph + j]);
                    v = 0.5 * (a + b);
                } else {
                    v = std::conj(W[((size_t)((P - k) % P) * P + (P - i) % P) * Ph + (P - j)]);
                }
                full[((size_t)k * P + i) * P + j] = v;
            }
    dft3(full, P, +1);
    /* real-space window to ori (Xmipp origins: P/2 and ori/2), FFT normalisation */
    const double normfft = (p->data_dim == 3) ? pf * pf * pf : pf * pf * pf * (double)ori;
    std::vector<double> vol((size_t)ori * ori * ori);
    const long oP = P / 2, oO = ori / 2;   /* -FIRST_XMIPP_INDEX (both even) */
    for (long z = 0; z < ori; z++)
        for (long y = 0; y < ori; y++)
            for (long x = 0; x < ori; x++) {
                const long zz = z - oO + oP, yy = y - oO + oP, xx = x - oO + oP;
                vol[((size_t)z * ori + y) * ori + x] = full[((size_t)zz * P + yy) * P + xx].real() / normfft;
            }
    soft_mask_outside(vol, ori);
    /* griddingCorrect (trilinear): divide by sinc^2(r / (ori * pf)) */
    for (long z = 0; z < ori; z++)
        for (long y = 0; y < ori; y++)
            for (long x = 0; x < ori; x++) {
                const long k = z - oO, i = y - oO, j = x - oO;
                const double r = std::sqrt((double)(k * k + i * i + j * j));
                if (r > 0.) {
                    const double rval = r / (ori * pf);
                    const double sinc = std::sin(kPi * rval) / (kPi * rval);
                    vol[((size_t)z * ori + y) * ori + x] /= sinc * sinc;
                }
            }
    for (size_t e = 0; e < vol.size(); e++) vol_out[e] = vol[e];
    return 0;
}

/* getFSC of two real n^3 maps: forward transforms (1/size), shells ROUND(|k|) < n/2+1 */
extern "C" int orc_fsc(int n, const double* m1, const double* m2, double* fsc)
{
    const long N = n, xh = N / 2 + 1;
    std::vector<cplx> a((size_t)N * N * N), b((size_t)N * N * N);
    for (size_t e = 0; e < a.size(); e++) { a[e] = m1[e]; b[e] = m2[e]; }
    dft3(a, N, -1);
    dft3(b, N, -1);
    std::vector<double> num(xh, 0.), d1(xh, 0.), d2(xh, 0.);
    const double s = 1.0 / ((double)N * N * N);
    for (long k = 0; k < N; k++)
        for (long i = 0; i < N; i++)
            for (long j = 0; j < xh; j++) {
                const long kp = fftw_logical(k, xh, N), ip = fftw_logical(i, xh, N), jp = j;
                const long idx = std::lround(std::sqrt((double)(kp * kp + ip * ip + jp * jp)));
                if (idx >= xh) continue;
                const cplx z1 = a[((size_t)k * N + i) * N + j] * s, z2 = b[((size_t)k * N + i) * N + j] * s;
                num[idx] += (std::conj(z1) * z2).real();
                d1[idx] += std::norm(z1);
                d2[idx] += std::norm(z2);
            }
    for (long i = 0; i < xh; i++) fsc[i] = num[i] / std::sqrt(d1[i] * d2[i]);
    return 0;
}
