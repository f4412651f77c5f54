/*
 * estep_oracle.cpp — FP64 CPU restatement of relion_refine's CPU E-step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for libt2x). Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library; the product never does.
 *
 * Each function cites the reference code it restates (paths relative to the reference tree).
 * All arithmetic is double precision, as relion_refine is built with DoublePrec_CPU
 * (CMakeLists.txt:102).
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cplx;
static const double PI = 3.14159265358979323846;

/* RELION's rounding macros (macros.h:195-224) */
static inline int R_ROUND(double x) { return (x > 0) ? (int)(x + 0.5) : (int)(x - 0.5); }
static inline int R_CEIL(double x) { return (x == (int)x) ? (int)x : ((x > 0) ? (int)(x + 1) : (int)x); }
static inline int R_FLOOR(double x) { return (x == (int)x) ? (int)x : ((x > 0) ? (int)x : (int)(x - 1)); }

extern "C" int orc_abi_version(void) { return T2X_ABI_VERSION; }

/* =====================================================================================
 * HEALPix NEST scheme. Restates Healpix_Base::nest2xyf / xyf2nest / pix2ang_z_phi
 * (src/Healpix_2.15a/healpix_base.cc) from the published HEALPix algorithm
 * (Gorski et al. 2005, ApJ 622:759): bit (de)interleaving of the face coordinates and the
 * ring index jr / pixel-in-ring jp of the base-resolution face tables.
 * ===================================================================================== */
static const int kJrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
static const int kJpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

static inline long hp_compact(long v) /* keep the even bits of v, packed */
{
    long r = 0;
    for (int b = 0; b < 31; b++)
        if (v & (1L << (2 * b))) r |= (1L << b);
    return r;
}
static inline long hp_spread(long v) /* inverse of hp_compact */
{
    long r = 0;
    for (int b = 0; b < 31; b++)
        if (v & (1L << b)) r |= (1L << (2 * b));
    return r;
}

extern "C" void orc_healpix_nest2xyf(int order, long pix, int* ix, int* iy, int* face)
{
    long npface = 1L << (2 * order);
    *face = (int)(pix >> (2 * order));
    long p = pix & (npface - 1);
    *ix = (int)hp_compact(p);
    *iy = (int)hp_compact(p >> 1);
}

extern "C" long orc_healpix_xyf2nest(int order, int ix, int iy, int face)
{
    return ((long)face << (2 * order)) + hp_spread(ix) + (hp_spread(iy) << 1);
}

extern "C" void orc_healpix_pix2ang_z_phi(int order, long pix, double* z, double* phi)
{
    long nside = 1L << order;
    long npix = 12 * nside * nside;
    long nl4 = 4 * nside;
    double fact2 = 4.0 / (double)npix;
    double fact1 = (double)(nside << 1) * fact2;
    int ix, iy, face;
    orc_healpix_nest2xyf(order, pix, &ix, &iy, &face);
    long jr = ((long)kJrll[face] << order) - ix - iy - 1;
    long nr, kshift;
    if (jr < nside) {
        nr = jr;
        *z = 1.0 - (double)(nr * nr) * fact2;
        kshift = 0;
    } else if (jr > 3 * nside) {
        nr = nl4 - jr;
        *z = (double)(nr * nr) * fact2 - 1.0;
        kshift = 0;
    } else {
        nr = nside;
        *z = (double)(2 * nside - jr) * fact1;
        kshift = (jr - nside) & 1;
    }
    long jp = (kJpll[face] * nr + ix - iy + 1 + kshift) / 2;
    if (jp > nl4) jp -= nl4;
    if (jp < 1) jp += nl4;
    *phi = ((double)jp - (double)(kshift + 1) * 0.5) * ((PI / 2.0) / (double)nr);
}

/* =====================================================================================
 * Sampling grid: HealpixSampling (src/healpix_sampling.cpp)
 * ===================================================================================== */
struct Sampling {
    int is_3D = 1;
    int order = 0;
    double psi_step = 0, offset_step = 0, offset_range = 0;
    std::vector<double> rot, tilt, psi, tx, ty; /* translations in A */
    std::vector<long> ipix;

    /* HealpixSampling::checkDirection (healpix_sampling.cpp:1617-1640) */
    static void checkDirection(double& r, double& t)
    {
        if (t < 0.) { t = -t; r += 180.; }
        for (;;) {
            if (r > 180.) r -= 360.;
            else if (r < -180.) r += 360.;
            else break;
        }
    }

    void setup(const t2x_estep_config* cfg)
    {
        is_3D = (cfg->ref_dim == 3);
        order = cfg->healpix_order;
        rot.clear(); tilt.clear(); psi.clear(); tx.clear(); ty.clear(); ipix.clear();
        /* setOrientations (healpix_sampling.cpp:534-608); C1: no symmetry-equivalent points */
        if (is_3D) {
            long npix = 12L << (2 * order);
            for (long p = 0; p < npix; p++) {
                double zz, phi;
                orc_healpix_pix2ang_z_phi(order, p, &zz, &phi);
                double r = phi * 180.0 / PI, t = std::acos(zz) * 180.0 / PI;
                checkDirection(r, t); /* getDirectionFromHealPix (:1642-1649) */
                rot.push_back(r); tilt.push_back(t); ipix.push_back(p);
            }
        } else {
            rot.push_back(0.); tilt.push_back(0.); ipix.push_back(-1);
        }
        int nr_psi = R_CEIL(360. / cfg->psi_step);
        psi_step = 360. / (double)nr_psi;
        for (int i = 0; i < nr_psi; i++) psi.push_back(i * psi_step);
        /* setTranslations (healpix_sampling.cpp:290-440), 2D, not helical */
        offset_step = cfg->offset_step;
        offset_range = cfg->offset_range;
        int maxp = R_CEIL(offset_range / offset_step);
        for (long ix = -maxp; ix <= maxp; ix++) {
            double xoff = ix * offset_step;
            for (long iy = -maxp; iy <= maxp; iy++) {
                double yoff = iy * offset_step;
                double max2 = xoff * xoff + yoff * yoff;
                if (max2 < (offset_range * offset_range) + 0.001) {
                    tx.push_back(xoff);
                    ty.push_back(yoff);
                }
            }
        }
    }
    long nr_dir() const { return (long)rot.size(); }
    long nr_psi() const { return (long)psi.size(); }
    long nr_trans() const { return (long)tx.size(); }
    /* oversamplingFactorOrientations / Translations (healpix_sampling.cpp:1720-1735) */
    int over_rot(int os) const { return is_3D ? (1 << (3 * os)) : (1 << os); }
    int over_trans(int os) const { return 1 << (2 * os); }

    /* pushbackOversampledPsiAngles (healpix_sampling.cpp:2023-2050) */
    void push_psi(long ipsi, int os, double r, double t, std::vector<double>& orot,
                  std::vector<double>& otilt, std::vector<double>& opsi) const
    {
        if (os == 0) {
            orot.push_back(r); otilt.push_back(t); opsi.push_back(psi[ipsi]);
            return;
        }
        int n = 1 << os;
        for (int k = 0; k < n; k++) {
            double overpsi = psi[ipsi] - 0.5 * psi_step + (0.5 + k) * psi_step / n;
            orot.push_back(r); otilt.push_back(t);
            if (!is_3D && overpsi > 180.) overpsi -= 360.;
            opsi.push_back(overpsi);
        }
    }
    /* getOrientations (healpix_sampling.cpp:1914-2021), no priors, no perturbation */
    void orientations(long idir, long ipsi, int os, std::vector<double>& orot,
                      std::vector<double>& otilt, std::vector<double>& opsi) const
    {
        orot.clear(); otilt.clear(); opsi.clear();
        if (os == 0) {
            orot.push_back(rot[idir]); otilt.push_back(tilt[idir]); opsi.push_back(psi[ipsi]);
        } else if (!is_3D) {
            push_psi(ipsi, os, 0., 0., orot, otilt, opsi);
        } else {
            int over_order = order + os;
            int fact = 1 << os;
            int x, y, face;
            orc_healpix_nest2xyf(order, ipix[idir], &x, &y, &face);
            for (int j = fact * y; j < fact * (y + 1); ++j)
                for (int i = fact * x; i < fact * (x + 1); ++i) {
                    long overpix = orc_healpix_xyf2nest(over_order, i, j, face);
                    double zz, phi;
                    orc_healpix_pix2ang_z_phi(over_order, overpix, &zz, &phi);
                    double r = phi * 180.0 / PI, t = std::acos(zz) * 180.0 / PI;
                    checkDirection(r, t);
                    push_psi(ipsi, os, r, t, orot, otilt, opsi);
                }
        }
    }
    /* getTranslationsInPixel (healpix_sampling.cpp:1800-1912), 2D, no helix, no perturbation */
    void translations_px(long itrans, int os, double pix, std::vector<double>& x,
                         std::vector<double>& y) const
    {
        x.clear(); y.clear();
        if (os == 0) {
            x.push_back(tx[itrans] / pix);
            y.push_back(ty[itrans] / pix);
            return;
        }
        int n = 1 << os;
        for (int ox = 0; ox < n; ox++) {
            double over_xoff = tx[itrans] - 0.5 * offset_step + (0.5 + ox) * offset_step / n;
            for (int oy = 0; oy < n; oy++) {
                double over_yoff = ty[itrans] - 0.5 * offset_step + (0.5 + oy) * offset_step / n;
                x.push_back(over_xoff / pix);
                y.push_back(over_yoff / pix);
            }
        }
    }
};

static int check_config(const t2x_estep_config* c)
{
    if (!c) return T2X_EINVAL;
    if (c->ori_size < 4 || (c->ori_size & 1)) return T2X_EINVAL;
    if (c->current_size < 2 || (c->current_size & 1) || c->current_size > c->ori_size) return T2X_EINVAL;
    if (c->coarse_size < 2 || (c->coarse_size & 1) || c->coarse_size > c->current_size) return T2X_EINVAL;
    if (c->ref_dim != 3 && c->ref_dim != 2) return T2X_EINVAL;
    if (c->nr_classes < 1) return T2X_EINVAL;
    if (c->adaptive_oversampling < 0 || c->adaptive_oversampling > 2) return T2X_EINVAL;
    if (c->offset_step <= 0 || c->offset_range < 0 || c->psi_step <= 0) return T2X_EINVAL;
    if (c->nr_groups < 0 || c->nr_optics_groups < 0) return T2X_EINVAL;
    if (c->orientational_prior_mode == T2X_PRIOR_ROTTILT_PSI) {
        /* local searches with one Gaussian on the angle between directions (sigma_rot and
         * sigma_tilt both set) or none (both 0); the rot-only / tilt-only branches are not restated */
        if (c->sigma_rot < 0 || c->sigma_tilt < 0 || c->sigma_psi < 0) return T2X_EINVAL;
        if ((c->sigma_rot > 0.) != (c->sigma_tilt > 0.)) return T2X_EINVAL;
    } else if (c->orientational_prior_mode != T2X_NOPRIOR) {
        return T2X_EINVAL;
    }
    return T2X_OK;
}

/* gaussian1D (funcs.cpp:356-360), evaluated left to right */
static double gaussian1D(double x, double sigma, double mu)
{
    x -= mu;
    return 1 / std::sqrt(2 * PI * sigma * sigma) * std::exp(-0.5 * ((x / sigma) * (x / sigma)));
}

/* Euler_angles2direction (euler.cpp:94-114) */
static void angles2direction(double rot, double tilt, double v[3])
{
    double a = rot * PI / 180, b = tilt * PI / 180;
    double ca = std::cos(a), cb = std::cos(b), sa = std::sin(a), sb = std::sin(b);
    v[0] = sb * ca;
    v[1] = sb * sa;
    v[2] = cb;
}

/* HealpixSampling::selectOrientationsWithNonZeroPriorProbability (healpix_sampling.cpp:769-1145)
 * for C1 (no symmetry mates), no bimodal psi, no multi-body terms, sigma_cutoff 3: the normalised
 * prior of every direction and psi angle, 0 outside the search range; the nearest one alone when
 * none is within range. Prior angles of 999 fall back to the current angles
 * (ml_optimiser.cpp:5851-5857). Returns the numbers of selected directions and psi angles. */
static void local_priors(const Sampling& S, const t2x_estep_config* cfg, const t2x_particle_in& pin,
                         std::vector<double>& dpr, std::vector<double>& ppr, long* nd, long* np)
{
    auto unset = [](double v) { return v > 998.99 && v < 999.01; };
    const double prot = unset(pin.prior_rot) ? pin.rot : pin.prior_rot;
    const double ptilt = unset(pin.prior_tilt) ? pin.tilt : pin.prior_tilt;
    const double ppsi = unset(pin.prior_psi) ? pin.psi : pin.prior_psi;
    const double cutoff = 3.;
    const long Nd = S.nr_dir(), Npsi = S.nr_psi();
    dpr.assign(Nd, 0.);
    ppr.assign(Npsi, 0.);
    *nd = *np = 0;
    if (S.is_3D) {
        double sum = 0., best_ang = 9999.;
        long best = -999;
        if (cfg->sigma_rot > 0. && cfg->sigma_tilt > 0.) {
            double pv[3], v[3];
            angles2direction(prot, ptilt, pv);
            const double sig = std::max(cfg->sigma_rot, cfg->sigma_tilt);
            for (long d = 0; d < Nd; d++) {
                angles2direction(S.rot[d], S.tilt[d], v);
                double dot = 0.;
                for (int i = 0; i < 3; i++) dot += v[i] * pv[i];   /* dotProduct (matrix1d.h) */
                double diffang = std::acos(dot) * 180. / PI;        /* ACOSD (macros.h:301) */
                if (diffang > 180.) diffang = std::fabs(diffang - 360.);
                if (diffang < cutoff * sig) {
                    dpr[d] = gaussian1D(diffang, sig, 0.);
                    sum += dpr[d];
                    (*nd)++;
                }
                if (diffang < best_ang) { best = d; best_ang = diffang; }
            }
        } else {
            for (long d = 0; d < Nd; d++) { dpr[d] = 1.; sum += 1.; (*nd)++; }
        }
        for (long d = 0; d < Nd; d++) if (dpr[d] > 0.) dpr[d] /= sum;
        if (*nd == 0) { dpr[best] = 1.; *nd = 1; }
    } else {
        dpr[0] = 1.;
        *nd = 1;
    }
    double sum = 0., best_diff = 9999.;
    long best = -999;
    for (long p = 0; p < Npsi; p++) {
        if (cfg->sigma_psi > 0.) {
            double diffpsi = std::fabs(S.psi[p] - ppsi);
            if (diffpsi > 180.) diffpsi = std::fabs(diffpsi - 360.);
            if (diffpsi < cutoff * cfg->sigma_psi) {
                ppr[p] = gaussian1D(diffpsi, cfg->sigma_psi, 0.);
                sum += ppr[p];
                (*np)++;
            }
            if (diffpsi < best_diff) { best = p; best_diff = diffpsi; }
        } else {
            ppr[p] = 1.;
            sum += 1.;
            (*np)++;
        }
    }
    for (long p = 0; p < Npsi; p++) if (ppr[p] > 0.) ppr[p] /= sum;
    if (*np == 0) { ppr[best] = 1.; *np = 1; }
}

extern "C" int orc_sampling_info(const t2x_estep_config* cfg, t2x_sampling_info* info)
{
    if (check_config(cfg) != T2X_OK || !info) return T2X_EINVAL;
    Sampling s;
    s.setup(cfg);
    info->nr_dir = s.nr_dir();
    info->nr_psi = s.nr_psi();
    info->nr_trans = s.nr_trans();
    info->nr_over_rot = s.over_rot(cfg->adaptive_oversampling);
    info->nr_over_trans = s.over_trans(cfg->adaptive_oversampling);
    info->nr_coarse_per_class = info->nr_dir * info->nr_psi;
    info->nr_hidden_coarse = cfg->nr_classes * info->nr_coarse_per_class * info->nr_trans;
    return T2X_OK;
}

extern "C" int orc_sampling_coarse(const t2x_estep_config* cfg, double* rot, double* tilt,
                                   double* psi, double* tx, double* ty)
{
    if (check_config(cfg) != T2X_OK) return T2X_EINVAL;
    Sampling s;
    s.setup(cfg);
    for (long i = 0; i < s.nr_dir(); i++) { if (rot) rot[i] = s.rot[i]; if (tilt) tilt[i] = s.tilt[i]; }
    for (long i = 0; i < s.nr_psi(); i++) if (psi) psi[i] = s.psi[i];
    for (long i = 0; i < s.nr_trans(); i++) { if (tx) tx[i] = s.tx[i]; if (ty) ty[i] = s.ty[i]; }
    return T2X_OK;
}

extern "C" int orc_sampling_oversampled_orientations(const t2x_estep_config* cfg, long idir,
                                                     long ipsi, int os, double* rot, double* tilt,
                                                     double* psi)
{
    if (check_config(cfg) != T2X_OK) return T2X_EINVAL;
    Sampling s;
    s.setup(cfg);
    std::vector<double> r, t, p;
    s.orientations(idir, ipsi, os, r, t, p);
    for (size_t i = 0; i < r.size(); i++) { rot[i] = r[i]; tilt[i] = t[i]; psi[i] = p[i]; }
    return (int)r.size();
}

extern "C" int orc_sampling_oversampled_translations(const t2x_estep_config* cfg, long itrans,
                                                     int os, double* x, double* y)
{
    if (check_config(cfg) != T2X_OK) return T2X_EINVAL;
    Sampling s;
    s.setup(cfg);
    std::vector<double> a, b;
    s.translations_px(itrans, os, cfg->pixel_size, a, b);
    for (size_t i = 0; i < a.size(); i++) { x[i] = a[i]; y[i] = b[i]; }
    return (int)a.size();
}

/* Local-search priors of one particle: dpr[nr_dir], ppr[nr_psi] (0 outside the range). */
extern "C" int orc_local_priors(const t2x_estep_config* cfg, const t2x_particle_in* pin, double* dpr,
                                double* ppr, long* nd, long* np)
{
    if (check_config(cfg) != T2X_OK || !pin) return T2X_EINVAL;
    Sampling s;
    s.setup(cfg);
    std::vector<double> d, p;
    local_priors(s, cfg, *pin, d, p, nd, np);
    std::copy(d.begin(), d.end(), dpr);
    std::copy(p.begin(), p.end(), ppr);
    return T2X_OK;
}

/* =====================================================================================
 * Euler matrix: Euler_angles2matrix (src/euler.cpp:52-91), ZYZ convention.
 * ===================================================================================== */
extern "C" void orc_euler_angles2matrix(double alpha, double beta, double gamma, double* A)
{
    /* DEG2RAD(d) = ((d) * PI / 180) (macros.h:275) */
    alpha = alpha * PI / 180; beta = beta * PI / 180; gamma = gamma * PI / 180;
    double ca = std::cos(alpha), cb = std::cos(beta), cg = std::cos(gamma);
    double sa = std::sin(alpha), sb = std::sin(beta), sg = std::sin(gamma);
    double cc = cb * ca, cs = cb * sa, sc = sb * ca, ss = sb * sa;
    A[0] = cg * cc - sg * sa;  A[1] = cg * cs + sg * ca;  A[2] = -cg * sb;
    A[3] = -sg * cc - cg * sa; A[4] = -sg * cs + cg * ca; A[5] = sg * sb;
    A[6] = sc;                 A[7] = ss;                 A[8] = cb;
}

/* =====================================================================================
 * CTF: CTF::initialise (ctf.cpp:206-260) and CTF::getCTF (ctf.h:181-230); no
 * magnification matrices, no aberrations, do_damping = true, no abs / phase flip.
 * ===================================================================================== */
struct Ctf {
    double K1, K2, K3, K4, K5, Axx, Axy, Ayy, scale;
    explicit Ctf(const t2x_ctf_params& p)
    {
        double local_Cs = p.cs_mm * 1e7;
        double local_kV = p.voltage_kv * 1e3;
        double rad_azimuth = p.defocus_angle * PI / 180.;
        double lambda = 12.2643247 / std::sqrt(local_kV * (1. + local_kV * 0.978466e-6));
        K1 = PI / 2 * 2 * lambda;
        K2 = PI / 2 * local_Cs * lambda * lambda * lambda;
        K3 = std::atan(p.q0 / std::sqrt(1 - p.q0 * p.q0));
        K4 = -p.bfactor / 4.;
        K5 = p.phase_shift * PI / 180.;
        double s = std::sin(rad_azimuth), c = std::cos(rad_azimuth);
        /* A = Q^T * D * Q with Q = [[c, s], [-s, c]], D = diag(-dU, -dV) */
        double dU = -p.defocus_u, dV = -p.defocus_v;
        Axx = c * c * dU + s * s * dV;
        Axy = c * s * dU - s * c * dV;
        Ayy = s * s * dU + c * c * dV;
        scale = p.scale;
    }
    double value(double X, double Y) const
    {
        double u2 = X * X + Y * Y;
        double u4 = u2 * u2;
        double gamma = K1 * (Axx * X * X + 2.0 * Axy * X * Y + Ayy * Y * Y) + K2 * u4 - K5 - K3;
        double r = -std::sin(gamma);
        r *= std::exp(K4 * u2);
        r *= scale;
        if (std::fabs(r) < 1e-8) r = ((r >= 0) ? 1 : -1) * 1e-8;
        return r;
    }
};

extern "C" double orc_ctf_value(const t2x_ctf_params* p, double X, double Y)
{
    return Ctf(*p).value(X, Y);
}

/* CTF::getFftwImage without padding (ctf.cpp:437-447): x = jp/(N*angpix), y = ip/(N*angpix) */
extern "C" void orc_ctf_image(const t2x_ctf_params* p, int ori_size, double angpix, int size,
                              double* out)
{
    Ctf c(*p);
    int xs = size / 2 + 1;
    double xsA = ori_size * angpix, ysA = ori_size * angpix;
    for (int i = 0; i < size; i++) {
        int ip = (i < xs) ? i : i - size;
        for (int j = 0; j < xs; j++) out[i * xs + j] = c.value(j / xsA, ip / ysA);
    }
}

/* =====================================================================================
 * FFT: mixed-radix recursive DFT (any n; radix = smallest prime factor).
 * RELION's FourierTransformer normalises the forward transform by 1/size (fftw.cpp:335-358).
 * ===================================================================================== */
static void dft_rec(const cplx* in, long stride, long n, cplx* out, int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    long p = 2;
    while (p * p <= n && n % p) p++;
    if (n % p) p = n; /* prime */
    long m = n / p;
    std::vector<cplx> sub(n);
    for (long r = 0; r < p; r++) dft_rec(in + r * stride, stride * p, m, &sub[r * m], sign);
    for (long k = 0; k < m; k++)
        for (long q = 0; q < p; q++) {
            long kk = k + m * q;
            cplx acc = 0;
            for (long r = 0; r < p; r++) {
                double ang = sign * 2.0 * PI * (double)((r * kk) % n) / (double)n;
                acc += sub[r * m + k] * cplx(std::cos(ang), std::sin(ang));
            }
            out[kk] = acc;
        }
}

static void fft2_full(std::vector<cplx>& a, int n, int sign)
{
    std::vector<cplx> row(n), res(n);
    for (int i = 0; i < n; i++) {
        dft_rec(&a[(size_t)i * n], 1, n, res.data(), sign);
        std::copy(res.begin(), res.end(), a.begin() + (size_t)i * n);
    }
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) row[i] = a[(size_t)i * n + j];
        dft_rec(row.data(), 1, n, res.data(), sign);
        for (int i = 0; i < n; i++) a[(size_t)i * n + j] = res[i];
    }
}

extern "C" void orc_fft2_r2c(int n, const double* in, double* out)
{
    std::vector<cplx> a((size_t)n * n);
    for (size_t i = 0; i < a.size(); i++) a[i] = in[i];
    fft2_full(a, n, -1);
    int xs = n / 2 + 1;
    double norm = 1.0 / ((double)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < xs; j++) {
            cplx v = a[(size_t)i * n + j] * norm;
            out[2 * ((size_t)i * xs + j)] = v.real();
            out[2 * ((size_t)i * xs + j) + 1] = v.imag();
        }
}

/* Hermitian completion of the half array, then an unnormalised inverse (FFTW c2r convention:
 * imaginary parts on self-conjugate frequencies are dropped). */
extern "C" void orc_fft2_c2r(int n, const double* in, double* out)
{
    int xs = n / 2 + 1;
    std::vector<cplx> a((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            cplx v;
            if (j < xs) {
                v = cplx(in[2 * ((size_t)i * xs + j)], in[2 * ((size_t)i * xs + j) + 1]);
                if (j == 0 || (j == n / 2)) {
                    /* column j is its own mirror: use the Hermitian average */
                    int im = (n - i) % n;
                    cplx w(in[2 * ((size_t)im * xs + j)], -in[2 * ((size_t)im * xs + j) + 1]);
                    v = 0.5 * (v + w);
                }
            } else {
                int im = (n - i) % n, jm = n - j;
                v = cplx(in[2 * ((size_t)im * xs + jm)], -in[2 * ((size_t)im * xs + jm) + 1]);
            }
            a[(size_t)i * n + j] = v;
        }
    fft2_full(a, n, +1);
    for (size_t i = 0; i < a.size(); i++) out[i] = a[i].real();
}

/* Matrix2D::inv for 3x3 matrices (matrix2d.h:1108-1123): adjugate divided by the determinant.
 * For a rotation this equals the transpose up to rounding; the reference uses this form. */
static void inv3(const double* A, double* R)
{
    auto M = [&](int r, int c) { return A[r * 3 + c]; };
    R[0] = M(2, 2) * M(1, 1) - M(2, 1) * M(1, 2);
    R[1] = -(M(2, 2) * M(0, 1) - M(2, 1) * M(0, 2));
    R[2] = M(1, 2) * M(0, 1) - M(1, 1) * M(0, 2);
    R[3] = -(M(2, 2) * M(1, 0) - M(2, 0) * M(1, 2));
    R[4] = M(2, 2) * M(0, 0) - M(2, 0) * M(0, 2);
    R[5] = -(M(1, 2) * M(0, 0) - M(1, 0) * M(0, 2));
    R[6] = M(2, 1) * M(1, 0) - M(2, 0) * M(1, 1);
    R[7] = -(M(2, 1) * M(0, 0) - M(2, 0) * M(0, 1));
    R[8] = M(1, 1) * M(0, 0) - M(1, 0) * M(0, 1);
    double tmp = M(0, 0) * R[0] + M(1, 0) * R[1] + M(2, 0) * R[2];
    for (int i = 0; i < 9; i++) R[i] /= tmp;
}

/* =====================================================================================
 * Projector::project (projector.cpp:524-691), trilinear branch.
 * Ainv = A.inv() * padding_factor.
 * ===================================================================================== */
static inline bool project_pixel(const orc_volume* v, const double* Ai, int x, int y, bool clean,
                                 cplx* res)
{
    const int pad = v->pad_size, xdim = pad / 2 + 1, start = -(pad / 2);
    double xp = Ai[0] * x + Ai[1] * y;
    double yp = Ai[3] * x + Ai[4] * y;
    double zp = Ai[6] * x + Ai[7] * y;
    const int r_max_ref = (int)(v->r_max * v->padding_factor);
    const int r_max_ref_2 = r_max_ref * r_max_ref;
    if (clean) {
        /* rotation-invariant form of the same test: |Ainv (x,y)|^2 = pf^2 (x^2+y^2) */
        double r2 = v->padding_factor * v->padding_factor * (double)(x * x + y * y);
        if (r2 > r_max_ref_2) return false;
    } else {
        double r_ref_2 = xp * xp + yp * yp + zp * zp;
        if (r_ref_2 > r_max_ref_2) return false;
    }
    bool is_neg_x = xp < 0;
    if (is_neg_x) { xp = -xp; yp = -yp; zp = -zp; }
    int x0 = R_FLOOR(xp);
    double fx = xp - x0;
    int x1 = x0 + 1;
    int y0 = R_FLOOR(yp);
    double fy = yp - y0;
    y0 -= start;
    int y1 = y0 + 1;
    int z0 = R_FLOOR(zp);
    double fz = zp - z0;
    z0 -= start;
    int z1 = z0 + 1;
    if (x0 < 0 || x0 + 1 >= xdim || y0 < 0 || y0 + 1 >= pad || z0 < 0 || z0 + 1 >= pad) return false;
    const cplx* d = reinterpret_cast<const cplx*>(v->data);
    auto at = [&](int z, int yy, int xx) { return d[((size_t)z * pad + yy) * xdim + xx]; };
    cplx d000 = at(z0, y0, x0), d001 = at(z0, y0, x1), d010 = at(z0, y1, x0), d011 = at(z0, y1, x1);
    cplx d100 = at(z1, y0, x0), d101 = at(z1, y0, x1), d110 = at(z1, y1, x0), d111 = at(z1, y1, x1);
    cplx dx00 = d000 + (d001 - d000) * fx;
    cplx dx01 = d100 + (d101 - d100) * fx;
    cplx dx10 = d010 + (d011 - d010) * fx;
    cplx dx11 = d110 + (d111 - d110) * fx;
    cplx dxy0 = dx00 + (dx10 - dx00) * fy;
    cplx dxy1 = dx01 + (dx11 - dx01) * fy;
    cplx val = dxy0 + (dxy1 - dxy0) * fz;
    *res = is_neg_x ? std::conj(val) : val;
    return true;
}

/* Projector::rotate2D (projector.cpp:904-1016), bilinear branch, for 2D references
 * (pad x (pad/2+1), y centred). Its radius test truncates xp^2 + yp^2 to an integer (the
 * reference declares `const int r_ref_2`); the clean form applies the same truncation to the
 * rotation-invariant pf^2 (x^2 + y^2). */
static inline bool project_pixel_2d(const orc_volume* v, const double* Ai, int x, int y, bool clean,
                                    cplx* res)
{
    const int pad = v->pad_size, xdim = pad / 2 + 1, start = -(pad / 2);
    double xp = Ai[0] * x + Ai[1] * y;
    double yp = Ai[3] * x + Ai[4] * y;
    const int r_max_ref = (int)(v->r_max * v->padding_factor);
    const int r_max_ref_2 = r_max_ref * r_max_ref;
    if (clean) {
        double r2 = v->padding_factor * v->padding_factor * (double)(x * x + y * y);
        if ((int)r2 > r_max_ref_2) return false;
    } else {
        const int r_ref_2 = (int)(xp * xp + yp * yp);
        if (r_ref_2 > r_max_ref_2) return false;
    }
    bool is_neg_x = xp < 0;
    if (is_neg_x) { xp = -xp; yp = -yp; }
    int x0 = R_FLOOR(xp);
    double fx = xp - x0;
    int x1 = x0 + 1;
    int y0 = R_FLOOR(yp);
    double fy = yp - y0;
    y0 -= start;
    int y1 = y0 + 1;
    if (x0 < 0 || x1 >= xdim || y0 < 0 || y1 >= pad) return false;   /* never within r_max */
    const cplx* d = reinterpret_cast<const cplx*>(v->data);
    cplx d00 = d[(size_t)y0 * xdim + x0], d01 = d[(size_t)y0 * xdim + x1];
    cplx d10 = d[(size_t)y1 * xdim + x0], d11 = d[(size_t)y1 * xdim + x1];
    cplx dx0 = d00 + (d01 - d00) * fx;
    cplx dx1 = d10 + (d11 - d10) * fx;
    cplx val = dx0 + (dx1 - dx0) * fy;
    *res = is_neg_x ? std::conj(val) : val;
    return true;
}

static void project_into(const orc_volume* v, const double* A, int size, cplx* f2d, bool clean)
{
    double Ai[9];
    inv3(A, Ai);
    for (int i = 0; i < 9; i++) Ai[i] *= v->padding_factor;
    const int xs = size / 2 + 1;
    const int r_max_out = xs - 1;
    const int r_max_out_2 = r_max_out * r_max_out;
    for (int i = 0; i < size; i++) {
        const int y = (i <= r_max_out) ? i : i - size;
        const int x_max = R_FLOOR(std::sqrt((double)(r_max_out_2 - y * y)));
        for (int x = 0; x <= x_max; x++) {
            cplx val;
            const bool in = v->ref_dim == 2 ? project_pixel_2d(v, Ai, x, y, clean, &val)
                                             : project_pixel(v, Ai, x, y, clean, &val);
            if (in) f2d[i * xs + x] = val;
        }
    }
}

extern "C" void orc_project(const orc_volume* v, const double* A9, int size, double* out)
{
    int xs = size / 2 + 1;
    std::memset(out, 0, sizeof(double) * 2 * size * xs);
    project_into(v, A9, size, reinterpret_cast<cplx*>(out), true);
}

/* BackProjector::backproject2Dto3D (backprojector.cpp:55-357), trilinear, no Ewald sphere,
 * no magnification matrix. */
static void backproject_from(orc_volume* v, const cplx* f2d, const double* fw, int s, const double* A)
{
    const double pf = v->padding_factor;
    double Ai[9];
    inv3(A, Ai);
    for (int i = 0; i < 9; i++) Ai[i] *= pf;
    const int max_r2 = R_ROUND(v->r_max * pf) * R_ROUND(v->r_max * pf);
    const double Am_Xx = Ai[0], Am_Xy = Ai[1], Am_Yx = Ai[3], Am_Yy = Ai[4], Am_Zx = Ai[6], Am_Zy = Ai[7];
    const double AtA_xx = Am_Xx * Am_Xx + Am_Yx * Am_Yx + Am_Zx * Am_Zx;
    const double AtA_xy = Am_Xx * Am_Xy + Am_Yx * Am_Yy + Am_Zx * Am_Zy;
    const double AtA_yy = Am_Xy * Am_Xy + Am_Yy * Am_Yy + Am_Zy * Am_Zy;
    const double AtA_xy2 = AtA_xy * AtA_xy;
    const int sh = s / 2 + 1;
    const int pad = v->pad_size, xdim = pad / 2 + 1, start = -(pad / 2);
    cplx* d = reinterpret_cast<cplx*>(v->data);
    double* w = v->weight;
    for (int i = 0; i < s; i++) {
        int y, first_allowed_x;
        if (i < sh) { y = i; first_allowed_x = 0; }
        else { y = i - s; first_allowed_x = 1; }
        double discr = AtA_xy2 * y * y - AtA_xx * (AtA_yy * y * y - max_r2);
        if (discr < 0.0) continue;
        double dd = std::sqrt(discr) / AtA_xx;
        double q = -AtA_xy * y / AtA_xx;
        int first_x = R_CEIL(q - dd);
        int last_x = R_FLOOR(q + dd);
        if (first_x < first_allowed_x) first_x = first_allowed_x;
        if (last_x > sh - 1) last_x = sh - 1;
        for (int x = first_x; x <= last_x; x++) {
            cplx my_val = f2d[i * sh + x];
            double my_weight = fw ? fw[i * sh + x] : 1.0;
            if (my_weight <= 0.) continue;
            double xp = Ai[0] * x + Ai[1] * y;
            double yp = Ai[3] * x + Ai[4] * y;
            double zp = Ai[6] * x + Ai[7] * y;
            double r2_3D = xp * xp + yp * yp + zp * zp;
            if (r2_3D > max_r2) continue;
            bool is_neg_x = xp < 0;
            if (is_neg_x) { xp = -xp; yp = -yp; zp = -zp; }
            int x0 = R_FLOOR(xp);
            double fx = xp - x0;
            int x1 = x0 + 1;
            int y0 = R_FLOOR(yp);
            double fy = yp - y0;
            y0 -= start;
            int y1 = y0 + 1;
            int z0 = R_FLOOR(zp);
            double fz = zp - z0;
            z0 -= start;
            int z1 = z0 + 1;
            if (x0 < 0 || x0 + 1 >= xdim || y0 < 0 || y0 + 1 >= pad || z0 < 0 || z0 + 1 >= pad) continue;
            double mfx = 1. - fx, mfy = 1. - fy, mfz = 1. - fz;
            double dd000 = mfz * mfy * mfx, dd001 = mfz * mfy * fx, dd010 = mfz * fy * mfx, dd011 = mfz * fy * fx;
            double dd100 = fz * mfy * mfx, dd101 = fz * mfy * fx, dd110 = fz * fy * mfx, dd111 = fz * fy * fx;
            if (is_neg_x) my_val = std::conj(my_val);
            auto idx = [&](int z, int yy, int xx) { return ((size_t)z * pad + yy) * xdim + xx; };
            d[idx(z0, y0, x0)] += dd000 * my_val; d[idx(z0, y0, x1)] += dd001 * my_val;
            d[idx(z0, y1, x0)] += dd010 * my_val; d[idx(z0, y1, x1)] += dd011 * my_val;
            d[idx(z1, y0, x0)] += dd100 * my_val; d[idx(z1, y0, x1)] += dd101 * my_val;
            d[idx(z1, y1, x0)] += dd110 * my_val; d[idx(z1, y1, x1)] += dd111 * my_val;
            w[idx(z0, y0, x0)] += dd000 * my_weight; w[idx(z0, y0, x1)] += dd001 * my_weight;
            w[idx(z0, y1, x0)] += dd010 * my_weight; w[idx(z0, y1, x1)] += dd011 * my_weight;
            w[idx(z1, y0, x0)] += dd100 * my_weight; w[idx(z1, y0, x1)] += dd101 * my_weight;
            w[idx(z1, y1, x0)] += dd110 * my_weight; w[idx(z1, y1, x1)] += dd111 * my_weight;
        }
    }
}

/* BackProjector::backrotate2D (backprojector.cpp:471-668), bilinear, no magnification matrix:
 * the ellipse x-range of every row (integer-truncated r_max * pf), no further radius test. */
static void backrotate_from(orc_volume* v, const cplx* f2d, const double* fw, int s, const double* A)
{
    const double pf = v->padding_factor;
    double Ai[9];
    inv3(A, Ai);
    for (int i = 0; i < 9; i++) Ai[i] *= pf;
    const int r_max_ref = (int)(v->r_max * pf);
    const int r_max_ref_2 = r_max_ref * r_max_ref;
    const double Am_Xx = Ai[0] * 1.0 + Ai[1] * 0.0, Am_Xy = Ai[0] * 0.0 + Ai[1] * 1.0;
    const double Am_Yx = Ai[3] * 1.0 + Ai[4] * 0.0, Am_Yy = Ai[3] * 0.0 + Ai[4] * 1.0;
    const double AtA_xx = Am_Xx * Am_Xx + Am_Yx * Am_Yx;
    const double AtA_xy = Am_Xx * Am_Xy + Am_Yx * Am_Yy;
    const double AtA_yy = Am_Xy * Am_Xy + Am_Yy * Am_Yy;
    const double AtA_xy2 = AtA_xy * AtA_xy;
    const int sh = s / 2 + 1;
    const int pad = v->pad_size, xdim = pad / 2 + 1, start = -(pad / 2);
    cplx* d = reinterpret_cast<cplx*>(v->data);
    double* w = v->weight;
    for (int i = 0; i < s; i++) {
        int y, first_allowed_x;
        if (i < sh) { y = i; first_allowed_x = 0; }
        else { y = i - s; first_allowed_x = 1; }
        double discr = AtA_xy2 * y * y - AtA_xx * (AtA_yy * y * y - r_max_ref_2);
        if (discr < 0.0) continue;
        double dd = std::sqrt(discr) / AtA_xx;
        double q = -AtA_xy * y / AtA_xx;
        int first_x = R_CEIL(q - dd);
        int last_x = R_FLOOR(q + dd);
        if (first_x < first_allowed_x) first_x = first_allowed_x;
        if (last_x > sh - 1) last_x = sh - 1;
        for (int x = first_x; x <= last_x; x++) {
            double my_weight = fw ? fw[i * sh + x] : 1.0;
            if (my_weight <= 0.) continue;
            cplx my_val = f2d[i * sh + x];
            double xu = 1.0 * x + 0.0 * y, yu = 0.0 * x + 1.0 * y;
            double xp = Ai[0] * xu + Ai[1] * yu;
            double yp = Ai[3] * xu + Ai[4] * yu;
            bool is_neg_x = xp < 0;
            if (is_neg_x) { xp = -xp; yp = -yp; }
            int x0 = R_FLOOR(xp);
            double fx = xp - x0;
            int x1 = x0 + 1;
            int y0 = R_FLOOR(yp);
            double fy = yp - y0;
            y0 -= start;
            int y1 = y0 + 1;
            if (x0 < 0 || x1 >= xdim || y0 < 0 || y1 >= pad) continue;   /* never within r_max */
            double mfx = 1. - fx, mfy = 1. - fy;
            double dd00 = mfy * mfx, dd01 = mfy * fx, dd10 = fy * mfx, dd11 = fy * fx;
            if (is_neg_x) my_val = std::conj(my_val);
            d[(size_t)y0 * xdim + x0] += dd00 * my_val; d[(size_t)y0 * xdim + x1] += dd01 * my_val;
            d[(size_t)y1 * xdim + x0] += dd10 * my_val; d[(size_t)y1 * xdim + x1] += dd11 * my_val;
            w[(size_t)y0 * xdim + x0] += dd00 * my_weight; w[(size_t)y0 * xdim + x1] += dd01 * my_weight;
            w[(size_t)y1 * xdim + x0] += dd10 * my_weight; w[(size_t)y1 * xdim + x1] += dd11 * my_weight;
        }
    }
}

extern "C" void orc_backproject(orc_volume* v, const double* f2d, const double* fweight, int size,
                                const double* A9)
{
    if (v->ref_dim == 2) backrotate_from(v, reinterpret_cast<const cplx*>(f2d), fweight, size, A9);
    else backproject_from(v, reinterpret_cast<const cplx*>(f2d), fweight, size, A9);
}

/* =====================================================================================
 * Synthetic data (shared formulas with the GPU generator; both in the build's own design).
 * A Gaussian blob amp*exp(-|r-c|^2/(2 s^2)) has the continuous transform
 *   amp*(2 pi s^2)^{3/2} exp(-2 pi^2 s^2 |f|^2) exp(-2 pi i f.c)    (f in cycles/px).
 * The projector grid holds V(f)/N^2 (the scale of Projector::computeFourierTransformMap for
 * 3D references and 2D images, projector.cpp:146-160, 407).
 * ===================================================================================== */
static inline cplx blob_ft(const t2x_blob* b, int nb, double fx, double fy, double fz)
{
    double f2 = fx * fx + fy * fy + fz * fz;
    double re = 0, im = 0;
    for (int k = 0; k < nb; k++) {
        double s2 = b[k].sigma * b[k].sigma;
        double mag = b[k].amplitude * std::pow(2 * PI * s2, 1.5) * std::exp(-2 * PI * PI * s2 * f2);
        double ph = -2 * PI * (fx * b[k].x + fy * b[k].y + fz * b[k].z);
        re += mag * std::cos(ph);
        im += mag * std::sin(ph);
    }
    return cplx(re, im);
}

static inline double lowpass_weight(double shell, double cutoff_shell)
{
    if (cutoff_shell <= 0) return 1.0;
    double lo = cutoff_shell - 1.0, hi = cutoff_shell + 1.0;
    if (shell <= lo) return 1.0;
    if (shell >= hi) return 0.0;
    return 0.5 + 0.5 * std::cos(PI * (shell - lo) / (hi - lo));
}

extern "C" int orc_synth_reference(int N, double angpix, int pad, int r_max, double pf,
                                   const t2x_blob* blobs, int nb, double lowpass_A, double* out)
{
    int xdim = pad / 2 + 1, h = pad / 2;
    int max_r2 = R_ROUND(r_max * pf) * R_ROUND(r_max * pf);
    double cutoff_shell = lowpass_A > 0 ? N * angpix / lowpass_A : 0;
    double inv = 1.0 / ((double)N * N);
#pragma omp parallel for schedule(dynamic)
    for (int kz = 0; kz < pad; kz++)
        for (int ky = 0; ky < pad; ky++)
            for (int kx = 0; kx < xdim; kx++) {
                int z = kz - h, y = ky - h, x = kx;
                size_t idx = ((size_t)kz * pad + ky) * xdim + kx;
                int r2 = x * x + y * y + z * z;
                cplx v = 0;
                if (r2 <= max_r2) {
                    /* voxels beyond the low-pass edge are zero: their blob sum is not evaluated */
                    const double lw = lowpass_weight(std::sqrt((double)r2) / pf, cutoff_shell);
                    double f = 1.0 / (pf * N);
                    if (lw > 0.) v = blob_ft(blobs, nb, x * f, y * f, z * f) * inv * lw;
                }
                out[2 * idx] = v.real();
                out[2 * idx + 1] = v.imag();
            }
    return T2X_OK;
}

/* The central (kz = 0) plane of orc_synth_reference's volume alone: the transform of the blobs'
 * z-projection, i.e. a 2D class average (pad x (pad/2+1) complex). */
extern "C" int orc_synth_reference_plane(int N, double angpix, int pad, int r_max, double pf,
                                         const t2x_blob* blobs, int nb, double lowpass_A, double* out)
{
    int xdim = pad / 2 + 1, h = pad / 2;
    int max_r2 = R_ROUND(r_max * pf) * R_ROUND(r_max * pf);
    double cutoff_shell = lowpass_A > 0 ? N * angpix / lowpass_A : 0;
    double inv = 1.0 / ((double)N * N);
#pragma omp parallel for schedule(dynamic)
    for (int ky = 0; ky < pad; ky++)
        for (int kx = 0; kx < xdim; kx++) {
            int y = ky - h, x = kx;
            size_t idx = (size_t)ky * xdim + kx;
            int r2 = x * x + y * y;
            cplx v = 0;
            if (r2 <= max_r2) {
                double f = 1.0 / (pf * N);
                v = blob_ft(blobs, nb, x * f, y * f, 0 * f) * inv;
                v *= lowpass_weight(std::sqrt((double)r2) / pf, cutoff_shell);
            }
            out[2 * idx] = v.real();
            out[2 * idx + 1] = v.imag();
        }
    return T2X_OK;
}

/* splitmix64 finaliser */
static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Counter-based Gaussian: Box-Muller on two 53-bit uniforms keyed by (seed, particle, pixel). */
extern "C" double orc_noise_sample(uint64_t seed, long particle, long pixel)
{
    uint64_t k = mix64(seed ^ mix64((uint64_t)particle * 0x100000001B3ULL + 0x51ED27ULL));
    uint64_t a = mix64(k + 2 * (uint64_t)pixel);
    uint64_t b = mix64(k + 2 * (uint64_t)pixel + 1);
    double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740992.0); /* (0,1] */
    double u2 = (b >> 11) * (1.0 / 9007199254740992.0);       /* [0,1) */
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
}

extern "C" int orc_synth_particles(int N, double angpix, int n, const t2x_blob* blobs,
                                   const int* boff, const t2x_particle_in* pin,
                                   const double* trot, const double* ttilt, const double* tpsi,
                                   const double* tx, const double* ty, const int* tcls,
                                   const double* snr, uint64_t seed, float* images,
                                   double* noise_var)
{
    int xs = N / 2 + 1;
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < n; p++) {
        double A[9];
        orc_euler_angles2matrix(trot[p], ttilt[p], tpsi[p], A);
        Ctf ctf(pin[p].ctf);
        const t2x_blob* b = blobs + boff[tcls[p]];
        int nb = boff[tcls[p] + 1] - boff[tcls[p]];
        std::vector<double> F((size_t)2 * N * xs), img((size_t)N * N);
        for (int i = 0; i < N; i++) {
            int ip = (i < xs) ? i : i - N;
            for (int j = 0; j < xs; j++) {
                double gx = (double)j / N, gy = (double)ip / N;
                /* slice theorem with the projector convention (xp,yp,zp) = A^T (x,y,0) */
                double fx = A[0] * gx + A[3] * gy, fy = A[1] * gx + A[4] * gy, fz = A[2] * gx + A[5] * gy;
                cplx v = blob_ft(b, nb, fx, fy, fz) / ((double)N * N);
                double ph = -2 * PI * (gx * tx[p] + gy * ty[p]);
                v *= cplx(std::cos(ph), std::sin(ph));
                v *= ctf.value(j / (N * angpix), ip / (N * angpix));
                if ((ip + j) & 1) v = -v; /* origin of the array at index 0 */
                F[2 * ((size_t)i * xs + j)] = v.real();
                F[2 * ((size_t)i * xs + j) + 1] = v.imag();
            }
        }
        orc_fft2_c2r(N, F.data(), img.data());
        double mean = 0, var = 0;
        for (double v : img) mean += v;
        mean /= (double)img.size();
        for (double v : img) var += (v - mean) * (v - mean);
        var /= (double)img.size();
        double sd = std::sqrt(var / snr[p]);
        if (noise_var) noise_var[p] = var / snr[p];
        for (long q = 0; q < (long)N * N; q++)
            images[(size_t)p * N * N + q] = (float)(img[q] + sd * orc_noise_sample(seed, p, q));
    }
    return T2X_OK;
}

/* =====================================================================================
 * Preprocessing: getFourierTransformsAndCtfs (ml_optimiser.cpp:5689-6593), single image,
 * no helix, no bodies, no subtomo, zero mask (softMaskOutsideMap with average background,
 * mask.cpp:43-95), no phase demodulation / MTF.
 * ===================================================================================== */
struct Prep {
    int cs = 0;
    std::vector<cplx> Fimg, Fimg_nomask; /* current size, CenterFFTbySign applied */
    std::vector<double> Fctf;            /* current size */
    std::vector<double> power;           /* [N/2+1] */
    double highres_Xi2 = 0;
    double old_x = 0, old_y = 0, prior_x = 0, prior_y = 0;
};

/* windowFourierTransform (fftw.h:737-787), shrinking case: out at logical (ip, jp) = in at (ip, jp) */
template <class T>
static void window_ft(const T* in, int nin, T* out, int nout)
{
    int xin = nin / 2 + 1, xout = nout / 2 + 1;
    for (int i = 0; i < nout; i++) {
        int ip = (i < xout) ? i : i - nout;
        int ii = ip < 0 ? ip + nin : ip;
        for (int j = 0; j < xout; j++) out[i * xout + j] = in[ii * xin + j];
    }
}

/* CenterFFTbySign (fftw.h:389-402): multiply by (-1)^(i+j) */
static void center_by_sign(cplx* a, int n)
{
    int xs = n / 2 + 1;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < xs; j++)
            if ((i ^ j) & 1) a[i * xs + j] = -a[i * xs + j];
}

static void fft_real(const std::vector<double>& img, int N, std::vector<cplx>& F)
{
    F.assign((size_t)N * (N / 2 + 1), 0);
    orc_fft2_r2c(N, img.data(), reinterpret_cast<double*>(F.data()));
}

/* The background of the soft mask when do_zero_mask is off (ml_optimiser.cpp:6173-6226): a noise
 * image whose Fourier components (FFTW half layout, N x (N/2+1)) within shell N/2 are complex
 * Gaussians of standard deviation sqrt(sigma2_fudge * sigma2_noise[shell]) per part, transformed
 * back without normalisation. The reference draws them from the process-wide rand() stream,
 * reseeded per particle (init_random_generator(random_seed + part_id)); here each component is
 * the counter-based sample keyed by (random_seed, part_id, 2 q + part) with q the component's
 * index, and the x = 0 and x = N/2 columns are made Hermitian (row -y the conjugate of row y, real
 * at the self-conjugate frequencies) so that the real inverse transform is unambiguous. */
static void background_noise(const t2x_estep_config* cfg, const double* sigma2, long part_id,
                             std::vector<double>& noise)
{
    const int N = cfg->ori_size, xN = N / 2 + 1;
    const uint64_t seed = cfg->random_seed ^ 0x6d61736b6e6f6973ULL;
    std::vector<double> F((size_t)2 * N * xN, 0.);
    for (int i = 0; i < N; i++) {
        const int ip = (i < xN) ? i : i - N;
        for (int j = 0; j < xN; j++) {
            const int ires = R_ROUND(std::sqrt((double)(ip * ip + j * j)));
            if (ires >= xN) continue;
            const bool edge = (j == 0 || j == N / 2);
            int src = i;                        /* row whose draws this component uses */
            bool conj = false;
            if (edge && ip < 0) { src = N - i; conj = true; }
            const long q = (long)src * xN + j;
            const double sd = std::sqrt(cfg->sigma2_fudge * sigma2[ires]);
            double re = sd * orc_noise_sample(seed, part_id, 2 * q);
            double im = sd * orc_noise_sample(seed, part_id, 2 * q + 1);
            if (edge && (i == 0 || i == N / 2)) im = 0.;
            F[2 * ((size_t)i * xN + j)] = re;
            F[2 * ((size_t)i * xN + j) + 1] = conj ? -im : im;
        }
    }
    noise.assign((size_t)N * N, 0.);
    orc_fft2_c2r(N, F.data(), noise.data());
}

extern "C" int orc_background_noise(const t2x_estep_config* cfg, const double* sigma2_noise, long part_id,
                                    double* noise)
{
    if (check_config(cfg) != T2X_OK || !sigma2_noise || !noise) return T2X_EINVAL;
    std::vector<double> v;
    background_noise(cfg, sigma2_noise, part_id, v);
    std::copy(v.begin(), v.end(), noise);
    return T2X_OK;
}

static void preprocess(const t2x_estep_config* cfg, const float* image, const t2x_particle_in& pin,
                       const double* sigma2, Prep& P)
{
    const int N = cfg->ori_size;
    const int cs = cfg->current_size;
    P.cs = cs;
    std::vector<double> img((size_t)N * N);
    for (size_t q = 0; q < img.size(); q++) img[q] = image[q];
    if (cfg->do_norm_correction) {
        double f = cfg->avg_norm_correction / pin.norm_correction;
        for (double& v : img) v *= f;
    }
    /* my_old_offset.selfROUND(); selfTranslate(img, my_old_offset, DONT_WRAP) — integer shift
     * with zero fill (transformations.h:744-772) */
    int sx = R_ROUND(pin.old_offset_x), sy = R_ROUND(pin.old_offset_y);
    if (sx != 0 || sy != 0) {
        std::vector<double> t((size_t)N * N, 0.0);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                int si = i - sy, sj = j - sx;
                if (si >= 0 && si < N && sj >= 0 && sj < N) t[(size_t)i * N + j] = img[(size_t)si * N + sj];
            }
        img.swap(t);
    }
    P.old_x = sx;
    P.old_y = sy;
    double px = pin.prior_offset_x, py = pin.prior_offset_y;
    if (px > 998.99 && px < 999.01) px = 0.;
    if (py > 998.99 && py < 999.01) py = 0.;
    P.prior_x = px;
    P.prior_y = py;

    std::vector<cplx> Faux;
    fft_real(img, N, Faux);
    P.Fimg_nomask.assign((size_t)cs * (cs / 2 + 1), 0);
    window_ft(Faux.data(), N, P.Fimg_nomask.data(), cs);
    center_by_sign(P.Fimg_nomask.data(), cs);

    /* softMaskOutsideMap(img, radius, width) with the average-background fill (--zero_mask),
     * or with the noise image Mnoise (mask.cpp:43-97) */
    double radius = cfg->particle_diameter / (2. * cfg->pixel_size);
    double cw = (double)cfg->width_mask_edge;
    double radius_p = radius + cw;
    std::vector<double> Mnoise;
    if (!cfg->do_zero_mask) background_noise(cfg, sigma2, pin.part_id, Mnoise);
    double sum_bg = 0, sum = 0;
    for (int i = 0; i < N && cfg->do_zero_mask; i++)
        for (int j = 0; j < N; j++) {
            int y = i - N / 2, x = j - N / 2;
            double r = std::sqrt((double)(x * x + y * y));
            double v = img[(size_t)i * N + j];
            if (r < radius) continue;
            if (r > radius_p) { sum += 1.; sum_bg += v; }
            else {
                double rc = 0.5 + 0.5 * std::cos(PI * (radius_p - r) / cw);
                sum += rc;
                sum_bg += rc * v;
            }
        }
    if (cfg->do_zero_mask) sum_bg /= sum;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            int y = i - N / 2, x = j - N / 2;
            double r = std::sqrt((double)(x * x + y * y));
            double& v = img[(size_t)i * N + j];
            if (r < radius) continue;
            const double add = cfg->do_zero_mask ? sum_bg : Mnoise[(size_t)i * N + j];
            if (r > radius_p) v = add;
            else {
                double rc = 0.5 + 0.5 * std::cos(PI * (radius_p - r) / cw);
                v = (1 - rc) * v + rc * add;
            }
        }

    fft_real(img, N, Faux);
    const int xN = N / 2 + 1;
    P.power.assign(xN, 0.0);
    P.highres_Xi2 = 0;
    if (cs < N) {
        for (int i = 0; i < N; i++) {
            int ip = (i < xN) ? i : i - N;
            for (int j = 0; j < xN; j++) {
                int ires = R_ROUND(std::sqrt((double)(ip * ip + j * j)));
                if (ires > 0 && ires < xN && !(j == 0 && ip < 0)) {
                    double nf = std::norm(Faux[(size_t)i * xN + j]);
                    P.power[ires] += nf;
                    if (ires >= cs / 2 + 1) P.highres_Xi2 += nf;
                }
            }
        }
    }
    P.Fimg.assign((size_t)cs * (cs / 2 + 1), 0);
    window_ft(Faux.data(), N, P.Fimg.data(), cs);
    center_by_sign(P.Fimg.data(), cs);

    P.Fctf.assign((size_t)cs * (cs / 2 + 1), 1.0);
    if (cfg->do_ctf_correction) orc_ctf_image(&pin.ctf, N, cfg->pixel_size, cs, P.Fctf.data());
}

extern "C" int orc_preprocess(const t2x_estep_config* cfg, const float* image,
                              const t2x_particle_in* pin, const double* sigma2_noise, double* Fimg,
                              double* Fimg_nomask, double* Fctf, double* power, double* highres,
                              double* old_xy)
{
    if (check_config(cfg) != T2X_OK) return T2X_EINVAL;
    if (!cfg->do_zero_mask && !sigma2_noise) return T2X_EINVAL;
    Prep P;
    preprocess(cfg, image, *pin, sigma2_noise, P);
    size_t nc = P.Fimg.size();
    if (Fimg) std::memcpy(Fimg, P.Fimg.data(), nc * sizeof(cplx));
    if (Fimg_nomask) std::memcpy(Fimg_nomask, P.Fimg_nomask.data(), nc * sizeof(cplx));
    if (Fctf) std::memcpy(Fctf, P.Fctf.data(), nc * sizeof(double));
    if (power) std::memcpy(power, P.power.data(), P.power.size() * sizeof(double));
    if (highres) *highres = P.highres_Xi2;
    if (old_xy) { old_xy[0] = P.old_x; old_xy[1] = P.old_y; }
    return T2X_OK;
}

/* Mresol map (updateImageSizeAndResolutionPointers, ml_optimiser.cpp:5596-5625) */
static std::vector<int> mresol(int s)
{
    int xs = s / 2 + 1;
    std::vector<int> m((size_t)s * xs, -1);
    for (int i = 0; i < s; i++) {
        int ip = (i < xs) ? i : i - s;
        for (int j = 0; j < xs; j++) {
            int ires = R_ROUND(std::sqrt((double)(ip * ip + j * j)));
            if (ires < xs && !(j == 0 && ip < 0)) m[(size_t)i * xs + j] = ires;
        }
    }
    return m;
}

/* shiftImageInFourierTransform (fftw.cpp:842-930), 2D: out = in * exp(i 2 pi (x*sx + y*sy)), s = -shift/oridim */
static void shift_image(const cplx* in, int s, double oridim, double xshift, double yshift, cplx* out)
{
    int xs = s / 2 + 1;
    xshift /= -oridim;
    yshift /= -oridim;
    if (std::fabs(xshift) < 1e-6 && std::fabs(yshift) < 1e-6) { /* XMIPP_EQUAL_ACCURACY */
        std::copy(in, in + (size_t)s * xs, out);
        return;
    }
    for (int i = 0; i < s; i++) {
        double y = (i < xs) ? i : i - s;
        for (int j = 0; j < xs; j++) {
            double dotp = 2 * PI * (j * xshift + y * yshift);
            double a = std::cos(dotp), b = std::sin(dotp);
            double c = in[i * xs + j].real(), d = in[i * xs + j].imag();
            double ac = a * c, bd = b * d, ab_cd = (a + b) * (c + d);
            out[i * xs + j] = cplx(ac - bd, ab_cd - ac - bd);
        }
    }
}

/* The sorted-threshold rule of convertAllSquaredDifferencesToWeights (ml_optimiser.cpp:7913-8031):
 * positive weights sorted ascending, accumulated from the top until the fraction is exceeded. */
static void significance_threshold(std::vector<double> w, double sum_weight, double fraction,
                                   long max_sig, bool count_sig, double* sig_weight, long* nr_sig)
{
    std::sort(w.begin(), w.end());
    double frac = 0.;
    double my_sig = 0.;
    long nsig = 0;
    for (long i = (long)w.size() - 1; i >= 0; i--) {
        if (max_sig > 0) {
            if (nsig < max_sig) {
                if (count_sig) nsig++;
                my_sig = w[i];
            }
        } else {
            if (count_sig) nsig++;
            my_sig = w[i];
        }
        frac += w[i];
        if (frac > fraction * sum_weight) break;
    }
    *sig_weight = my_sig;
    *nr_sig = nsig;
}

extern "C" int orc_weights_from_diff2(long n, const double* diff2, const double* prior,
                                      double min_diff2, double fraction, long max_sig,
                                      int count_sig, double* weights, double* sum_weight,
                                      double* sig_weight, long* nr_sig)
{
    double sum = 0;
    std::vector<double> pos;
    for (long i = 0; i < n; i++) {
        double w;
        if (diff2[i] < 0.) w = 0.;
        else {
            double d = diff2[i] - min_diff2;
            w = (d > 700.) ? 0. : prior[i] * std::exp(-d);
        }
        weights[i] = w;
        sum += w;
        if (w > 0.) pos.push_back(w);
    }
    *sum_weight = sum;
    significance_threshold(pos, sum, fraction, max_sig, count_sig != 0, sig_weight, nr_sig);
    return T2X_OK;
}

/* =====================================================================================
 * The E-step of one particle: expectationOneParticle (ml_optimiser.cpp:4252-4608) with
 * getAllSquaredDifferences (:6922-7507), convertAllSquaredDifferencesToWeights (:7510-8041)
 * and storeWeightedSums (:8043-8992).
 * ===================================================================================== */
struct Shared {
    const t2x_estep_config* cfg;
    const t2x_model* model;
    const orc_volume* proj;
    orc_volume* bp;
    Sampling S;
    std::vector<int> mres_coarse, mres_fine;
    std::vector<std::mutex>* class_mutex;
    std::mutex* global_mutex;
    bool clean;
};

struct FineBlock { std::vector<double> v; };   /* the nover fine samples of one coarse sample */

static int estep_particle(Shared& sh, const float* image, const t2x_particle_in& pin,
                          t2x_particle_out& out, t2x_wsum* wsum, t2x_estep_debug* dbg, int ipart)
{
    const t2x_estep_config* cfg = sh.cfg;
    const Sampling& S = sh.S;
    const int K = cfg->nr_classes;
    const long Nd = S.nr_dir(), Npsi = S.nr_psi(), Nt = S.nr_trans();
    const long Norient = Nd * Npsi;
    const long Nhid = (long)K * Norient * Nt;
    const int AO = cfg->adaptive_oversampling;
    const int nr_passes = AO > 0 ? 2 : 1;
    const double pix = cfg->pixel_size;
    const int N = cfg->ori_size;
    const int nsig2 = N / 2 + 1;
    /* the particle's optics group selects its noise spectrum (ml_optimiser.cpp:6703-6720) and
     * its image group the scale of the references (:7119-7125, 8326-8350) */
    const int n_og = cfg->nr_optics_groups > 0 ? cfg->nr_optics_groups : 1;
    const int n_grp = cfg->nr_groups > 0 ? cfg->nr_groups : 1;
    if (pin.optics_group < 0 || pin.optics_group >= n_og || pin.group_id < 0 || pin.group_id >= n_grp)
        return T2X_EINVAL;
    const double* sigma2 = sh.model->sigma2_noise + (size_t)pin.optics_group * nsig2;
    double scale_d2 = 1., scale_st = 1., scale_raw = 1.;
    if (cfg->do_scale_correction) {
        scale_raw = sh.model->scale_correction[pin.group_id];
        if (!(scale_raw > 0.) || scale_raw > 10000.) return T2X_EINVAL;
        scale_d2 = scale_raw;
        scale_st = std::max(0.001, scale_raw);   /* the store clamps small scales (:8333-8343) */
    }

    Prep P;
    preprocess(cfg, image, pin, sigma2, P);
    /* the offset prior per class: the particle's own (3D references), or for 2D references the
     * class's prior_offset_class (ml_optimiser.cpp:7627-7631, 7678-7682, 8283-8287) */
    std::vector<double> qx(sh.cfg->nr_classes, P.prior_x), qy(sh.cfg->nr_classes, P.prior_y);
    if (cfg->ref_dim == 2)
        for (int c = 0; c < cfg->nr_classes; c++) {
            qx[c] = sh.model->prior_offset_class ? sh.model->prior_offset_class[2 * c] : 0.;
            qy[c] = sh.model->prior_offset_class ? sh.model->prior_offset_class[2 * c + 1] : 0.;
        }

    /* orientation prior: NOPRIOR pdf_direction[class][idir]; local searches the product of the
     * particle's direction and psi priors (ml_optimiser.cpp:7031-7045, 7700-7713) */
    const bool local = cfg->orientational_prior_mode == T2X_PRIOR_ROTTILT_PSI;
    std::vector<double> dpr, ppr;
    long nd_loc = 0, np_loc = 0;
    if (local) local_priors(S, cfg, pin, dpr, ppr, &nd_loc, &np_loc);
    auto pdf_orient = [&](int c, long idir, long ipsi) {
        return local ? dpr[idir] * ppr[ipsi] : sh.model->pdf_direction[(size_t)c * Nd + idir];
    };
    const double sig2off = cfg->sigma2_offset;
    auto pdf_offset = [&](int c, long itrans) {
        double ox = P.old_x + S.tx[itrans], oy = P.old_y + S.ty[itrans];
        double tdiff2 = (ox - qx[c]) * (ox - qx[c]) + (oy - qy[c]) * (oy - qy[c]);
        tdiff2 *= pix * pix;
        if (sig2off < 0.0001) return (tdiff2 > 0.) ? 0. : 1.;
        return std::exp(tdiff2 / (-2. * sig2off)) / (2. * PI * sig2off);
    };

    std::vector<double> Mw_coarse(Nhid, -999.);
    std::vector<char> Msig(Nhid, 0);
    std::unordered_map<long, int> fine_index; /* coarse ihidden -> FineBlock */
    std::vector<FineBlock> fine_blocks;
    const int over_r = S.over_rot(AO), over_t = S.over_trans(AO);
    const int nover = over_r * over_t;
    double min_diff2 = 0, sum_weight = 0, sig_weight = 0;
    long nr_sig_coarse = 0, nr_sig_samples = 0;
    std::vector<double> fine_dbg_d2;
    std::vector<long> fine_dbg_h;

    for (int ipass = 0; ipass < nr_passes; ipass++) {
        const int os = (ipass == 0) ? 0 : AO;
        const int s = (AO > 0) ? (ipass == 0 ? cfg->coarse_size : cfg->current_size) : cfg->current_size;
        const int xs = s / 2 + 1;
        const size_t np = (size_t)s * xs;
        /* precalculateShiftedImagesCtfsAndInvSigma2s (ml_optimiser.cpp:6594-6894) */
        std::vector<cplx> Fimg(np);
        window_ft(P.Fimg.data(), P.cs, Fimg.data(), s);
        std::vector<double> Fctf(np);
        window_ft(P.Fctf.data(), P.cs, Fctf.data(), s);
        const std::vector<int>& mres = (s == cfg->coarse_size) ? sh.mres_coarse : sh.mres_fine;
        std::vector<double> Minv(np, 0.);
        for (size_t n = 0; n < np; n++) {
            int ires = mres[n];
            if (ires > 0 && ires < nsig2) Minv[n] = 1. / (cfg->sigma2_fudge * sigma2[ires]);
        }
        std::vector<std::vector<cplx>> Fshift((size_t)Nt * (os ? over_t : 1));
        {
            std::vector<double> ox, oy;
            for (long it = 0; it < Nt; it++) {
                S.translations_px(it, os, pix, ox, oy);
                for (size_t k = 0; k < ox.size(); k++) {
                    std::vector<cplx>& f = Fshift[it * ox.size() + k];
                    f.resize(np);
                    shift_image(Fimg.data(), s, (double)N, ox[k], oy[k], f.data());
                }
            }
        }
        const int nov_t = os ? over_t : 1, nov_r = os ? over_r : 1;
        double mymin = 99e99;
        /* the diff2 of every (oversampled) sample of one (class, direction, psi) */
        auto score = [&](int c, long idir, long ipsi, std::vector<cplx>& Fref, std::vector<cplx>& Frefctf,
                         double& mn) {
            std::vector<double> orot, otilt, opsi;
            double A[9];
            long iorientclass = (long)c * Norient + idir * Npsi + ipsi;
            double pdf_o = pdf_orient(c, idir, ipsi);
            bool proceed = true;
            if (ipass > 0) {
                proceed = false;
                for (long it = 0; it < Nt; it++)
                    if (Msig[iorientclass * Nt + it]) { proceed = true; break; }
            }
            if (!(proceed && pdf_o > 0.)) return;
            S.orientations(idir, ipsi, os, orot, otilt, opsi);
            for (int ir = 0; ir < nov_r; ir++) {
                orc_euler_angles2matrix(orot[ir], otilt[ir], opsi[ir], A);
                if (sh.clean) std::fill(Fref.begin(), Fref.end(), cplx(0, 0));
                project_into(&sh.proj[c], A, s, Fref.data(), sh.clean);
                for (size_t n = 0; n < np; n++) Frefctf[n] = Fref[n] * Fctf[n];
                if (cfg->do_scale_correction)
                    for (size_t n = 0; n < np; n++) Frefctf[n] *= scale_d2;
                long ihidden = iorientclass * Nt;
                for (long it = 0; it < Nt; it++, ihidden++) {
                    if (ipass > 0 && !Msig[ihidden]) continue;
                    for (int iot = 0; iot < nov_t; iot++) {
                        const cplx* Fs = Fshift[it * nov_t + iot].data();
                        double diff2 = P.highres_Xi2 / 2.;
                        for (size_t n = 0; n < np; n++) {
                            double dr = Frefctf[n].real() - Fs[n].real();
                            double di = Frefctf[n].imag() - Fs[n].imag();
                            diff2 += (dr * dr + di * di) * 0.5 * Minv[n];
                        }
                        if (ipass == 0) {
                            Mw_coarse[ihidden] = diff2;
                        } else {
                            auto itb = fine_index.find(ihidden);
                            int bi;
                            if (itb == fine_index.end()) {
                                bi = (int)fine_blocks.size();
                                fine_index[ihidden] = bi;
                                FineBlock fb;
                                fb.v.assign((size_t)nover, -999.);
                                fine_blocks.push_back(fb);
                            } else bi = itb->second;
                            fine_blocks[bi].v[ir * over_t + iot] = diff2;
                            fine_dbg_d2.push_back(diff2);
                            fine_dbg_h.push_back(ihidden * nover + ir * over_t + iot);
                        }
                        if (diff2 < mn) mn = diff2;
                    }
                }
            }
        };
        {
            std::vector<cplx> Fref(np, 0.), Frefctf(np);
            for (int c = 0; c < K; c++) {
                if (!(sh.model->pdf_class[c] > 0.)) continue;
                for (long idir = 0; idir < Nd; idir++)
                    for (long ipsi = 0; ipsi < Npsi; ipsi++) score(c, idir, ipsi, Fref, Frefctf, mymin);
            }
        }
        min_diff2 = mymin;
        if (ipass == 0 && dbg && ipart < dbg->n_debug && dbg->coarse_diff2)
            for (long h = 0; h < Nhid; h++) dbg->coarse_diff2[(size_t)ipart * Nhid + h] = (float)Mw_coarse[h];

        /* convertAllSquaredDifferencesToWeights (ml_optimiser.cpp:7510-8041); the means run
         * over the particle's orientations (all of them, or the local ones) and translations */
        double pdf_orientation_mean = 0, pdf_offset_mean = 0;
        long cnt_o = 0, cnt_t = 0;
        for (int c = 0; c < K; c++) {
            for (long idir = 0; idir < Nd; idir++) {
                if (local && !(dpr[idir] > 0.)) continue;
                for (long ipsi = 0; ipsi < Npsi; ipsi++) {
                    if (local && !(ppr[ipsi] > 0.)) continue;
                    pdf_orientation_mean += pdf_orient(c, idir, ipsi);
                    cnt_o++;
                }
            }
            for (long it = 0; it < Nt; it++) { pdf_offset_mean += pdf_offset(c, it); cnt_t++; }
        }
        pdf_orientation_mean /= (double)cnt_o;
        pdf_offset_mean /= (double)cnt_t;
        std::vector<double> pdf_t((size_t)K * Nt);   /* [class][translation] */
        for (int c = 0; c < K; c++)
            for (long it = 0; it < Nt; it++) {
                double v = pdf_offset(c, it);
                if (pdf_offset_mean > 0.) v /= pdf_offset_mean;
                pdf_t[(size_t)c * Nt + it] = v;
            }
        double sumw = 0;
        std::vector<double> positive;
        auto weigh = [&](double d2, double prior) {
            if (d2 < 0.) return 0.;
            double w = prior;
            double d = d2 - min_diff2;
            if (d > 700.) w = 0.;
            else w *= std::exp(-d);
            return w;
        };
        for (int c = 0; c < K; c++)
            for (long idir = 0; idir < Nd; idir++)
                for (long ipsi = 0; ipsi < Npsi; ipsi++) {
                    long iorientclass = (long)c * Norient + idir * Npsi + ipsi;
                    double pdf_o = pdf_orient(c, idir, ipsi);
                    if (pdf_orientation_mean != 0.) pdf_o /= pdf_orientation_mean;
                    for (long it = 0; it < Nt; it++) {
                        long ihidden = iorientclass * Nt + it;
                        double prior = pdf_o * pdf_t[(size_t)c * Nt + it];
                        if (ipass == 0) {
                            double w = weigh(Mw_coarse[ihidden], prior);
                            Mw_coarse[ihidden] = w;
                            sumw += w;
                            if (w > 0.) positive.push_back(w);
                        } else {
                            auto itb = fine_index.find(ihidden);
                            if (itb == fine_index.end()) continue;
                            FineBlock& fb = fine_blocks[itb->second];
                            for (int q = 0; q < nover; q++) {
                                double w = weigh(fb.v[q], prior);
                                fb.v[q] = w;
                                sumw += w;
                                if (w > 0.) positive.push_back(w);
                            }
                        }
                    }
                }
        sum_weight = sumw;
        if (sumw == 0. || std::isnan(sumw)) return T2X_EZEROSUM;
        long nsig = 0;
        significance_threshold(positive, sumw, cfg->adaptive_fraction, cfg->maximum_significants,
                               ipass == 0, &sig_weight, &nsig);
        if (ipass == 0) {
            if (nsig == 0) return T2X_ENOSIG;
            nr_sig_coarse = nsig;
            for (long h = 0; h < Nhid; h++) {
                Msig[h] = (Mw_coarse[h] >= sig_weight) ? 1 : 0;
                nr_sig_samples += Msig[h];
            }
            if (dbg && ipart < dbg->n_debug) {
                if (dbg->coarse_significant)
                    for (long h = 0; h < Nhid; h++) dbg->coarse_significant[(size_t)ipart * Nhid + h] = Msig[h];
                if (dbg->coarse_min_diff2) dbg->coarse_min_diff2[ipart] = min_diff2;
                if (dbg->coarse_sum_weight) dbg->coarse_sum_weight[ipart] = sumw;
                if (dbg->coarse_sig_weight) dbg->coarse_sig_weight[ipart] = sig_weight;
            }
        }
    }
    if (dbg && ipart < dbg->n_debug && dbg->fine_count) {
        long nf = (long)fine_dbg_d2.size();
        dbg->fine_count[ipart] = nf;
        for (long q = 0; q < nf && q < dbg->fine_capacity; q++) {
            if (dbg->fine_diff2) dbg->fine_diff2[(size_t)ipart * dbg->fine_capacity + q] = (float)fine_dbg_d2[q];
            if (dbg->fine_hidden) dbg->fine_hidden[(size_t)ipart * dbg->fine_capacity + q] = fine_dbg_h[q];
            if (dbg->fine_weight) {   /* the fine pass's weight of the sample (priors included) */
                const long h = fine_dbg_h[q];
                const auto itb = fine_index.find(h / nover);
                dbg->fine_weight[(size_t)ipart * dbg->fine_capacity + q] =
                    itb == fine_index.end() ? 0. : fine_blocks[itb->second].v[h % nover];
            }
        }
    }

    /* storeWeightedSums (ml_optimiser.cpp:8043-8992) */
    const int s = cfg->current_size, xs = s / 2 + 1;
    const size_t np = (size_t)s * xs;
    const int os = AO;
    const int nov_r = os ? over_r : 1, nov_t = os ? over_t : 1;
    std::vector<double> Minv(np, 0.);
    for (size_t n = 0; n < np; n++) {
        int ires = sh.mres_fine[n];
        if (ires > 0 && ires < nsig2) Minv[n] = 1. / (cfg->sigma2_fudge * sigma2[ires]);
    }
    Minv[0] = 1. / (cfg->sigma2_fudge * sigma2[0]);
    std::vector<std::vector<cplx>> Fsh((size_t)Nt * nov_t), Fsh_nm((size_t)Nt * nov_t);
    std::vector<double> trx((size_t)Nt * nov_t), try_((size_t)Nt * nov_t);
    {
        std::vector<double> ox, oy;
        for (long it = 0; it < Nt; it++) {
            S.translations_px(it, os, pix, ox, oy);
            for (int k = 0; k < nov_t; k++) {
                size_t q = it * nov_t + k;
                trx[q] = ox[k]; try_[q] = oy[k];
                Fsh[q].resize(np);
                Fsh_nm[q].resize(np);
                shift_image(P.Fimg.data(), s, (double)N, ox[k], oy[k], Fsh[q].data());
                shift_image(P.Fimg_nomask.data(), s, (double)N, ox[k], oy[k], Fsh_nm[q].data());
            }
        }
    }
    std::vector<double> thr_sigma2(nsig2, 0.);
    std::vector<double> thr_pdf_dir((size_t)K * Nd, 0.), thr_pdf_class(K, 0.);
    double thr_sumw_group = 0, thr_sigma2_offset = 0, wsum_norm = 0;
    double wsum_XA = 0, wsum_AA = 0;   /* scale correction (ml_optimiser.cpp:8477-8486) */
    const double* dvp = sh.model->data_vs_prior_class;
    std::vector<double> thr_prior_off((size_t)2 * K, 0.);   /* 2D references (ml_optimiser.cpp:8504-8508) */
    double max_weight = -1.;
    long best_hidden = -1;
    std::vector<cplx> Fref(np, 0.), Frefctf(np), Facc(np);
    std::vector<double> Fw(np);
    std::vector<double> orot, otilt, opsi;
    double A[9];
    long nfine_sig = 0;
    auto fine_weight = [&](long ihidden, int q) -> double {
        if (AO == 0) return Mw_coarse[ihidden];
        auto itb = fine_index.find(ihidden);
        if (itb == fine_index.end()) return 0.;
        double w = fine_blocks[itb->second].v[q];
        return w < 0 ? 0. : w;
    };
    for (int c = 0; c < K; c++)
        for (long idir = 0; idir < Nd; idir++)
            for (long ipsi = 0; ipsi < Npsi; ipsi++) {
                long iorientclass = (long)c * Norient + idir * Npsi + ipsi;
                bool any = false;
                for (long it = 0; it < Nt; it++)
                    if (Msig[iorientclass * Nt + it]) { any = true; break; }
                if (!any) continue;
                S.orientations(idir, ipsi, os, orot, otilt, opsi);
                for (int ir = 0; ir < nov_r; ir++) {
                    orc_euler_angles2matrix(orot[ir], otilt[ir], opsi[ir], A);
                    if (sh.clean) std::fill(Fref.begin(), Fref.end(), cplx(0, 0));
                    project_into(&sh.proj[c], A, s, Fref.data(), sh.clean);
                    /* Frefctf and the CTF of the backprojection both carry the group's scale */
                    for (size_t n = 0; n < np; n++) Frefctf[n] = Fref[n] * (P.Fctf[n] * scale_st);
                    std::fill(Facc.begin(), Facc.end(), cplx(0, 0));
                    std::fill(Fw.begin(), Fw.end(), 0.);
                    long ihidden = iorientclass * Nt;
                    for (long it = 0; it < Nt; it++, ihidden++)
                        for (int iot = 0; iot < nov_t; iot++) {
                            double weight = fine_weight(ihidden, ir * nov_t + iot);
                            if (!(weight >= sig_weight)) continue;
                            nfine_sig++;
                            weight /= sum_weight;
                            const size_t q = it * nov_t + iot;
                            const cplx* Fs = Fsh[q].data();
                            for (size_t n = 0; n < np; n++) {
                                int ires = sh.mres_fine[n];
                                if (ires > -1) {
                                    double dr = Frefctf[n].real() - Fs[n].real();
                                    double di = Frefctf[n].imag() - Fs[n].imag();
                                    double wd = weight * (dr * dr + di * di);
                                    thr_sigma2[ires] += wd;
                                    wsum_norm += wd;
                                    if (cfg->do_scale_correction && dvp && dvp[(size_t)c * nsig2 + ires] > 3.) {
                                        wsum_XA += weight * (Frefctf[n].real() * Fs[n].real() +
                                                             Frefctf[n].imag() * Fs[n].imag());
                                        wsum_AA += weight * std::norm(Frefctf[n]);
                                    }
                                }
                            }
                            thr_sumw_group += weight;
                            thr_pdf_class[c] += weight;
                            if (cfg->ref_dim == 2) {
                                thr_prior_off[2 * c] += weight * pix * (P.old_x + trx[q]);
                                thr_prior_off[2 * c + 1] += weight * pix * (P.old_y + try_[q]);
                            }
                            double dx = qx[c] - P.old_x - trx[q];
                            thr_sigma2_offset += weight * pix * pix * dx * dx;
                            double dy = qy[c] - P.old_y - try_[q];
                            thr_sigma2_offset += weight * pix * pix * dy * dy;
                            thr_pdf_dir[(size_t)c * Nd + idir] += weight;
                            const cplx* Fn = Fsh_nm[q].data();
                            for (size_t n = 0; n < np; n++) {
                                double myctf = P.Fctf[n] * scale_st;
                                double wxs = weight * myctf * Minv[n];
                                Facc[n] += Fn[n] * wxs;
                                Fw[n] += wxs * myctf;
                            }
                            if (weight > max_weight) {
                                max_weight = weight;
                                out.rot = orot[ir]; out.tilt = otilt[ir]; out.psi = opsi[ir];
                                out.xoff = P.old_x + trx[q];
                                out.yoff = P.old_y + try_[q];
                                out.iclass = c + 1;
                                out.pmax = max_weight;
                                best_hidden = ihidden * (long)(nov_r * nov_t) + ir * nov_t + iot;
                            }
                        }
                    {
                        std::lock_guard<std::mutex> lk((*sh.class_mutex)[c]);
                        if (sh.bp[c].ref_dim == 2) backrotate_from(&sh.bp[c], Facc.data(), Fw.data(), s, A);
                        else backproject_from(&sh.bp[c], Facc.data(), Fw.data(), s, A);
                    }
                }
            }
    for (int ires = s / 2 + 1; ires < nsig2; ires++) {
        thr_sigma2[ires] += P.power[ires];
        wsum_norm += P.power[ires];
    }
    double normcorr = pin.norm_correction;
    if (cfg->do_norm_correction) {
        double old = pin.norm_correction / cfg->avg_norm_correction;
        normcorr = old * std::sqrt(wsum_norm * 2.);
    }
    double logsigma2 = 0.;
    for (size_t n = 0; n < np; n++) {
        int ires = sh.mres_fine[n];
        if (ires > 0 && ires < nsig2) logsigma2 += std::log(2. * PI * sigma2[ires]);
    }
    double dLL = std::log(sum_weight) - min_diff2 - logsigma2;
    out.dLL = dLL;
    out.norm_correction = normcorr;
    out.sum_weight = sum_weight;
    out.min_diff2 = min_diff2;
    out.significant_weight = sig_weight;
    out.nr_significant_coarse = nr_sig_coarse;
    /* samples whose weights the sums below read: the fine pass's, or with one pass the
     * significant coarse samples */
    out.nr_fine_samples = (AO > 0) ? (long)fine_dbg_d2.size() : nr_sig_samples;
    out.nr_fine_significant = nfine_sig;
    out.best_hidden = best_hidden;
    if (wsum) {
        std::lock_guard<std::mutex> lk(*sh.global_mutex);
        const size_t og = (size_t)pin.optics_group;
        if (wsum->sigma2_noise)
            for (int i = 0; i < nsig2; i++) wsum->sigma2_noise[og * nsig2 + i] += thr_sigma2[i];
        if (wsum->sumw_group) wsum->sumw_group[og] += thr_sumw_group;
        if (cfg->do_scale_correction) {
            /* XA / scale and AA / scale^2 with the group's unclamped scale (:8855-8862) */
            if (wsum->wsum_signal_product) wsum->wsum_signal_product[pin.group_id] += wsum_XA / scale_raw;
            if (wsum->wsum_reference_power)
                wsum->wsum_reference_power[pin.group_id] += wsum_AA / (scale_raw * scale_raw);
        }
        if (wsum->pdf_class)
            for (int c = 0; c < K; c++) wsum->pdf_class[c] += thr_pdf_class[c];
        if (wsum->pdf_direction)
            for (size_t i = 0; i < thr_pdf_dir.size(); i++) wsum->pdf_direction[i] += thr_pdf_dir[i];
        if (wsum->sigma2_offset) wsum->sigma2_offset[0] += thr_sigma2_offset;
        if (wsum->prior_offset_class && cfg->ref_dim == 2)
            for (int c = 0; c < 2 * K; c++) wsum->prior_offset_class[c] += thr_prior_off[c];
        if (wsum->avg_norm_correction && cfg->do_norm_correction) wsum->avg_norm_correction[0] += normcorr;
        if (wsum->LL) wsum->LL[0] += dLL;
        if (wsum->ave_Pmax) wsum->ave_Pmax[0] += out.pmax;
    }
    return T2X_OK;
}

extern "C" int orc_estep(const t2x_estep_config* cfg, const t2x_model* model,
                         const orc_volume* proj, orc_volume* bp, int n, const float* images,
                         const t2x_particle_in* pin, t2x_particle_out* out, t2x_wsum* wsum,
                         t2x_estep_debug* dbg, int proj_mode, int nthreads)
{
    if (check_config(cfg) != T2X_OK || !model || !proj || !bp || n < 0) return T2X_EINVAL;
    if (cfg->do_scale_correction && !model->scale_correction) return T2X_EINVAL;
    for (int c = 0; c < cfg->nr_classes; c++)
        if (proj[c].ref_dim != cfg->ref_dim || bp[c].ref_dim != cfg->ref_dim) return T2X_EINVAL;
    Shared sh;
    sh.cfg = cfg;
    sh.model = model;
    sh.proj = proj;
    sh.bp = bp;
    sh.S.setup(cfg);
    sh.mres_coarse = mresol(cfg->coarse_size);
    sh.mres_fine = mresol(cfg->current_size);
    std::vector<std::mutex> cm(cfg->nr_classes);
    std::mutex gm;
    sh.class_mutex = &cm;
    sh.global_mutex = &gm;
    sh.clean = (proj_mode == ORC_PROJ_CLEAN);
    int status = T2X_OK;
    const size_t NN = (size_t)cfg->ori_size * cfg->ori_size;
    auto one = [&](int p) {
        std::memset(&out[p], 0, sizeof(t2x_particle_out));
        int st = estep_particle(sh, images + p * NN, pin[p], out[p], wsum, dbg, p);
        if (st != T2X_OK) {
#ifdef _OPENMP
#pragma omp critical
#endif
            status = st;
        }
    };
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
    for (int p = 0; p < n; p++) one(p);
    return status;
}
