// t2x_sampling.cpp — orientation and translation grids of the E-step (host, double precision).
//
// HEALPix NEST pixel centres follow the published HEALPix scheme (Gorski et al. 2005) as used by
// HealpixSampling (src/healpix_sampling.cpp:534-608, 1642-1649, 1914-2050) over
// Healpix_Base (src/Healpix_2.15a/healpix_base.cc); translations follow
// HealpixSampling::setTranslations / getTranslationsInPixel (healpix_sampling.cpp:290-440, 1800-1912).
#include <cmath>

#include "t2x_internal.h"

namespace t2x {

static const double kPi = 3.14159265358979323846;

namespace {

// Face coordinates <-> NEST index: the NEST index interleaves the bits of (ix, iy) within a face.
long interleave(long v)
{
    long r = 0;
    for (int b = 0; b < 31; b++)
        if (v & (1L << b)) r |= 1L << (2 * b);
    return r;
}
long deinterleave(long v)
{
    long r = 0;
    for (int b = 0; b < 31; b++)
        if (v & (1L << (2 * b))) r |= 1L << b;
    return r;
}

void nest_to_face(int order, long pix, long* ix, long* iy, int* face)
{
    *face = (int)(pix >> (2 * order));
    long local = pix & ((1L << (2 * order)) - 1);
    *ix = deinterleave(local);
    *iy = deinterleave(local >> 1);
}

long face_to_nest(int order, long ix, long iy, int face)
{
    return ((long)face << (2 * order)) + interleave(ix) + (interleave(iy) << 1);
}

// Pixel centre (z = cos(theta), phi) of NEST pixel `pix` at resolution 2^order.
void pixel_centre(int order, long pix, double* z, double* phi)
{
    static const int ring_of_face[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
    static const int phase_of_face[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
    const long ns = 1L << order;
    const double f2 = 4.0 / (double)(12 * ns * ns);
    long ix, iy;
    int face;
    nest_to_face(order, pix, &ix, &iy, &face);
    long ring = ((long)ring_of_face[face] << order) - ix - iy - 1;   // 1 .. 4ns-1
    long nr, shift;
    if (ring < ns) {                 // north polar cap
        nr = ring;
        *z = 1.0 - (double)(nr * nr) * f2;
        shift = 0;
    } else if (ring > 3 * ns) {      // south polar cap
        nr = 4 * ns - ring;
        *z = (double)(nr * nr) * f2 - 1.0;
        shift = 0;
    } else {                         // equatorial belt
        nr = ns;
        const double f1 = (double)(2 * ns) * f2;
        *z = (double)(2 * ns - ring) * f1;
        shift = (ring - ns) & 1;
    }
    long jp = (phase_of_face[face] * nr + ix - iy + 1 + shift) / 2;
    if (jp > 4 * ns) jp -= 4 * ns;
    if (jp < 1) jp += 4 * ns;
    *phi = ((double)jp - 0.5 * (double)(shift + 1)) * (0.5 * kPi / (double)nr);
}

void normalise_direction(double& r, double& t)
{
    if (t < 0.) { t = -t; r += 180.; }
    while (r > 180. || r < -180.) r += (r > 180.) ? -360. : 360.;
}

int ceil_relion(double x) { return (x == (int)x) ? (int)x : ((x > 0) ? (int)(x + 1) : (int)x); }

}  // namespace

int check_config(const t2x_estep_config* c)
{
    if (!c) { set_error("null configuration"); return T2X_EINVAL; }
    if (c->ori_size < 4 || (c->ori_size & 1)) { set_error("ori_size must be even and >= 4"); return T2X_EINVAL; }
    if (c->current_size < 2 || (c->current_size & 1) || c->current_size > c->ori_size) {
        set_error("current_size must be even and <= ori_size");
        return T2X_EINVAL;
    }
    if (c->coarse_size < 2 || (c->coarse_size & 1) || c->coarse_size > c->current_size) {
        set_error("coarse_size must be even and <= current_size");
        return T2X_EINVAL;
    }
    if (c->ref_dim != 3 && c->ref_dim != 2) { set_error("ref_dim must be 3 or 2"); return T2X_EINVAL; }
    if (c->nr_classes < 1) { set_error("nr_classes < 1"); return T2X_EINVAL; }
    if (c->adaptive_oversampling < 0 || c->adaptive_oversampling > 2) {
        set_error("adaptive_oversampling must be 0 (one pass), 1 or 2");
        return T2X_EINVAL;
    }
    if (c->offset_step <= 0 || c->offset_range < 0 || c->psi_step <= 0) {
        set_error("bad offset or psi sampling");
        return T2X_EINVAL;
    }
    if (c->orientational_prior_mode == T2X_PRIOR_ROTTILT_PSI) {
        if (!(c->sigma_rot >= 0.) || !(c->sigma_tilt >= 0.) || !(c->sigma_psi >= 0.)) {
            set_error("local searches: negative angular sigma");
            return T2X_EINVAL;
        }
        if ((c->sigma_rot > 0.) != (c->sigma_tilt > 0.)) {
            set_error("local searches: sigma_rot and sigma_tilt must both be set or both be 0 "
                      "(rot-only and tilt-only priors are not supported)");
            return T2X_EINVAL;
        }
    } else if (c->orientational_prior_mode != T2X_NOPRIOR) {
        set_error("orientational_prior_mode must be T2X_NOPRIOR or T2X_PRIOR_ROTTILT_PSI");
        return T2X_EINVAL;
    }
    if (c->precision != T2X_PREC_FP32 && c->precision != T2X_PREC_BF16X3 && c->precision != T2X_PREC_FP16X3 &&
        c->precision != T2X_PREC_FP16X2) {
        set_error("precision mode not supported");
        return T2X_EINVAL;
    }
    if (c->ref_dim == 3 && (c->healpix_order < 0 || c->healpix_order > 8)) {
        set_error("healpix_order out of range");
        return T2X_EINVAL;
    }
    return T2X_OK;
}

int build_sampling(const t2x_estep_config* cfg, HostSampling& s)
{
    T2X_TRY(check_config(cfg));
    const bool is3d = cfg->ref_dim == 3;
    const int order = cfg->healpix_order;
    const long npix = is3d ? (12L << (2 * order)) : 1;
    s = HostSampling();
    if (is3d) {
        for (long p = 0; p < npix; p++) {
            double z, phi;
            pixel_centre(order, p, &z, &phi);
            double r = phi * 180.0 / kPi, t = std::acos(z) * 180.0 / kPi;
            normalise_direction(r, t);
            s.rot.push_back(r);
            s.tilt.push_back(t);
            s.ipix.push_back(p);
        }
    } else {   // 2D references: in-plane rotations only (healpix_sampling.cpp:590-595)
        s.rot.push_back(0.);
        s.tilt.push_back(0.);
        s.ipix.push_back(-1);
    }
    s.nr_dir = npix;
    int nr_psi = ceil_relion(360. / cfg->psi_step);
    double psi_step = 360. / nr_psi;
    for (int i = 0; i < nr_psi; i++) s.psi.push_back(i * psi_step);
    s.nr_psi = nr_psi;
    const double step = cfg->offset_step, range = cfg->offset_range;
    int maxp = ceil_relion(range / step);
    for (long ix = -maxp; ix <= maxp; ix++) {
        double xo = ix * step;
        for (long iy = -maxp; iy <= maxp; iy++) {
            double yo = iy * step;
            if (xo * xo + yo * yo < range * range + 0.001) {
                s.tx.push_back(xo);
                s.ty.push_back(yo);
            }
        }
    }
    s.nr_trans = (long)s.tx.size();
    const int os = cfg->adaptive_oversampling;
    // oversamplingFactorOrientations (healpix_sampling.cpp:1720-1735): 2D references oversample psi only
    s.over_rot = is3d ? 1 << (3 * os) : 1 << os;
    s.over_trans = 1 << (2 * os);
    s.n_subdir = is3d ? 1 << (2 * os) : 1;
    s.n_subpsi = 1 << os;
    // sub-directions: children of each coarse pixel at order+os, j (y) outer, i (x) inner
    const int fact = 1 << os;
    s.sub_rot.resize(npix * s.n_subdir);
    s.sub_tilt.resize(npix * s.n_subdir);
    if (!is3d) s.sub_rot[0] = s.sub_tilt[0] = 0.;
    for (long d = 0; is3d && d < npix; d++) {
        long x, y;
        int face;
        nest_to_face(order, s.ipix[d], &x, &y, &face);
        int k = 0;
        for (long j = fact * y; j < fact * (y + 1); ++j)
            for (long i = fact * x; i < fact * (x + 1); ++i, ++k) {
                long child = face_to_nest(order + os, i, j, face);
                double z, phi;
                pixel_centre(order + os, child, &z, &phi);
                double r = phi * 180.0 / kPi, t = std::acos(z) * 180.0 / kPi;
                normalise_direction(r, t);
                s.sub_rot[d * s.n_subdir + k] = r;
                s.sub_tilt[d * s.n_subdir + k] = t;
            }
    }
    s.sub_psi.resize(nr_psi * s.n_subpsi);
    for (int i = 0; i < nr_psi; i++)
        for (int k = 0; k < s.n_subpsi; k++)
        {
            // pushbackOversampledPsiAngles (healpix_sampling.cpp:2023-2050): 2D wraps to (-180, 180]
            double over = s.psi[i] - 0.5 * psi_step + (0.5 + k) * psi_step / s.n_subpsi;
            if (!is3d && over > 180.) over -= 360.;
            s.sub_psi[i * s.n_subpsi + k] = over;
        }
    // sines and cosines of the fine angles (host libm, as the CPU path's Euler_angles2matrix)
    s.sub_dir_sc.resize(s.sub_rot.size() * 4);
    for (size_t k = 0; k < s.sub_rot.size(); k++) {
        double a = s.sub_rot[k] * kPi / 180, b = s.sub_tilt[k] * kPi / 180;
        s.sub_dir_sc[4 * k + 0] = std::sin(a);
        s.sub_dir_sc[4 * k + 1] = std::cos(a);
        s.sub_dir_sc[4 * k + 2] = std::sin(b);
        s.sub_dir_sc[4 * k + 3] = std::cos(b);
    }
    s.sub_psi_sc.resize(s.sub_psi.size() * 2);
    for (size_t k = 0; k < s.sub_psi.size(); k++) {
        double g = s.sub_psi[k] * kPi / 180;
        s.sub_psi_sc[2 * k + 0] = std::sin(g);
        s.sub_psi_sc[2 * k + 1] = std::cos(g);
    }
    const double pix = cfg->pixel_size;
    for (long t = 0; t < s.nr_trans; t++) {
        s.coarse_tx_px.push_back(s.tx[t] / pix);
        s.coarse_ty_px.push_back(s.ty[t] / pix);
        for (int ox = 0; ox < fact; ox++) {
            double xo = s.tx[t] - 0.5 * step + (0.5 + ox) * step / fact;
            for (int oy = 0; oy < fact; oy++) {
                double yo = s.ty[t] - 0.5 * step + (0.5 + oy) * step / fact;
                s.fine_tx_px.push_back(xo / pix);
                s.fine_ty_px.push_back(yo / pix);
            }
        }
    }
    return T2X_OK;
}

}  // namespace t2x

extern "C" int t2x_sampling_info_get(const t2x_estep_config* cfg, t2x_sampling_info* info)
{
    if (!info) { t2x::set_error("null info"); return T2X_EINVAL; }
    t2x::HostSampling s;
    T2X_TRY(t2x::build_sampling(cfg, s));
    info->nr_dir = s.nr_dir;
    info->nr_psi = s.nr_psi;
    info->nr_trans = s.nr_trans;
    info->nr_over_rot = s.over_rot;
    info->nr_over_trans = s.over_trans;
    info->nr_coarse_per_class = s.nr_dir * s.nr_psi;
    info->nr_hidden_coarse = (long)cfg->nr_classes * s.nr_dir * s.nr_psi * s.nr_trans;
    return T2X_OK;
}

extern "C" int t2x_sampling_coarse_get(const t2x_estep_config* cfg, double* rot, double* tilt,
                                       double* psi, double* tx, double* ty)
{
    t2x::HostSampling s;
    T2X_TRY(t2x::build_sampling(cfg, s));
    for (long i = 0; i < s.nr_dir; i++) { if (rot) rot[i] = s.rot[i]; if (tilt) tilt[i] = s.tilt[i]; }
    for (long i = 0; i < s.nr_psi; i++) if (psi) psi[i] = s.psi[i];
    for (long i = 0; i < s.nr_trans; i++) { if (tx) tx[i] = s.tx[i]; if (ty) ty[i] = s.ty[i]; }
    return T2X_OK;
}

extern "C" int t2x_sampling_fine_get(const t2x_estep_config* cfg, double* sub_rot, double* sub_tilt,
                                     double* sub_psi, double* ftx_px, double* fty_px)
{
    t2x::HostSampling s;
    T2X_TRY(t2x::build_sampling(cfg, s));
    for (size_t i = 0; i < s.sub_rot.size(); i++) { if (sub_rot) sub_rot[i] = s.sub_rot[i]; if (sub_tilt) sub_tilt[i] = s.sub_tilt[i]; }
    for (size_t i = 0; i < s.sub_psi.size(); i++) if (sub_psi) sub_psi[i] = s.sub_psi[i];
    for (size_t i = 0; i < s.fine_tx_px.size(); i++) { if (ftx_px) ftx_px[i] = s.fine_tx_px[i]; if (fty_px) fty_px[i] = s.fine_ty_px[i]; }
    return T2X_OK;
}
