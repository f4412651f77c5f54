// t2x_recon.hip — BackProjector::reconstruct (default skip_gridding branch) and getFSC on the
// GPU, all arithmetic in FP64 (backprojector.cpp:1344-1770, :2486-2600; fftw.cpp:481-523).
//
// Data flow for one class: the projector-layout data/weight volumes are uploaded once; one
// kernel per FFTW-layout voxel of the padded transform decentres, adds the MAP term and
// accumulates the per-shell weight sums (LDS partials per block, then the blocks summed in a
// fixed order; the partials' LDS atomics may reorder additions within a block, so the floor
// can differ in the last bits between runs); a second kernel writes the 2*ori-sized half-complex grid directly —
// division by max(weight, floor), Fourier window, CenterFFTbySign and Hermitian symmetrisation of
// the x = 0 / Nyquist planes fused per output voxel — followed by one hipFFT Z2D, a real-space
// window with the FFT normalisation, the soft spherical mask and the trilinear gridding
// correction. Volumes: 512^3 doubles for a 256 box (1 GiB each) fit easily in HBM.
#include "t2x_internal.h"

#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace t2x {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kShellBlocks = 256;   // partial-sum blocks of the shell reductions

__host__ __device__ inline long long fftw_logical(long long k, long long xsize, long long zdim)
{
    return k < xsize ? k : k - zdim;
}
__host__ __device__ inline long long wrapi(long long k, long long n) { return k < 0 ? k + n : k; }

struct ReconGeom {
    long long pad, xh, P, Ph, ori;
    double pf;
    long long max_r2, round_max_r2, win_r2;
    int r_max, do_map, minres_map;
    double tau2_fudge, normfft;
};

// decenter + MAP term: weights of the pad^3 FFTW-layout grid, and per-shell partial sums
// (shell = FLOOR(|k| / pf) over |k|^2 < round_max_r2) of the final weights.
__global__ void k_recon_weights(ReconGeom g, const double* __restrict__ w_in, const double* __restrict__ tau2,
                                double* __restrict__ Fw, double* __restrict__ part_sum, double* __restrict__ part_cnt,
                                int* err)
{
    extern __shared__ double sh[];
    double* ssum = sh;
    double* scnt = sh + g.r_max;
    for (int s = threadIdx.x; s < g.r_max; s += blockDim.x) { ssum[s] = 0.; scnt[s] = 0.; }
    __syncthreads();
    const long long nv = g.pad * g.pad * g.xh;
    const long long c0 = g.pad / 2;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nv; e += (long long)gridDim.x * blockDim.x) {
        const long long j = e % g.xh, i = (e / g.xh) % g.pad, k = e / (g.xh * g.pad);
        const long long kp = fftw_logical(k, g.xh, g.pad), ip = fftw_logical(i, g.xh, g.pad), jp = j;
        const long long r2 = kp * kp + ip * ip + jp * jp;
        double w = 0.;
        if (r2 <= g.max_r2) w = w_in[((kp + c0) * g.pad + (ip + c0)) * g.xh + jp];
        if (g.do_map && r2 < g.max_r2) {
            const long long ires = llround(sqrt((double)r2) / g.pf);
            const double t = tau2[ires];
            double invtau2;
            if (t > 0.) invtau2 = 1. / (g.pf * g.pf * g.pf * g.tau2_fudge * t);
            else if (t < 1e-20) invtau2 = (w > 1e-20) ? 1. / (0.001 * w) : 0.;
            else { *err = 2; invtau2 = 0.; }
            if (ires >= g.minres_map) w = w + invtau2;
        }
        Fw[e] = w;
        if (r2 < g.round_max_r2) {
            const long long ires = (long long)floor(sqrt((double)r2) / g.pf);
            if (ires >= g.r_max) { *err = 3; continue; }
            atomicAdd(&ssum[ires], w);   // LDS partial: order within the block only
            atomicAdd(&scnt[ires], 1.);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < g.r_max; s += blockDim.x) {
        part_sum[(size_t)blockIdx.x * g.r_max + s] = ssum[s];
        part_cnt[(size_t)blockIdx.x * g.r_max + s] = scnt[s];
    }
}

// floor[s] = (sum over blocks in order) / (1000 * count)
__global__ void k_recon_floor(ReconGeom g, const double* part_sum, const double* part_cnt, int nblocks,
                              double* floorv, int* err)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.r_max) return;
    double a = 0., c = 0.;
    for (int b = 0; b < nblocks; b++) {
        a += part_sum[(size_t)b * g.r_max + s];
        c += part_cnt[(size_t)b * g.r_max + s];
    }
    if (c > 0. || a > 0.) floorv[s] = a / (1000. * c);
    else { floorv[s] = 0.; *err = 4; }
}

// data/weight at FFTW logical (kp, ip, jp) of the padded grid (0 outside it)
__device__ inline double2 recon_value(const ReconGeom& g, long long kp, long long ip, long long jp,
                                      const double* __restrict__ d_re, const double* __restrict__ d_im,
                                      const double* __restrict__ Fw, const double* __restrict__ floorv)
{
    const long long half = g.pad / 2;
    if (jp >= g.xh || kp > half || kp < -half || ip > half || ip < -half) return make_double2(0., 0.);
    const long long r2 = kp * kp + ip * ip + jp * jp;
    const long long e = (wrapi(kp, g.pad) * g.pad + wrapi(ip, g.pad)) * g.xh + jp;
    double re = 0., im = 0.;
    if (r2 <= g.max_r2) {
        const long long c0 = g.pad / 2;
        const long long src = ((kp + c0) * g.pad + (ip + c0)) * g.xh + jp;
        re = d_re[src];
        im = d_im[src];
    }
    const long long ires = (long long)floor(sqrt((double)r2) / g.pf);
    const double w = fmax(Fw[e], floorv[(ires < g.r_max) ? ires : (g.r_max - 1)]);
    if (w != 0.) { re /= w; im /= w; }
    return make_double2(re, im);
}

// 2*ori half-complex grid: window, CenterFFTbySign, Hermitian x = 0 / Nyquist planes
__global__ void k_recon_window(ReconGeom g, const double* __restrict__ d_re, const double* __restrict__ d_im,
                               const double* __restrict__ Fw, const double* __restrict__ floorv,
                               double2* __restrict__ W)
{
    const long long n = g.P * g.P * g.Ph;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const long long j = e % g.Ph, i = (e / g.Ph) % g.P, k = e / (g.Ph * g.P);
        const long long kp = fftw_logical(k, g.Ph, g.P), ip = fftw_logical(i, g.Ph, g.P), jp = j;
        auto at = [&](long long a, long long b, long long c) -> double2 {
            if (g.Ph > g.xh) {   // zero-padding window: keep |k|^2 <= (XSIZE(in)-1)^2
                if (a * a + b * b + c * c > g.win_r2) return make_double2(0., 0.);
            }
            return recon_value(g, a, b, c, d_re, d_im, Fw, floorv);
        };
        double2 v = at(kp, ip, jp);
        if (jp == 0 || jp == g.P / 2) {
            // partner (-kp, -ip): logical index of the mirrored direct index (P-k)%P, (P-i)%P
            const long long k2 = (g.P - k) % g.P, i2 = (g.P - i) % g.P;
            const double2 u = at(fftw_logical(k2, g.Ph, g.P), fftw_logical(i2, g.Ph, g.P), jp);
            v = make_double2(0.5 * (v.x + u.x), 0.5 * (v.y - u.y));
        }
        if (((k ^ i ^ j) & 1) != 0) { v.x = -v.x; v.y = -v.y; }
        W[e] = v;
    }
}

// real-space window (Xmipp origins P/2 and ori/2) with the FFT normalisation, and the partial
// sums of the soft-mask background (r >= ori/2) per block
__global__ void k_recon_crop(ReconGeom g, const double* __restrict__ R, double* __restrict__ vol,
                             double* __restrict__ part_bg, double* __restrict__ part_w)
{
    const long long n = g.ori * g.ori * g.ori;
    const long long oP = g.P / 2, oO = g.ori / 2;
    const double radius = (double)g.ori / 2., radius_p = radius + 3.;
    double sbg = 0., sw = 0.;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const long long x = e % g.ori, y = (e / g.ori) % g.ori, z = e / (g.ori * g.ori);
        const double v = R[((z - oO + oP) * g.P + (y - oO + oP)) * g.P + (x - oO + oP)] / g.normfft;
        vol[e] = v;
        const long long kk = z - oO, ii = y - oO, jj = x - oO;
        const double r = sqrt((double)(kk * kk + ii * ii + jj * jj));
        if (r < radius) continue;
        if (r > radius_p) { sw += 1.; sbg += v; }
        else {
            const double rc = 0.5 + 0.5 * cos(kPi * (radius_p - r) / 3.);
            sw += rc;
            sbg += rc * v;
        }
    }
    __shared__ double rb[1024], rw[1024];
    rb[threadIdx.x] = sbg;
    rw[threadIdx.x] = sw;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { rb[threadIdx.x] += rb[threadIdx.x + s]; rw[threadIdx.x] += rw[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part_bg[blockIdx.x] = rb[0]; part_w[blockIdx.x] = rw[0]; }
}

// soft mask with the background average, then the trilinear gridding correction
__global__ void k_recon_finish(ReconGeom g, const double* part_bg, const double* part_w, int nblocks,
                               double* __restrict__ vol)
{
    __shared__ double bg;
    if (threadIdx.x == 0) {
        double a = 0., c = 0.;
        for (int b = 0; b < nblocks; b++) { a += part_bg[b]; c += part_w[b]; }
        bg = a / c;
    }
    __syncthreads();
    const long long n = g.ori * g.ori * g.ori;
    const long long oO = g.ori / 2;
    const double radius = (double)g.ori / 2., radius_p = radius + 3.;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const long long x = e % g.ori, y = (e / g.ori) % g.ori, z = e / (g.ori * g.ori);
        const long long kk = z - oO, ii = y - oO, jj = x - oO;
        const double r = sqrt((double)(kk * kk + ii * ii + jj * jj));
        double v = vol[e];
        if (r >= radius) {
            if (r > radius_p) v = bg;
            else {
                const double rc = 0.5 + 0.5 * cos(kPi * (radius_p - r) / 3.);
                v = (1 - rc) * v + rc * bg;
            }
        }
        if (r > 0.) {
            const double rval = r / ((double)g.ori * g.pf);
            const double sinc = sin(kPi * rval) / (kPi * rval);
            v /= sinc * sinc;
        }
        vol[e] = v;
    }
}

// FSC partial sums per block and shell (shell = ROUND(|k|) < n/2+1), forward transforms / size
__global__ void k_fsc_partial(long long n, const double2* __restrict__ A, const double2* __restrict__ B,
                              double* __restrict__ part)   // [blocks][3][xh]
{
    extern __shared__ double sh[];
    const long long xh = n / 2 + 1;
    for (long long s = threadIdx.x; s < 3 * xh; s += blockDim.x) sh[s] = 0.;
    __syncthreads();
    const double scale = 1.0 / ((double)n * n * n);
    const long long tot = n * n * xh;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
        const long long j = e % xh, i = (e / xh) % n, k = e / (xh * n);
        const long long kp = fftw_logical(k, xh, n), ip = fftw_logical(i, xh, n), jp = j;
        const long long idx = llround(sqrt((double)(kp * kp + ip * ip + jp * jp)));
        if (idx >= xh) continue;
        const double2 a = A[e], b = B[e];
        const double ar = a.x * scale, ai = a.y * scale, br = b.x * scale, bi = b.y * scale;
        atomicAdd(&sh[idx], ar * br + ai * bi);
        atomicAdd(&sh[xh + idx], ar * ar + ai * ai);
        atomicAdd(&sh[2 * xh + idx], br * br + bi * bi);
    }
    __syncthreads();
    for (long long s = threadIdx.x; s < 3 * xh; s += blockDim.x) part[(size_t)blockIdx.x * 3 * xh + s] = sh[s];
}

__global__ void k_fsc_final(long long n, const double* part, int nblocks, double* fsc)
{
    const long long xh = n / 2 + 1;
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= xh) return;
    double num = 0., d1 = 0., d2 = 0.;
    for (int b = 0; b < nblocks; b++) {
        num += part[(size_t)b * 3 * xh + s];
        d1 += part[(size_t)b * 3 * xh + xh + s];
        d2 += part[(size_t)b * 3 * xh + 2 * xh + s];
    }
    fsc[s] = num / sqrt(d1 * d2);
}

// The device backprojector's FP32 complex data and weight (projector layout, as accumulated by the
// E-step) widened to the FP64 inputs of the reconstruction: the same values the host path reads
// back into BackProjector::data and ::weight.
__global__ void k_bp_widen(size_t n, const float2* __restrict__ d, const float* __restrict__ w,
                           double* __restrict__ re, double* __restrict__ im, double* __restrict__ wo)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const float2 v = d[e];
        re[e] = v.x;
        im[e] = v.y;
        wo[e] = w[e];
    }
}

struct FftPlan {
    hipfftHandle h = 0;
    ~FftPlan() { if (h) hipfftDestroy(h); }
};

}  // namespace

}  // namespace t2x

using namespace t2x;

// Shared body of t2x_reconstruct (inputs from host arrays) and t2x_reconstruct_class (inputs from
// the device backprojector `bp`); the caller holds the device lock.
static int reconstruct_run(Device* d, const t2x_recon_params* p, const double* data_re, const double* data_im,
                           const double* weight, const ClassVolumes* bp, const double* tau2, double* vol_out)
{
    if (p->ref_dim != 3) { set_error("t2x_reconstruct: only 3D references are supported"); return T2X_EINVAL; }
    if (!p->skip_gridding) {
        set_error("t2x_reconstruct: the iterative gridding branch (--dont_skip_gridding) is not supported");
        return T2X_EINVAL;
    }
    if (p->ori_size < 2 || p->ori_size % 2 || p->r_max < 1 || p->padding_factor <= 0. ||
        p->pad_size != 2 * ((int)std::lround(p->padding_factor * p->r_max) + 1) + 1) {
        set_error("t2x_reconstruct: inconsistent sizes (pad_size must be 2*(ROUND(pf*r_max)+1)+1, even ori_size)");
        return T2X_EINVAL;
    }
    if (p->do_map && !tau2) { set_error("t2x_reconstruct: do_map needs tau2"); return T2X_EINVAL; }
    if (bp && (bp->b_ref_dim != p->ref_dim || bp->b_pad != p->pad_size || bp->b_rmax != p->r_max ||
               bp->b_pf != p->padding_factor)) {
        set_error("t2x_reconstruct_class: the backprojector's ref_dim/pad_size/r_max/padding_factor differ from the params");
        return T2X_EINVAL;
    }
    T2X_HIP(hipSetDevice(d->id));
    hipStream_t st = d->stream;

    ReconGeom g{};
    g.pad = p->pad_size;
    g.xh = g.pad / 2 + 1;
    g.pf = p->padding_factor;
    g.ori = p->ori_size;
    long long P = std::llround(g.pf * g.ori);
    P += P % 2;
    g.P = P;
    g.Ph = P / 2 + 1;
    const long long rr = std::llround(p->r_max * g.pf);
    g.max_r2 = rr * rr;
    g.round_max_r2 = std::llround(p->r_max * g.pf * p->r_max * g.pf);
    g.win_r2 = (g.xh - 1) * (g.xh - 1);
    g.r_max = p->r_max;
    g.do_map = p->do_map;
    g.minres_map = p->minres_map;
    g.tau2_fudge = p->tau2_fudge;
    g.normfft = (p->data_dim == 3) ? g.pf * g.pf * g.pf : g.pf * g.pf * g.pf * (double)g.ori;

    const size_t nv = (size_t)g.pad * g.pad * g.xh;
    const size_t nw = (size_t)P * P * g.Ph;
    const size_t nr = (size_t)P * P * P;
    const size_t no = (size_t)g.ori * g.ori * g.ori;
    const int nshell_tau = g.ori / 2 + 1;
    DevBuf b_re, b_im, b_w, b_tau, b_Fw, b_ps, b_pc, b_floor, b_err, b_W, b_R, b_vol, b_bg, b_bw;
    T2X_TRY(b_re.ensure(nv * 8));
    T2X_TRY(b_im.ensure(nv * 8));
    T2X_TRY(b_w.ensure(nv * 8));
    T2X_TRY(b_tau.ensure((size_t)nshell_tau * 8));
    T2X_TRY(b_Fw.ensure(nv * 8));
    T2X_TRY(b_ps.ensure((size_t)kShellBlocks * g.r_max * 8));
    T2X_TRY(b_pc.ensure((size_t)kShellBlocks * g.r_max * 8));
    T2X_TRY(b_floor.ensure((size_t)g.r_max * 8));
    T2X_TRY(b_err.ensure(sizeof(int)));
    T2X_TRY(b_W.ensure(nw * 16));
    T2X_TRY(b_R.ensure(nr * 8));
    T2X_TRY(b_vol.ensure(no * 8));
    T2X_TRY(b_bg.ensure(kShellBlocks * 8));
    T2X_TRY(b_bw.ensure(kShellBlocks * 8));
    if (bp) {
        k_bp_widen<<<(unsigned)std::min<size_t>((nv + 255) / 256, 65536), 256, 0, st>>>(
            nv, bp->bp_data, bp->bp_weight, b_re.as<double>(), b_im.as<double>(), b_w.as<double>());
        T2X_HIP(hipGetLastError());
    } else {
        T2X_HIP(hipMemcpyAsync(b_re.ptr, data_re, nv * 8, hipMemcpyHostToDevice, st));
        T2X_HIP(hipMemcpyAsync(b_im.ptr, data_im, nv * 8, hipMemcpyHostToDevice, st));
        T2X_HIP(hipMemcpyAsync(b_w.ptr, weight, nv * 8, hipMemcpyHostToDevice, st));
    }
    if (p->do_map) T2X_HIP(hipMemcpyAsync(b_tau.ptr, tau2, (size_t)nshell_tau * 8, hipMemcpyHostToDevice, st));
    else T2X_HIP(hipMemsetAsync(b_tau.ptr, 0, (size_t)nshell_tau * 8, st));
    T2X_HIP(hipMemsetAsync(b_err.ptr, 0, sizeof(int), st));
    // the MAP shell index ROUND(|k|/pf) for |k|^2 < max_r2 stays below r_max + 1 <= ori/2 + 1
    if (p->do_map && (long long)p->r_max + 1 > nshell_tau) {
        set_error("t2x_reconstruct: r_max exceeds the tau2 spectrum (ori_size/2+1 shells)");
        return T2X_EINVAL;
    }

    k_recon_weights<<<kShellBlocks, 256, 2 * sizeof(double) * g.r_max, st>>>(
        g, b_w.as<double>(), b_tau.as<double>(), b_Fw.as<double>(), b_ps.as<double>(), b_pc.as<double>(),
        b_err.as<int>());
    T2X_HIP(hipGetLastError());
    k_recon_floor<<<(g.r_max + 255) / 256, 256, 0, st>>>(g, b_ps.as<double>(), b_pc.as<double>(), kShellBlocks,
                                                         b_floor.as<double>(), b_err.as<int>());
    T2X_HIP(hipGetLastError());
    k_recon_window<<<(unsigned)std::min<size_t>((nw + 255) / 256, 65536), 256, 0, st>>>(
        g, b_re.as<double>(), b_im.as<double>(), b_Fw.as<double>(), b_floor.as<double>(), b_W.as<double2>());
    T2X_HIP(hipGetLastError());
    int herr = 0;
    T2X_HIP(hipMemcpyAsync(&herr, b_err.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
    T2X_HIP(hipStreamSynchronize(st));
    if (herr == 2) { set_error("t2x_reconstruct: negative tau2 values"); return T2X_EINVAL; }
    if (herr == 3 || herr == 4) { set_error("t2x_reconstruct: empty or inconsistent weight shells"); return T2X_EINVAL; }

    FftPlan plan;
    if (hipfftPlan3d(&plan.h, (int)P, (int)P, (int)P, HIPFFT_Z2D) != HIPFFT_SUCCESS ||
        hipfftSetStream(plan.h, st) != HIPFFT_SUCCESS) {
        set_error("t2x_reconstruct: hipFFT Z2D plan failed");
        return T2X_EDEVICE;
    }
    if (hipfftExecZ2D(plan.h, b_W.as<hipfftDoubleComplex>(), b_R.as<hipfftDoubleReal>()) != HIPFFT_SUCCESS) {
        set_error("t2x_reconstruct: hipFFT Z2D failed");
        return T2X_EDEVICE;
    }
    k_recon_crop<<<kShellBlocks, 1024, 0, st>>>(g, b_R.as<double>(), b_vol.as<double>(), b_bg.as<double>(),
                                                b_bw.as<double>());
    T2X_HIP(hipGetLastError());
    k_recon_finish<<<(unsigned)std::min<size_t>((no + 255) / 256, 65536), 256, 0, st>>>(
        g, b_bg.as<double>(), b_bw.as<double>(), kShellBlocks, b_vol.as<double>());
    T2X_HIP(hipGetLastError());
    T2X_HIP(hipMemcpyAsync(vol_out, b_vol.ptr, no * 8, hipMemcpyDeviceToHost, st));
    T2X_HIP(hipStreamSynchronize(st));
    return T2X_OK;
}

extern "C" int t2x_reconstruct(t2x_device* hdev, const t2x_recon_params* p, const double* data_re,
                               const double* data_im, const double* weight, const double* tau2, double* vol_out)
{
    if (!hdev || !p || !data_re || !data_im || !weight || !vol_out) { set_error("null argument"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(hdev);
    std::lock_guard<std::mutex> lk(d->mtx);
    return reconstruct_run(d, p, data_re, data_im, weight, nullptr, tau2, vol_out);
}

extern "C" int t2x_reconstruct_class(t2x_device* hdev, int iclass, const t2x_recon_params* p, const double* tau2,
                                     double* vol_out)
{
    if (!hdev || !p || !vol_out) { set_error("null argument"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(hdev);
    std::lock_guard<std::mutex> lk(d->mtx);
    auto it = d->classes.find(iclass);
    if (it == d->classes.end() || !it->second.bp_data) {
        set_error("t2x_reconstruct_class: backprojector not initialised");
        return T2X_ESTATE;
    }
    return reconstruct_run(d, p, nullptr, nullptr, nullptr, &it->second, tau2, vol_out);
}

extern "C" int t2x_fsc(t2x_device* hdev, int n, const double* map1, const double* map2, double* fsc)
{
    if (!hdev || !map1 || !map2 || !fsc) { set_error("null argument"); return T2X_EINVAL; }
    if (n < 2) { set_error("t2x_fsc: n < 2"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(hdev);
    std::lock_guard<std::mutex> lk(d->mtx);
    T2X_HIP(hipSetDevice(d->id));
    hipStream_t st = d->stream;
    const long long N = n, xh = N / 2 + 1;
    const size_t nr = (size_t)N * N * N, nc = (size_t)N * N * xh;
    DevBuf r1, r2, c1, c2, part, out;
    T2X_TRY(r1.ensure(nr * 8));
    T2X_TRY(r2.ensure(nr * 8));
    T2X_TRY(c1.ensure(nc * 16));
    T2X_TRY(c2.ensure(nc * 16));
    T2X_TRY(part.ensure((size_t)kShellBlocks * 3 * xh * 8));
    T2X_TRY(out.ensure((size_t)xh * 8));
    T2X_HIP(hipMemcpyAsync(r1.ptr, map1, nr * 8, hipMemcpyHostToDevice, st));
    T2X_HIP(hipMemcpyAsync(r2.ptr, map2, nr * 8, hipMemcpyHostToDevice, st));
    FftPlan plan;
    if (hipfftPlan3d(&plan.h, n, n, n, HIPFFT_D2Z) != HIPFFT_SUCCESS || hipfftSetStream(plan.h, st) != HIPFFT_SUCCESS) {
        set_error("t2x_fsc: hipFFT D2Z plan failed");
        return T2X_EDEVICE;
    }
    if (hipfftExecD2Z(plan.h, r1.as<hipfftDoubleReal>(), c1.as<hipfftDoubleComplex>()) != HIPFFT_SUCCESS ||
        hipfftExecD2Z(plan.h, r2.as<hipfftDoubleReal>(), c2.as<hipfftDoubleComplex>()) != HIPFFT_SUCCESS) {
        set_error("t2x_fsc: hipFFT D2Z failed");
        return T2X_EDEVICE;
    }
    k_fsc_partial<<<kShellBlocks, 256, 3 * sizeof(double) * xh, st>>>(N, c1.as<double2>(), c2.as<double2>(),
                                                                       part.as<double>());
    T2X_HIP(hipGetLastError());
    k_fsc_final<<<(unsigned)((xh + 255) / 256), 256, 0, st>>>(N, part.as<double>(), kShellBlocks, out.as<double>());
    T2X_HIP(hipGetLastError());
    T2X_HIP(hipMemcpyAsync(fsc, out.ptr, (size_t)xh * 8, hipMemcpyDeviceToHost, st));
    T2X_HIP(hipStreamSynchronize(st));
    return T2X_OK;
}
