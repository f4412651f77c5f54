// t2x_synth.hip — synthetic workloads on the GPU: analytic Fourier transforms of Gaussian-blob
// volumes on the projector grid, and CTF-modulated, shifted, noisy particle images.
//
// A blob amp*exp(-|r-c|^2/(2 s^2)) has the continuous transform
//   amp (2 pi s^2)^{3/2} exp(-2 pi^2 s^2 |f|^2) exp(-2 pi i f.c)     (f in cycles/px).
// The projector grid stores V(f)/N^2, the scale Projector::computeFourierTransformMap gives a
// 3D reference used against 2D images (projector.cpp:146-160, 407). Particle images use the
// projection-slice theorem with the projector's convention (xp,yp,zp) = A^T (x,y,0).
#include "t2x_internal.h"

namespace t2x {

static const double kPiS = 3.14159265358979323846;

__device__ __forceinline__ double2 blob_sum(const t2x_blob* __restrict__ bl, int nb, double fx,
                                            double fy, double fz)
{
    double f2 = fx * fx + fy * fy + fz * fz;
    double re = 0., im = 0.;
    for (int k = 0; k < nb; k++) {
        double s2 = bl[k].sigma * bl[k].sigma;
        double mag = bl[k].amplitude * pow(2 * kPiS * s2, 1.5) * exp(-2 * kPiS * kPiS * s2 * f2);
        double ph = -2 * kPiS * (fx * bl[k].x + fy * bl[k].y + fz * bl[k].z);
        double s, c;
        sincos(ph, &s, &c);
        re += mag * c;
        im += mag * s;
    }
    return make_double2(re, im);
}

__global__ void k_synth_reference(const t2x_blob* bl, int nb, int N, double angpix, int pad,
                                  int max_r2, double pf, double cutoff_shell, float* re, float* im)
{
    const int xdim = pad / 2 + 1, h = pad / 2;
    const size_t total = (size_t)pad * pad * xdim;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        int kx = (int)(idx % xdim);
        int ky = (int)((idx / xdim) % pad);
        int kz = (int)(idx / ((size_t)xdim * pad));
        int x = kx, y = ky - h, z = kz - h;
        int r2 = x * x + y * y + z * z;
        double vr = 0., vi = 0.;
        if (r2 <= max_r2) {
            double f = 1.0 / (pf * N);
            double2 v = blob_sum(bl, nb, x * f, y * f, z * f);
            double w = 1.0;
            if (cutoff_shell > 0) {
                double shell = sqrt((double)r2) / pf, lo = cutoff_shell - 1.0, hi = cutoff_shell + 1.0;
                if (shell >= hi) w = 0.;
                else if (shell > lo) w = 0.5 + 0.5 * cos(kPiS * (shell - lo) / (hi - lo));
            }
            double inv = w / ((double)N * N);
            vr = v.x * inv;
            vi = v.y * inv;
        }
        re[idx] = (float)vr;
        im[idx] = (float)vi;
    }
}

struct SynthPart {
    double A[9];
    double tx, ty;
    CtfConst ctf;
    int cls;
};

__global__ void k_synth_fourier(const t2x_blob* bl, const int* boff, const SynthPart* parts, int N,
                                double angpix, float2* F)
{
    const int p = blockIdx.y;
    const int xs = N / 2 + 1;
    const SynthPart& sp = parts[p];
    const t2x_blob* b = bl + boff[sp.cls];
    const int nb = boff[sp.cls + 1] - boff[sp.cls];
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < N * xs; q += gridDim.x * blockDim.x) {
        int i = q / xs, j = q % xs;
        int ip = (i < xs) ? i : i - N;
        double gx = (double)j / N, gy = (double)ip / N;
        double fx = sp.A[0] * gx + sp.A[3] * gy, fy = sp.A[1] * gx + sp.A[4] * gy, fz = sp.A[2] * gx + sp.A[5] * gy;
        double2 v = blob_sum(b, nb, fx, fy, fz);
        double inv = 1.0 / ((double)N * N);
        double vr = v.x * inv, vi = v.y * inv;
        double s, c;
        sincos(-2 * kPiS * (gx * sp.tx + gy * sp.ty), &s, &c);
        double r = vr * c - vi * s, im = vr * s + vi * c;
        double ctf = ctf_value(sp.ctf, j / (N * angpix), ip / (N * angpix));
        r *= ctf;
        im *= ctf;
        if ((ip + j) & 1) { r = -r; im = -im; }
        F[(size_t)p * N * xs + q] = make_float2((float)r, (float)im);
    }
}

// noise_sample (t2x_internal.h): the counter-based Gaussian of the particle noise.

__global__ __launch_bounds__(256) void k_synth_noise(float* img, int N, const double* snr, uint64_t seed,
                                                     long p_offset, double* noise_var)
{
    const int p = blockIdx.x;
    const long gp = p_offset + p;
    const size_t NN = (size_t)N * N;
    float* im = img + (size_t)p * NN;
    double s = 0., s2 = 0.;
    for (size_t q = threadIdx.x; q < NN; q += blockDim.x) {
        double v = im[q];
        s += v;
        s2 += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) { s += __shfl_xor(s, o); s2 += __shfl_xor(s2, o); }
    __shared__ double r1[4], r2[4];
    if ((threadIdx.x & 63) == 0) { r1[threadIdx.x >> 6] = s; r2[threadIdx.x >> 6] = s2; }
    __syncthreads();
    double mean = (r1[0] + r1[1] + r1[2] + r1[3]) / NN;
    double var = (r2[0] + r2[1] + r2[2] + r2[3]) / NN - mean * mean;
    double sd = sqrt(var / snr[p]);
    if (threadIdx.x == 0 && noise_var) noise_var[p] = var / snr[p];
    for (size_t q = threadIdx.x; q < NN; q += blockDim.x)
        im[q] = (float)((double)im[q] + sd * noise_sample(seed, gp, (long)q));
}

}  // namespace t2x

using namespace t2x;

extern "C" int t2x_synth_reference_device(t2x_device* h, int N, double angpix, int pad, int r_max,
                                          double pf, const t2x_blob* blobs, int nb, double lowpass_A,
                                          float* dre, float* dim)
{
    if (!h || !blobs || !dre || !dim || nb <= 0 || pad < 3) { set_error("bad argument"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    T2X_HIP(hipSetDevice(d->id));
    t2x_blob* db = nullptr;
    T2X_HIP(hipMalloc(&db, sizeof(t2x_blob) * nb));
    T2X_HIP(hipMemcpy(db, blobs, sizeof(t2x_blob) * nb, hipMemcpyHostToDevice));
    int rr = (int)(r_max * pf + 0.5);
    double cutoff = lowpass_A > 0 ? N * angpix / lowpass_A : 0.;
    k_synth_reference<<<4096, 256, 0, d->stream>>>(db, nb, N, angpix, pad, rr * rr, pf, cutoff, dre, dim);
    T2X_HIP(hipGetLastError());
    T2X_HIP(hipStreamSynchronize(d->stream));
    (void)hipFree(db);
    return T2X_OK;
}

extern "C" int t2x_synth_particles_device(t2x_device* h, int N, double angpix, int n,
                                          const t2x_blob* blobs, const int* boff,
                                          const t2x_particle_in* pin, const double* trot,
                                          const double* ttilt, const double* tpsi, const double* tx,
                                          const double* ty, const int* tcls, const double* snr,
                                          uint64_t seed, float* dimages, double* noise_var)
{
    if (!h || !blobs || !boff || !pin || !dimages || n < 0) { set_error("bad argument"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    T2X_HIP(hipSetDevice(d->id));
    int K = 0;
    for (int i = 0; i < n; i++) K = std::max(K, tcls[i] + 1);
    int nb_total = boff[K];
    const int xs = N / 2 + 1;
    const int chunk = 512;
    t2x_blob* db = nullptr;
    int* dboff = nullptr;
    SynthPart* dparts = nullptr;
    float2* dF = nullptr;
    double* dsnr = nullptr;
    double* dvar = nullptr;
    T2X_HIP(hipMalloc(&dvar, sizeof(double) * chunk));
    T2X_HIP(hipMalloc(&db, sizeof(t2x_blob) * nb_total));
    T2X_HIP(hipMalloc(&dboff, sizeof(int) * (K + 1)));
    T2X_HIP(hipMalloc(&dparts, sizeof(SynthPart) * chunk));
    T2X_HIP(hipMalloc(&dF, sizeof(float2) * (size_t)chunk * N * xs));
    T2X_HIP(hipMalloc(&dsnr, sizeof(double) * chunk));
    T2X_HIP(hipMemcpy(db, blobs, sizeof(t2x_blob) * nb_total, hipMemcpyHostToDevice));
    T2X_HIP(hipMemcpy(dboff, boff, sizeof(int) * (K + 1), hipMemcpyHostToDevice));
    hipfftHandle plan = 0;
    int plan_batch = -1;
    int status = T2X_OK;
    for (int p0 = 0; p0 < n && status == T2X_OK; p0 += chunk) {
        int m = std::min(chunk, n - p0);
        std::vector<SynthPart> hp(m);
        for (int i = 0; i < m; i++) {
            euler_matrix_host(trot[p0 + i], ttilt[p0 + i], tpsi[p0 + i], hp[i].A);
            hp[i].tx = tx[p0 + i];
            hp[i].ty = ty[p0 + i];
            hp[i].ctf = ctf_const(pin[p0 + i].ctf);
            hp[i].cls = tcls[p0 + i];
        }
        T2X_HIP(hipMemcpy(dparts, hp.data(), sizeof(SynthPart) * m, hipMemcpyHostToDevice));
        T2X_HIP(hipMemcpy(dsnr, snr + p0, sizeof(double) * m, hipMemcpyHostToDevice));
        dim3 grid((unsigned)((N * xs + 255) / 256), (unsigned)m);
        k_synth_fourier<<<grid, 256, 0, d->stream>>>(db, dboff, dparts, N, angpix, dF);
        T2X_HIP(hipGetLastError());
        if (plan_batch != m) {
            if (plan) (void)hipfftDestroy(plan);
            int dims[2] = {N, N};
            if (hipfftPlanMany(&plan, 2, dims, nullptr, 1, N * xs, nullptr, 1, N * N, HIPFFT_C2R, m) != HIPFFT_SUCCESS) {
                set_error("hipfftPlanMany C2R failed");
                status = T2X_EDEVICE;
                break;
            }
            (void)hipfftSetStream(plan, d->stream);
            plan_batch = m;
        }
        if (hipfftExecC2R(plan, reinterpret_cast<hipfftComplex*>(dF), dimages + (size_t)p0 * N * N) != HIPFFT_SUCCESS) {
            set_error("hipfftExecC2R failed");
            status = T2X_EDEVICE;
            break;
        }
        k_synth_noise<<<m, 256, 0, d->stream>>>(dimages + (size_t)p0 * N * N, N, dsnr, seed, p0, dvar);
        T2X_HIP(hipGetLastError());
        T2X_HIP(hipStreamSynchronize(d->stream));
        if (noise_var) T2X_HIP(hipMemcpy(noise_var + p0, dvar, sizeof(double) * m, hipMemcpyDeviceToHost));
    }
    if (plan) (void)hipfftDestroy(plan);
    (void)hipFree(db);
    (void)hipFree(dboff);
    (void)hipFree(dparts);
    (void)hipFree(dF);
    (void)hipFree(dsnr);
    (void)hipFree(dvar);
    return status;
}
