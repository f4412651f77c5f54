// t2x_proposal.hip — C ABI of the pose-proposal network: initialisation, parameter I/O, the
// gradient entry point used by the tests, and GPU training on projections of the current
// references.
//
// Training data are made on the fly: a fixed set of projection parameter sets (class, SO(3)
// pose, shift, CTF) is drawn once on the host; each step takes the next `batch` of them and
// synthesises their 64x64 inputs on the GPU with fresh noise: the Fourier slice of the class's
// projector at the 64x64 low-resolution crop of the particle box (|k| <= 32 of the ori-box
// frequency grid), times the CTF and the shift phase, plus complex Gaussian noise with the
// model's per-shell variance sigma2_noise; then a batched 64x64 inverse FFT and per-image
// standardisation. The label is the bin of the coarse HEALPix direction nearest to the
// projection direction (its parent one order down) for that class. Inputs of real particles
// are the same low-resolution crop of their preprocessed transforms (see the E-step).
#include "t2x_cnn.h"

#include <hipfft/hipfft.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace t2x {

int build_sampling(const t2x_estep_config* cfg, HostSampling& s);

namespace cnn {

template <class T>
static int upload_vec(DevBuf& buf, const std::vector<T>& v, hipStream_t st)
{
    T2X_TRY(buf.ensure(sizeof(T) * std::max<size_t>(1, v.size())));
    if (!v.empty()) T2X_HIP(hipMemcpyAsync(buf.ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
    return T2X_OK;
}

constexpr int kCropX = kIn / 2 + 1;   // 33 half-plane columns of the 64x64 transform
constexpr double kPiP = 3.14159265358979323846;

struct TrainSample {
    float Ai[9];      // inverse rotation x padding factor (projector coordinates of a 2D frequency)
    CtfConst ctf;
    float sx, sy;     // shift (px)
    int cls;
    int label;
};

__device__ __forceinline__ uint64_t pmix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// two independent standard normals keyed by (key, index)
__device__ __forceinline__ float2 gauss2(uint64_t key, uint64_t idx)
{
    const uint64_t a = pmix(key ^ pmix(2 * idx)), b = pmix(key ^ pmix(2 * idx + 1));
    const double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740992.0), u2 = (b >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincos(2.0 * kPiP * u2, &s, &c);
    return make_float2((float)(r * c), (float)(r * s));
}

// Fourier crop of one training sample: rows ky (FFTW layout of 64), columns kx in [0, 32]
__global__ void k_train_synth(const TrainSample* __restrict__ samples, long long first, int n_samples,
                              const float2* const* __restrict__ vols, const int* __restrict__ pads,
                              int r_max, int ori, double angpix, const double* __restrict__ sigma2,
                              uint64_t noise_key, float2* __restrict__ out)
{
    const int b = blockIdx.x;
    const TrainSample s = samples[(first + b) % n_samples];
    const float2* vol = vols[s.cls];
    const int pad = pads[s.cls];
    for (int q = threadIdx.x; q < kIn * kCropX; q += blockDim.x) {
        const int i = q / kCropX, j = q % kCropX;
        const int ky = (i < kCropX) ? i : i - kIn, kx = j;
        float2 v = make_float2(0.f, 0.f);
        const int r2 = kx * kx + ky * ky;
        if (r2 <= r_max * r_max) {
            const float xp = s.Ai[0] * kx + s.Ai[1] * ky, yp = s.Ai[3] * kx + s.Ai[4] * ky,
                        zp = s.Ai[6] * kx + s.Ai[7] * ky;
            float2 f;
            if (trilinear_sample(vol, pad, xp, yp, zp, &f)) v = f;
            const double ctf = ctf_value(s.ctf, kx / (ori * angpix), ky / (ori * angpix));
            double sn, cs;
            sincos(-2.0 * kPiP * ((double)kx * s.sx + (double)ky * s.sy) / ori, &sn, &cs);
            const double re = (v.x * cs - v.y * sn) * ctf, im = (v.x * sn + v.y * cs) * ctf;
            const int shell = (int)floor(sqrt((double)r2) + 0.5);
            const double sd = sqrt(fmax(sigma2[min(shell, ori / 2)], 0.));
            const float2 g = gauss2(noise_key ^ pmix((uint64_t)(first + b)), (uint64_t)q);
            v = make_float2((float)(re + sd * g.x), (float)(im + sd * g.y));
        }
        if ((ky + kx) & 1) { v.x = -v.x; v.y = -v.y; }   // image centre at (32, 32)
        out[(size_t)b * kIn * kCropX + q] = v;
    }
}

// per-image standardisation of [B][64][64] (zero mean, unit variance) into the network input
__global__ void k_standardise(const float* __restrict__ img, float* __restrict__ x)
{
    const int b = blockIdx.x;
    const float* p = img + (size_t)b * kIn * kIn;
    double s = 0., s2 = 0.;
    for (int q = threadIdx.x; q < kIn * kIn; q += blockDim.x) {
        const double v = p[q];
        s += v;
        s2 += v * v;
    }
    __shared__ double r1[256], r2[256];
    r1[threadIdx.x] = s;
    r2[threadIdx.x] = s2;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) { r1[threadIdx.x] += r1[threadIdx.x + o]; r2[threadIdx.x] += r2[threadIdx.x + o]; }
        __syncthreads();
    }
    const double n = (double)kIn * kIn;
    const double mean = r1[0] / n, var = fmax(r2[0] / n - mean * mean, 1e-30);
    const float inv = (float)(1.0 / sqrt(var));
    for (int q = threadIdx.x; q < kIn * kIn; q += blockDim.x)
        x[(size_t)b * kIn * kIn + q] = (float)(p[q] - mean) * inv;
}

__global__ void k_train_labels(const TrainSample* __restrict__ samples, long long first, int n_samples, int B,
                               double max_shift, int* __restrict__ label, float* __restrict__ shift_t)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const TrainSample s = samples[(first + b) % n_samples];
    label[b] = s.label;
    shift_t[2 * b] = (float)(s.sx / max_shift);
    shift_t[2 * b + 1] = (float)(s.sy / max_shift);
}

int crop_to_input(Net& net, const float2* crop, int B, hipStream_t st, DevBuf& img, hipfftHandle& plan, int& plan_b)
{
    T2X_TRY(img.ensure(sizeof(float) * (size_t)B * kIn * kIn));
    if (plan_b != B) {
        if (plan) hipfftDestroy(plan);
        plan = 0;
        int n[2] = {kIn, kIn};
        if (hipfftPlanMany(&plan, 2, n, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_C2R, B) != HIPFFT_SUCCESS) {
            set_error("proposal: hipFFT C2R plan failed");
            plan_b = 0;
            return T2X_EDEVICE;
        }
        plan_b = B;
    }
    if (hipfftSetStream(plan, st) != HIPFFT_SUCCESS ||
        hipfftExecC2R(plan, (hipfftComplex*)crop, img.as<hipfftReal>()) != HIPFFT_SUCCESS) {
        set_error("proposal: hipFFT C2R failed");
        return T2X_EDEVICE;
    }
    k_standardise<<<B, 256, 0, st>>>(img.as<float>(), net.act[0].as<float>());
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

// low-resolution crop of preprocessed particle transforms (cs x (cs/2+1), FFTW rows, centred
// image convention) into the 64x64 half-plane grid, with the centring sign of k_train_synth
__global__ void k_particle_crop(const float2* __restrict__ F, int cs, float2* __restrict__ out)
{
    const int b = blockIdx.x;
    const int xc = cs / 2 + 1;
    for (int q = threadIdx.x; q < kIn * kCropX; q += blockDim.x) {
        const int i = q / kCropX, j = q % kCropX;
        const int ky = (i < kCropX) ? i : i - kIn, kx = j;
        float2 v = make_float2(0.f, 0.f);
        if (kx < xc && ky < xc && -ky < xc) {
            const int row = ky >= 0 ? ky : ky + cs;
            v = F[(size_t)b * cs * xc + (size_t)row * xc + kx];
        }
        if ((ky + kx) & 1) { v.x = -v.x; v.y = -v.y; }
        out[(size_t)b * kIn * kCropX + q] = v;
    }
}

// per particle: softmax over the bin logits, top-k bins (descending logit, ties by index) and the
// top probability; fallback when it is below min_conf
__global__ void k_topk(const float* __restrict__ y, int n_bins, int topk, float min_conf, int gate,
                       int* __restrict__ bins, float* __restrict__ maxp, float* __restrict__ mass,
                       int* __restrict__ fallback)
{
    extern __shared__ float sl[];   // n_bins logits, then the topk logits of the selected bins
    float* tv = sl + n_bins;
    __shared__ float rv[256];
    __shared__ int ri[256];
    const int b = blockIdx.x;
    const float* yb = y + (size_t)b * (n_bins + 2);
    for (int i = threadIdx.x; i < n_bins; i += blockDim.x) sl[i] = yb[i];
    __syncthreads();
    float z = 0.f, m0 = 0.f;
    for (int k = 0; k < topk; k++) {
        float bv = -3.0e38f;
        int bi = 0x7fffffff;
        for (int i = threadIdx.x; i < n_bins; i += blockDim.x)
            if (sl[i] > bv || (sl[i] == bv && i < bi)) { bv = sl[i]; bi = i; }
        rv[threadIdx.x] = bv;
        ri[threadIdx.x] = bi;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                const float ov = rv[threadIdx.x + o];
                const int oi = ri[threadIdx.x + o];
                if (ov > rv[threadIdx.x] || (ov == rv[threadIdx.x] && oi < ri[threadIdx.x])) {
                    rv[threadIdx.x] = ov;
                    ri[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        const int sel = ri[0];
        if (k == 0) m0 = rv[0];
        if (threadIdx.x == 0) tv[k] = rv[0];
        if (threadIdx.x == 0) bins[(size_t)b * topk + k] = sel;
        __syncthreads();
        if (threadIdx.x == 0) sl[sel] = -3.0e38f;   // removed; the softmax below uses yb
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n_bins; i += blockDim.x) z += expf(yb[i] - m0);
    rv[threadIdx.x] = z;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) rv[threadIdx.x] += rv[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float p = 1.f / rv[0];
        float mk = 0.f;   // softmax mass of the selected bins
        for (int k = 0; k < topk; k++) mk += expf(tv[k] - m0);
        mk *= p;
        maxp[b] = p;
        mass[b] = mk;
        fallback[b] = (gate == T2X_GATE_TOPK_MASS ? mk : p) < min_conf;
    }
}

// Network inference on B preprocessed particles: top-k bins, top probability, fallback flags
// (host vectors). Fimg: device [B][cs*(cs/2+1)].
int proposal_infer(Device* d, const float2* Fimg, int cs, int B, int topk, double min_conf, int gate,
                   std::vector<int>& bins, std::vector<float>& maxp, std::vector<float>& mass,
                   std::vector<int>& fallback)
{
    Net& net = *d->net;
    hipStream_t st = d->stream;
    T2X_TRY(net.ensure_batch(B));
    DevBuf& crop = d->prop_crop;
    T2X_TRY(crop.ensure(sizeof(float2) * (size_t)B * kIn * kCropX));
    k_particle_crop<<<B, 256, 0, st>>>(Fimg, cs, crop.as<float2>());
    T2X_HIP(hipGetLastError());
    T2X_TRY(crop_to_input(net, crop.as<float2>(), B, st, d->prop_img, d->prop_plan, d->prop_plan_b));
    T2X_TRY(net.forward(B, true, st));
    DevBuf& outb = d->prop_out;
    T2X_TRY(outb.ensure(sizeof(int) * (size_t)B * (topk + 3)));
    int* d_bins = outb.as<int>();
    float* d_maxp = reinterpret_cast<float*>(d_bins + (size_t)B * topk);
    float* d_mass = d_maxp + B;
    int* d_fb = d_bins + (size_t)B * (topk + 2);
    k_topk<<<B, 256, sizeof(float) * (net.n_bins + topk), st>>>(net.y.as<float>(), net.n_bins, topk, (float)min_conf, gate,
                                                       d_bins, d_maxp, d_mass, d_fb);
    T2X_HIP(hipGetLastError());
    bins.resize((size_t)B * topk);
    maxp.resize(B);
    mass.resize(B);
    fallback.resize(B);
    T2X_HIP(hipMemcpyAsync(bins.data(), d_bins, sizeof(int) * (size_t)B * topk, hipMemcpyDeviceToHost, st));
    T2X_HIP(hipMemcpyAsync(maxp.data(), d_maxp, sizeof(float) * B, hipMemcpyDeviceToHost, st));
    T2X_HIP(hipMemcpyAsync(mass.data(), d_mass, sizeof(float) * B, hipMemcpyDeviceToHost, st));
    T2X_HIP(hipMemcpyAsync(fallback.data(), d_fb, sizeof(int) * B, hipMemcpyDeviceToHost, st));
    T2X_HIP(hipStreamSynchronize(st));
    return T2X_OK;
}

}  // namespace cnn
}  // namespace t2x

using namespace t2x;
using namespace t2x::cnn;

static Net* get_net(t2x_device* h)
{
    Device* d = reinterpret_cast<Device*>(h);
    return d->net.get();
}

extern "C" int t2x_proposal_init(t2x_device* h, int n_bins, uint64_t seed)
{
    if (!h || n_bins < 1) { set_error("t2x_proposal_init: bad arguments"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    T2X_HIP(hipSetDevice(d->id));
    d->net = std::make_shared<Net>();
    T2X_TRY(d->net->init(n_bins, seed, d->stream));
    T2X_HIP(hipStreamSynchronize(d->stream));
    return T2X_OK;
}

extern "C" int t2x_proposal_num_params(t2x_device* h, long long* n)
{
    if (!h || !n) { set_error("null argument"); return T2X_EINVAL; }
    Net* net = get_net(h);
    if (!net) { set_error("proposal network not initialised"); return T2X_ESTATE; }
    *n = net->n_params;
    return T2X_OK;
}

extern "C" int t2x_proposal_get_params(t2x_device* h, float* host, long long n)
{
    if (!h || !host) { set_error("null argument"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    Net* net = d->net.get();
    if (!net) { set_error("proposal network not initialised"); return T2X_ESTATE; }
    if (n != net->n_params) { set_error("t2x_proposal_get_params: wrong size"); return T2X_EINVAL; }
    T2X_HIP(hipMemcpyAsync(host, net->params.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, d->stream));
    T2X_HIP(hipStreamSynchronize(d->stream));
    return T2X_OK;
}

extern "C" int t2x_proposal_set_params(t2x_device* h, const float* host, long long n)
{
    if (!h || !host) { set_error("null argument"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    Net* net = d->net.get();
    if (!net) { set_error("proposal network not initialised"); return T2X_ESTATE; }
    if (n != net->n_params) { set_error("t2x_proposal_set_params: wrong size"); return T2X_EINVAL; }
    T2X_HIP(hipMemcpyAsync(net->params.ptr, host, sizeof(float) * n, hipMemcpyHostToDevice, d->stream));
    T2X_HIP(hipStreamSynchronize(d->stream));
    return T2X_OK;
}

extern "C" int t2x_proposal_mark_trained(t2x_device* h, const t2x_estep_config* cfg)
{
    if (!h || !cfg) { set_error("null argument"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    Net* net = d->net.get();
    if (!net) { set_error("proposal network not initialised"); return T2X_ESTATE; }
    HostSampling S;
    T2X_TRY(build_sampling(cfg, S));
    const int n_bins = (int)(cfg->nr_classes * ((long)S.rot.size() / 4));
    if (n_bins != net->n_bins) { set_error("t2x_proposal_mark_trained: bins do not match the configuration"); return T2X_EINVAL; }
    net->trained = true;
    net->train_ori = cfg->ori_size;
    net->train_K = cfg->nr_classes;
    net->train_order = cfg->healpix_order;
    return T2X_OK;
}

extern "C" int t2x_proposal_loss_grad(t2x_device* h, int B, const float* x, const int* label,
                                      const float* shift_target, double shift_weight, int bf16, double* loss,
                                      float* grad, float* outputs)
{
    if (!h || B < 1 || !x || !label || !shift_target) { set_error("t2x_proposal_loss_grad: bad arguments"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    T2X_HIP(hipSetDevice(d->id));
    Net* net = d->net.get();
    if (!net) { set_error("proposal network not initialised"); return T2X_ESTATE; }
    for (int b = 0; b < B; b++)
        if (label[b] < 0 || label[b] >= net->n_bins) { set_error("t2x_proposal_loss_grad: label out of range"); return T2X_EINVAL; }
    hipStream_t st = d->stream;
    T2X_TRY(net->ensure_batch(B));
    T2X_HIP(hipMemcpyAsync(net->act[0].ptr, x, sizeof(float) * (size_t)B * kIn * kIn, hipMemcpyHostToDevice, st));
    T2X_HIP(hipMemcpyAsync(net->label.ptr, label, sizeof(int) * B, hipMemcpyHostToDevice, st));
    T2X_HIP(hipMemcpyAsync(net->shift_t.ptr, shift_target, sizeof(float) * 2 * B, hipMemcpyHostToDevice, st));
    T2X_TRY(net->loss_grad(B, bf16 != 0, (float)shift_weight, st));
    std::vector<double> lb(B);
    T2X_HIP(hipMemcpyAsync(lb.data(), net->loss.ptr, sizeof(double) * B, hipMemcpyDeviceToHost, st));
    if (grad) T2X_HIP(hipMemcpyAsync(grad, net->grad.ptr, sizeof(float) * net->n_params, hipMemcpyDeviceToHost, st));
    if (outputs)
        T2X_HIP(hipMemcpyAsync(outputs, net->y.ptr, sizeof(float) * (size_t)B * (net->n_bins + 2), hipMemcpyDeviceToHost, st));
    T2X_HIP(hipStreamSynchronize(st));
    if (loss) {
        double s = 0.;
        for (int b = 0; b < B; b++) s += lb[b];
        *loss = s / B;
    }
    return T2X_OK;
}

extern "C" int t2x_proposal_train(t2x_device* h, const t2x_estep_config* cfg, const t2x_model* model,
                                  const t2x_proposal_train_params* tp, t2x_proposal_train_stats* out)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    if (!h || !cfg || !model || !model->sigma2_noise || !tp) { set_error("t2x_proposal_train: null argument"); return T2X_EINVAL; }
    if (tp->steps < 1 || tp->batch < 1 || tp->batch > 4096 || tp->n_projections < 1 || !(tp->max_shift_px > 0.) ||
        !(tp->lr > 0.)) {
        set_error("t2x_proposal_train: bad training parameters");
        return T2X_EINVAL;
    }
    HostSampling S;
    T2X_TRY(build_sampling(cfg, S));
    if (cfg->healpix_order < 1) { set_error("t2x_proposal_train: healpix_order must be >= 1"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    T2X_HIP(hipSetDevice(d->id));
    hipStream_t st = d->stream;
    const int K = cfg->nr_classes;
    const long Nd = (long)S.rot.size();
    const int n_bins = (int)(K * (Nd / 4));
    std::vector<const float2*> hv(K);
    std::vector<int> hp(K);
    int r_max = 0;
    double pf = 0.;
    for (int c = 0; c < K; c++) {
        auto it = d->classes.find(c);
        if (it == d->classes.end() || !it->second.proj) { set_error("t2x_proposal_train: projector of a class not set"); return T2X_ESTATE; }
        hv[c] = it->second.proj;
        hp[c] = it->second.p_pad;
        r_max = (c == 0) ? it->second.p_rmax : std::min(r_max, it->second.p_rmax);
        pf = it->second.p_pf;
    }
    if (!d->net || d->net->n_bins != n_bins) {
        d->net = std::make_shared<Net>();
        T2X_TRY(d->net->init(n_bins, tp->seed ^ 0x5eedULL, st));
    }
    Net& net = *d->net;

    // ---- projection parameter sets and labels (host, once) ----
    const int NP = tp->n_projections;
    std::vector<TrainSample> hs(NP);
    std::vector<double> dx(Nd), dy(Nd), dz(Nd);
    for (long i = 0; i < Nd; i++) {
        const double r = S.rot[i] * kPiP / 180., t = S.tilt[i] * kPiP / 180.;
        dx[i] = std::cos(r) * std::sin(t);
        dy[i] = std::sin(r) * std::sin(t);
        dz[i] = std::cos(t);
    }
    uint64_t ctr = tp->seed * 0x100000001B3ULL + 17;
    auto uni = [&]() {
        uint64_t z = (ctr += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int i = 0; i < NP; i++) {
        TrainSample& s = hs[i];
        s.cls = std::min(K - 1, (int)(uni() * K));
        const double rot = 360. * uni(), tilt = std::acos(2. * uni() - 1.) * 180. / kPiP, psi = 360. * uni();
        double A[9], R[9];
        euler_matrix_host(rot, tilt, psi, A);
        inv3(A, R);
        for (int k = 0; k < 9; k++) s.Ai[k] = (float)(R[k] * pf);
        s.sx = (float)((2. * uni() - 1.) * tp->max_shift_px);
        s.sy = (float)((2. * uni() - 1.) * tp->max_shift_px);
        t2x_ctf_params cp{};
        cp.voltage_kv = tp->voltage > 0 ? tp->voltage : 300.;
        cp.cs_mm = tp->cs > 0 ? tp->cs : 2.7;
        cp.q0 = tp->q0 > 0 ? tp->q0 : 0.1;
        const double dmin = tp->defocus_min > 0 ? tp->defocus_min : 10000., dmax = tp->defocus_max > 0 ? tp->defocus_max : 30000.;
        cp.defocus_u = dmin + (dmax - dmin) * uni();
        cp.defocus_v = dmin + (dmax - dmin) * uni();
        cp.defocus_angle = 180. * uni();
        cp.scale = 1.;
        s.ctf = ctf_const(cp);
        // nearest coarse direction of the projection direction; bin = its parent one order down
        const double r = rot * kPiP / 180., t = tilt * kPiP / 180.;
        const double ux = std::cos(r) * std::sin(t), uy = std::sin(r) * std::sin(t), uz = std::cos(t);
        long best = 0;
        double bd = -2.;
        for (long k = 0; k < Nd; k++) {
            const double dd = ux * dx[k] + uy * dy[k] + uz * dz[k];
            if (dd > bd) { bd = dd; best = k; }
        }
        s.label = (int)(s.cls * (Nd / 4) + (best >> 2));
    }
    DevBuf d_samples, d_vols, d_pads, d_sigma2, d_crop, d_img;
    T2X_TRY(upload_vec(d_samples, hs, st));
    T2X_TRY(upload_vec(d_vols, hv, st));
    T2X_TRY(upload_vec(d_pads, hp, st));
    std::vector<double> sig(model->sigma2_noise, model->sigma2_noise + cfg->ori_size / 2 + 1);
    T2X_TRY(upload_vec(d_sigma2, sig, st));
    const int B = tp->batch;
    T2X_TRY(net.ensure_batch(B));
    T2X_TRY(d_crop.ensure(sizeof(float2) * (size_t)B * kIn * kCropX));
    T2X_HIP(hipStreamSynchronize(st));
    const auto t_setup = clk::now();

    hipfftHandle plan = 0;
    int plan_b = 0;
    const int r_lim = std::min(r_max, kIn / 2);
    const int nlog = std::min(20, tp->steps);
    std::vector<double> lfirst, llast;
    std::vector<int> cor(B), rk(B);
    double loss_first = 0., loss_last = 0., top1 = 0., top16 = 0.;
    const int warm = std::max(1, std::min(100, tp->steps / 10));
    int rc = T2X_OK;
    for (int step = 0; step < tp->steps && rc == T2X_OK; step++) {
        const long long first = (long long)step * B;
        k_train_synth<<<B, 256, 0, st>>>(d_samples.as<TrainSample>(), first, NP, d_vols.as<const float2*>(),
                                         d_pads.as<int>(), r_lim, cfg->ori_size, cfg->pixel_size,
                                         d_sigma2.as<double>(), tp->seed * 31 + 7, d_crop.as<float2>());
        k_train_labels<<<(B + 255) / 256, 256, 0, st>>>(d_samples.as<TrainSample>(), first, NP, B, tp->max_shift_px,
                                                        net.label.as<int>(), net.shift_t.as<float>());
        if (hipGetLastError() != hipSuccess) { set_error("proposal: synthesis launch failed"); rc = T2X_EDEVICE; break; }
        rc = crop_to_input(net, d_crop.as<float2>(), B, st, d_img, plan, plan_b);
        if (rc != T2X_OK) break;
        rc = net.loss_grad(B, tp->bf16 != 0, (float)tp->shift_weight, st);
        if (rc != T2X_OK) break;
        // linear warm-up, then cosine decay to 5% of the peak rate
        double lr = tp->lr;
        if (step < warm) lr *= (double)(step + 1) / warm;
        else lr *= 0.05 + 0.95 * 0.5 * (1. + std::cos(kPiP * (double)(step - warm) / std::max(1, tp->steps - warm)));
        rc = net.adam_step((float)lr, st);
        if (rc != T2X_OK) break;
        if (step < nlog || step >= tp->steps - nlog) {
            std::vector<double> lb(B);
            if (hipMemcpyAsync(lb.data(), net.loss.ptr, sizeof(double) * B, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(cor.data(), net.correct.ptr, sizeof(int) * B, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(rk.data(), net.rank.ptr, sizeof(int) * B, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                set_error("proposal: readback failed");
                rc = T2X_EDEVICE;
                break;
            }
            double s = 0.;
            int c = 0, c16 = 0;
            for (int b = 0; b < B; b++) { s += lb[b]; c += cor[b]; c16 += rk[b] < 16; }
            if (step < nlog) loss_first += s / B / nlog;
            if (step >= tp->steps - nlog) {
                loss_last += s / B / nlog;
                top1 += (double)c / B / nlog;
                top16 += (double)c16 / B / nlog;
            }
        }
    }
    if (plan) hipfftDestroy(plan);
    T2X_TRY(rc);
    T2X_HIP(hipStreamSynchronize(st));
    net.trained = true;
    net.train_ori = cfg->ori_size;
    net.train_K = K;
    net.train_order = cfg->healpix_order;
    net.train_max_shift = tp->max_shift_px;
    if (out) {
        *out = t2x_proposal_train_stats{};
        out->ms_total = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        out->ms_setup = std::chrono::duration<double, std::milli>(t_setup - t0).count();
        out->loss_first = loss_first;
        out->loss_last = loss_last;
        out->top1_last = top1;
        out->top16_last = top16;
        out->steps = tp->steps;
        out->n_bins = n_bins;
    }
    return T2X_OK;
}

extern "C" int t2x_proposal_gate_values(t2x_device* h, int n, float* top1, float* topk_mass)
{
    if (!h || n < 0) { set_error("t2x_proposal_gate_values: bad arguments"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    if ((size_t)n != d->prop_gate_top1.size()) {
        set_error("t2x_proposal_gate_values: n differs from the particle count of the last proposal E-step");
        return T2X_EINVAL;
    }
    if (top1) std::copy(d->prop_gate_top1.begin(), d->prop_gate_top1.end(), top1);
    if (topk_mass) std::copy(d->prop_gate_mass.begin(), d->prop_gate_mass.end(), topk_mass);
    return T2X_OK;
}
