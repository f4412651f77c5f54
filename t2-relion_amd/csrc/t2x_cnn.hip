// t2x_cnn.hip — the pose-proposal network: forward, backward and Adam on the GPU.
//
// Network: 64x64 input, four 3x3 stride-2 pad-1 convolutions (32/64/128/128 channels, ReLU),
// flatten (y, x, c), linear 2048 -> n_bins + 2 (bin logits over classes x HEALPix order-2
// directions, and a 2-D shift regression in units of the maximum shift). Activations are
// channels-last [B][H][W][C] in FP32; weights [Cout][ky][kx][Cin] (FP32 master copy, Adam state).
//
// Every layer pass is one GEMM C[m][n] = sum_k A(m,k) B(n,k) of a single kernel template whose
// operands are either strided views of FP32 arrays or implicit im2col views of an activation
// tensor (no im2col buffer for the forward pass or the weight gradients). Operands are rounded
// to BF16 in LDS and multiplied on v_mfma_f32_16x16x32_bf16 with FP32 accumulation (BF16
// weights, FP32 accumulate); an FP32 operand mode (v_mfma_f32_16x16x4f32) exists for gradient
// checks against the float64 oracle. Epilogues fuse bias and ReLU.
#include "t2x_cnn.h"

#include <cmath>
#include <vector>

namespace t2x {
namespace cnn {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

enum { OP_STRIDED = 0, OP_IM2COL = 1, OP_IM2COL_T = 2 };
enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_RELU = 2 };

// An operand view X(r, k), r the output row (A: m, B: n), k the reduction index.
struct Operand {
    int mode;
    const float* p;
    long long sr, sk;                 // OP_STRIDED: p[r*sr + k*sk]
    int H, W, C, Ho, Wo;              // im2col: activation [Bn][H][W][C], output grid Ho x Wo
    int ones_r;                       // >= 0: row ones_r reads 1 (bias gradients as an extra column)
};

struct GemmArgs {
    int M, N, K;
    Operand a, b;
    float* c;
    long long ldc;
    const float* bias;
    int extra_col;                    // >= 0: output column extra_col goes to c_extra[m] instead
    float* c_extra;
};

// Operand tile loads. Each thread fills 8 entries of a 64 x 32 (row x k) tile:
//  row-major form (A = strided with unit k stride, or im2col): thread = (row, 8 consecutive k);
//    the im2col row (b, y, x) is decoded once and (ky, kx, ci) advanced incrementally.
//  k-major form (strided with unit row stride, or im2col transposed): thread = (k chunk, row),
//    lanes run along the row index so the loads coalesce.
#define CT 64
#define CK 32
#define CLD 33   // LDS row stride (floats): conflict-free row-wise writes and fragment reads

__device__ __forceinline__ bool row_major_form(const Operand& o)
{
    return o.mode == OP_IM2COL || (o.mode == OP_STRIDED && o.sk == 1);
}

template <class LT, int LD>
__device__ __forceinline__ void load_tile(const Operand& o, int R, int K, int r0, int k0, LT (*S)[LD], int tid)
{
    if (row_major_form(o)) {
        const int lr = tid >> 2, lk = (tid & 3) * 8;
        const int r = r0 + lr;
        float v[8];
        if (r == o.ones_r) {
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = (k0 + lk + q < K) ? 1.f : 0.f;
        } else if (o.mode == OP_STRIDED) {
            const float* p = o.p + (long long)r * o.sr;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int k = k0 + lk + q;
                v[q] = (r < R && k < K) ? p[k] : 0.f;
            }
        } else {
            // im2col: r = pixel (b, y, x) of the output grid, k = (ky, kx, ci)
            const int x = r % o.Wo, t = r / o.Wo, y = t % o.Ho, b = t / o.Ho;
            int k = k0 + lk;
            int ci = k % o.C, kk = k / o.C;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                float val = 0.f;
                if (r < R && k < K) {
                    const int ky = kk / 3, kx = kk - 3 * ky;
                    const int yy = 2 * y + ky - 1, xx = 2 * x + kx - 1;
                    if (yy >= 0 && yy < o.H && xx >= 0 && xx < o.W)
                        val = o.p[(((long long)b * o.H + yy) * o.W + xx) * o.C + ci];
                }
                v[q] = val;
                k++;
                if (++ci == o.C) { ci = 0; kk++; }
            }
        }
        if constexpr (sizeof(LT) == 2) {
            bf16x8 pk;
#pragma unroll
            for (int q = 0; q < 8; q++) pk[q] = (__bf16)v[q];
            *reinterpret_cast<bf16x8*>(&S[lr][lk]) = pk;   // one 16-byte store
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++) S[lr][lk + q] = v[q];
        }
    } else {
        const int lr = tid & 63, kc = (tid >> 6) * 8;
        const int r = r0 + lr;
        if (r == o.ones_r) {
#pragma unroll
            for (int q = 0; q < 8; q++) S[lr][kc + q] = (LT)((k0 + kc + q < K) ? 1.f : 0.f);
        } else if (o.mode == OP_STRIDED) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int k = k0 + kc + q;
                S[lr][kc + q] = (LT)((r < R && k < K) ? o.p[(long long)r * o.sr + (long long)k * o.sk] : 0.f);
            }
        } else {
            // im2col transposed: r = column (ky, kx, ci), k = pixel (b, y, x)
            const int ci = r % o.C, kk = r / o.C, ky = kk / 3, kx = kk - 3 * ky;
            int k = k0 + kc;
            int x = k % o.Wo, t = k / o.Wo, y = t % o.Ho, b = t / o.Ho;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                float val = 0.f;
                if (r < R && k < K) {
                    const int yy = 2 * y + ky - 1, xx = 2 * x + kx - 1;
                    if (yy >= 0 && yy < o.H && xx >= 0 && xx < o.W)
                        val = o.p[(((long long)b * o.H + yy) * o.W + xx) * o.C + ci];
                }
                S[lr][kc + q] = (LT)val;
                k++;
                if (++x == o.Wo) { x = 0; if (++y == o.Ho) { y = 0; b++; } }
            }
        }
    }
}

// C (+ blockIdx.z * split_stride for split-K partials) = A(m, k) B(n, k) over k in this block's range
// LDS tiles hold BF16 (row stride 40: 16-byte aligned fragment reads) or FP32 (stride 33)
template <bool BF16> struct LdsT;
template <> struct LdsT<true> { typedef __bf16 type; static constexpr int ld = 40; };
template <> struct LdsT<false> { typedef float type; static constexpr int ld = CLD; };

template <bool BF16, int EPI>
__global__ __launch_bounds__(256) void k_cnn_gemm(GemmArgs g, int k_chunk, long long split_stride)
{
    typedef typename LdsT<BF16>::type LT;
    constexpr int LD = LdsT<BF16>::ld;
    __shared__ __attribute__((aligned(16))) LT As[CT][LD];
    __shared__ __attribute__((aligned(16))) LT Bs[CT][LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.x * CT, n0 = blockIdx.y * CT;
    const int kb = blockIdx.z * k_chunk, ke = min(g.K, kb + k_chunk);
    floatx4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += CK) {
        __syncthreads();
        load_tile(g.a, g.M, ke, m0, k0, As, tid);
        load_tile(g.b, g.N, ke, n0, k0, Bs, tid);
        __syncthreads();
        if constexpr (BF16) {
            const int kc = (lane >> 4) * 8;
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + i * 16 + (lane & 15)][kc]);
#pragma unroll
            for (int j = 0; j < 2; j++) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn * 32 + j * 16 + (lane & 15)][kc]);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        } else {
#pragma unroll
            for (int s = 0; s < CK / 4; s++) {
                const int k = 4 * s + (lane >> 4);
                float a[2], b[2];   // FP32 operand mode
#pragma unroll
                for (int i = 0; i < 2; i++) a[i] = As[wm * 32 + i * 16 + (lane & 15)][k];
#pragma unroll
                for (int j = 0; j < 2; j++) b[j] = Bs[wn * 32 + j * 16 + (lane & 15)][k];
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
    }
    float* c = g.c + blockIdx.z * split_stride;
    const int col = lane & 15, rg = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm * 32 + i * 16 + rg + r;
                const int n = n0 + wn * 32 + j * 16 + col;
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][r];
                if (EPI != EPI_STORE) v += g.bias[n];
                if (EPI == EPI_BIAS_RELU) v = fmaxf(v, 0.f);
                if (split_stride == 0 && n == g.extra_col) g.c_extra[m] = v;
                else c[(long long)m * (split_stride ? g.N : g.ldc) + n] = v;
            }
}

// out[m][n] = sum over the split-K partials in order
__global__ void k_splitk_reduce(const float* __restrict__ part, int S, long long stride, int M, int N, long long ldc,
                                float* __restrict__ out, int extra_col, float* __restrict__ out_extra)
{
    const long long n_el = (long long)M * N;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n_el; e += (long long)gridDim.x * blockDim.x) {
        const long long m = e / N, n = e - m * N;
        float s = 0.f;
        for (int z = 0; z < S; z++) s += part[z * stride + m * N + n];
        if (n == extra_col) out_extra[m] = s;
        else out[m * ldc + n] = s;
    }
}

// dz = da * (out > 0)
__global__ void k_relu_bwd(const float* __restrict__ da, const float* __restrict__ out, float* __restrict__ dz,
                           long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dz[i] = out[i] > 0.f ? da[i] : 0.f;
}

// dact[b][yi][xi][ci] = sum over the (<= 4) output pixels whose 3x3 window covers (yi, xi)
__global__ void k_col2im(const float* __restrict__ dcols, int Bn, int H, int W, int C, float* __restrict__ dact)
{
    const int Ho = H / 2, Wo = W / 2;
    const long long n = (long long)Bn * H * W * C;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int ci = (int)(e % C);
        const long long pix = e / C;
        const int xi = (int)(pix % W), yi = (int)((pix / W) % H);
        const long long b = pix / ((long long)W * H);
        float s = 0.f;
        for (int ky = 0; ky < 3; ky++) {
            const int t = yi + 1 - ky;
            if (t < 0 || (t & 1)) continue;
            const int y = t >> 1;
            if (y >= Ho) continue;
            for (int kx = 0; kx < 3; kx++) {
                const int u = xi + 1 - kx;
                if (u < 0 || (u & 1)) continue;
                const int x = u >> 1;
                if (x >= Wo) continue;
                s += dcols[((b * Ho + y) * Wo + x) * (9LL * C) + (ky * 3 + kx) * C + ci];
            }
        }
        dact[e] = s;
    }
}

// softmax cross-entropy on the bins + shift_weight * squared shift error; dy = dloss/dy / B.
// One workgroup per sample; loss per sample into loss_out (summed on the host in order).
__global__ void k_cnn_loss(const float* __restrict__ y, int nbins, const int* __restrict__ label,
                           const float* __restrict__ shift_t, float shift_weight, int Bn, float* __restrict__ dy,
                           double* __restrict__ loss_out, int* __restrict__ correct, int* __restrict__ rank)
{
    const int b = blockIdx.x;
    const int ld = nbins + 2;
    const float* yb = y + (size_t)b * ld;
    __shared__ float red[256];
    __shared__ int redi[256];
    float mx = -3.0e38f;
    int arg = 0;
    for (int i = threadIdx.x; i < nbins; i += blockDim.x)
        if (yb[i] > mx) { mx = yb[i]; arg = i; }
    red[threadIdx.x] = mx;
    redi[threadIdx.x] = arg;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const float o = red[threadIdx.x + s];
            const int oi = redi[threadIdx.x + s];
            if (o > red[threadIdx.x] || (o == red[threadIdx.x] && oi < redi[threadIdx.x])) {
                red[threadIdx.x] = o;
                redi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const float m = red[0];
    const int am = redi[0];
    __syncthreads();
    float se = 0.f;
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) se += expf(yb[i] - m);
    red[threadIdx.x] = se;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const float Z = red[0];
    const int lab = label[b];
    __syncthreads();
    int above = 0;   // logits strictly above the label's: its rank
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) above += yb[i] > yb[lab];
    redi[threadIdx.x] = above;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) redi[threadIdx.x] += redi[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && rank) rank[b] = redi[0];
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) {
        const float p = expf(yb[i] - m) / Z;
        dy[(size_t)b * ld + i] = (p - (i == lab ? 1.f : 0.f)) / (float)Bn;
    }
    if (threadIdx.x < 2) {
        const float d = yb[nbins + threadIdx.x] - shift_t[2 * b + threadIdx.x];
        dy[(size_t)b * ld + nbins + threadIdx.x] = 2.f * shift_weight * d / (float)Bn;
    }
    if (threadIdx.x == 0) {
        const double dx = (double)yb[nbins] - shift_t[2 * b], dyv = (double)yb[nbins + 1] - shift_t[2 * b + 1];
        loss_out[b] = -((double)yb[lab] - m - log((double)Z)) + shift_weight * (dx * dx + dyv * dyv);
        if (correct) correct[b] = (am == lab);
    }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long long n, float lr, float b1, float b2, float eps, float c1,
                       float c2)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = v;
        p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
    }
}

// ---------------------------------------------------------------------------------------------

static unsigned grid_for(long long n, int bs = 256)
{
    long long g = (n + bs - 1) / bs;
    return (unsigned)std::min<long long>(std::max<long long>(g, 1), 65536);
}

// One layer GEMM. Outputs with few tiles and a long reduction (the weight gradients: M x N of
// at most 770 x 2048 against K up to B*32*32 pixels) are split over K into partial products that
// a second kernel sums in a fixed order; `scratch` holds the partials.
int gemm(bool bf16, int epi, const GemmArgs& a, hipStream_t st, DevBuf* scratch)
{
    if (a.M <= 0 || a.N <= 0 || a.K <= 0) return T2X_OK;
    const unsigned gx = (unsigned)((a.M + CT - 1) / CT), gy = (unsigned)((a.N + CT - 1) / CT);
    if (gy > 65535) { set_error("cnn gemm: N too large"); return T2X_EINVAL; }
    const long long tiles = (long long)gx * gy;
    int S = 1;
    if (epi == EPI_STORE && scratch && tiles < 256 && a.K >= 4 * 1024) {
        S = (int)std::min<long long>((512 + tiles - 1) / tiles, a.K / 1024);
        S = std::max(1, std::min(S, 1024));
    }
    int k_chunk = a.K;
    if (S > 1) k_chunk = (int)(((long long)(a.K + S - 1) / S + CK - 1) / CK * CK);
    S = (a.K + k_chunk - 1) / k_chunk;
    GemmArgs g = a;
    long long stride = 0;
    if (S > 1) {
        stride = (long long)a.M * a.N;
        T2X_TRY(scratch->ensure(sizeof(float) * (size_t)stride * S));
        g.c = scratch->as<float>();
        g.ldc = a.N;
        g.extra_col = -1;   // the reduction routes the extra column
    }
    const dim3 grid(gx, gy, (unsigned)S);
    if (bf16) {
        if (epi == EPI_STORE) k_cnn_gemm<true, EPI_STORE><<<grid, 256, 0, st>>>(g, k_chunk, stride);
        else if (epi == EPI_BIAS) k_cnn_gemm<true, EPI_BIAS><<<grid, 256, 0, st>>>(g, k_chunk, stride);
        else k_cnn_gemm<true, EPI_BIAS_RELU><<<grid, 256, 0, st>>>(g, k_chunk, stride);
    } else {
        if (epi == EPI_STORE) k_cnn_gemm<false, EPI_STORE><<<grid, 256, 0, st>>>(g, k_chunk, stride);
        else if (epi == EPI_BIAS) k_cnn_gemm<false, EPI_BIAS><<<grid, 256, 0, st>>>(g, k_chunk, stride);
        else k_cnn_gemm<false, EPI_BIAS_RELU><<<grid, 256, 0, st>>>(g, k_chunk, stride);
    }
    T2X_HIP(hipGetLastError());
    if (S > 1) {
        k_splitk_reduce<<<grid_for(stride), 256, 0, st>>>(scratch->as<float>(), S, stride, a.M, a.N, a.ldc, a.c,
                                                          a.extra_col, a.c_extra);
        T2X_HIP(hipGetLastError());
    }
    return T2X_OK;
}

static Operand strided(const float* p, long long sr, long long sk, int ones_r = -1)
{
    Operand o{};
    o.ones_r = ones_r;
    o.mode = OP_STRIDED;
    o.p = p;
    o.sr = sr;
    o.sk = sk;
    return o;
}

static Operand im2col_view(int mode, const float* act, int H, int W, int C, int ones_r = -1)
{
    Operand o{};
    o.ones_r = ones_r;
    o.mode = mode;
    o.p = act;
    o.H = H;
    o.W = W;
    o.C = C;
    o.Ho = H / 2;
    o.Wo = W / 2;
    return o;
}

void Net::layout(int n_bins_)
{
    n_bins = n_bins_;
    long long off = 0;
    for (int l = 0; l < 4; l++) {
        w_off[l] = off;
        off += (long long)kChannels[l + 1] * 9 * kChannels[l];
        b_off[l] = off;
        off += kChannels[l + 1];
    }
    wh_off = off;
    off += (long long)(n_bins + 2) * kFeat;
    bh_off = off;
    off += n_bins + 2;
    n_params = off;
}

int Net::ensure_batch(int B)
{
    if (B <= cap) return T2X_OK;
    long long a = 0;
    for (int l = 1; l <= 4; l++) {
        const int s = kIn >> l;
        a += (long long)B * s * s * kChannels[l];
    }
    (void)a;
    for (int l = 0; l < 5; l++) {
        const int s = kIn >> l;
        T2X_TRY(act[l].ensure(sizeof(float) * (size_t)B * s * s * kChannels[l]));
        T2X_TRY(dact[l].ensure(sizeof(float) * (size_t)B * s * s * kChannels[l]));
    }
    size_t dc = 0;
    for (int l = 1; l < 4; l++) {   // dcols of layers 1..3 (input gradients of layers 2..4)
        const int s = kIn >> (l + 1);
        dc = std::max(dc, sizeof(float) * (size_t)B * s * s * 9 * kChannels[l]);
    }
    T2X_TRY(dcols.ensure(dc));
    T2X_TRY(dz.ensure(sizeof(float) * (size_t)B * (kIn / 2) * (kIn / 2) * kChannels[1]));
    T2X_TRY(y.ensure(sizeof(float) * (size_t)B * (n_bins + 2)));
    T2X_TRY(dy.ensure(sizeof(float) * (size_t)B * (n_bins + 2)));
    T2X_TRY(loss.ensure(sizeof(double) * B));
    T2X_TRY(correct.ensure(sizeof(int) * B));
    T2X_TRY(rank.ensure(sizeof(int) * B));
    T2X_TRY(label.ensure(sizeof(int) * B));
    T2X_TRY(shift_t.ensure(sizeof(float) * 2 * B));
    cap = B;
    return T2X_OK;
}

// act[0] holds the input [B][64][64][1]; writes act[1..4] and y
int Net::forward(int B, bool bf16, hipStream_t st)
{
    const float* P = params.as<float>();
    for (int l = 0; l < 4; l++) {
        const int H = kIn >> l;
        GemmArgs g{};
        g.extra_col = -1;
        g.M = B * (H / 2) * (H / 2);
        g.N = kChannels[l + 1];
        g.K = 9 * kChannels[l];
        g.a = im2col_view(OP_IM2COL, act[l].as<float>(), H, H, kChannels[l]);
        g.b = strided(P + w_off[l], g.K, 1);
        g.c = act[l + 1].as<float>();
        g.ldc = g.N;
        g.bias = P + b_off[l];
        T2X_TRY(gemm(bf16, EPI_BIAS_RELU, g, st, nullptr));
    }
    GemmArgs h{};
    h.extra_col = -1;
    h.M = B;
    h.N = n_bins + 2;
    h.K = kFeat;
    h.a = strided(act[4].as<float>(), kFeat, 1);
    h.b = strided(P + wh_off, kFeat, 1);
    h.c = y.as<float>();
    h.ldc = h.N;
    h.bias = P + bh_off;
    return gemm(bf16, EPI_BIAS, h, st, nullptr);
}


// dy must hold dloss/dy; fills grad (n_params floats)
int Net::backward(int B, bool bf16, hipStream_t st)
{
    const float* P = params.as<float>();
    float* G = grad.as<float>();
    const int ldy = n_bins + 2;
    // head
    {
        GemmArgs g{};   // dWh[o][k] = sum_b dy[b][o] flat[b][k]; column kFeat (ones) gives dbh
        g.M = ldy;
        g.N = kFeat + 1;
        g.K = B;
        g.a = strided(dy.as<float>(), 1, ldy);
        g.b = strided(act[4].as<float>(), 1, kFeat, kFeat);
        g.c = G + wh_off;
        g.ldc = kFeat;
        g.extra_col = kFeat;
        g.c_extra = G + bh_off;
        T2X_TRY(gemm(bf16, EPI_STORE, g, st, &splitk));
        GemmArgs d{};   // dflat[b][k] = sum_o dy[b][o] Wh[o][k]
        d.extra_col = -1;
        d.M = B;
        d.N = kFeat;
        d.K = ldy;
        d.a = strided(dy.as<float>(), ldy, 1);
        d.b = strided(P + wh_off, 1, kFeat);
        d.c = dact[4].as<float>();
        d.ldc = kFeat;
        T2X_TRY(gemm(bf16, EPI_STORE, d, st, &splitk));
    }
    for (int l = 3; l >= 0; l--) {
        const int H = kIn >> l, Ho = H / 2;
        const long long M = (long long)B * Ho * Ho;
        const int Co = kChannels[l + 1], Ci = kChannels[l], K = 9 * Ci;
        float* dzp = dz.as<float>();
        k_relu_bwd<<<grid_for(M * Co), 256, 0, st>>>(dact[l + 1].as<float>(), act[l + 1].as<float>(), dzp, M * Co);
        T2X_HIP(hipGetLastError());
        GemmArgs g{};   // dW[co][q] = sum_p dz[p][co] im2col(act_l)[p][q]; column K (ones) gives db
        g.M = Co;
        g.N = K + 1;
        g.K = (int)M;
        g.a = strided(dzp, 1, Co);
        g.b = im2col_view(OP_IM2COL_T, act[l].as<float>(), H, H, Ci, K);
        g.c = G + w_off[l];
        g.ldc = K;
        g.extra_col = K;
        g.c_extra = G + b_off[l];
        T2X_TRY(gemm(bf16, EPI_STORE, g, st, &splitk));
        if (l > 0) {
            GemmArgs d{};   // dcols[p][q] = sum_co dz[p][co] W[co][q]
            d.extra_col = -1;
            d.M = (int)M;
            d.N = K;
            d.K = Co;
            d.a = strided(dzp, Co, 1);
            d.b = strided(P + w_off[l], 1, K);
            d.c = dcols.as<float>();
            d.ldc = K;
            T2X_TRY(gemm(bf16, EPI_STORE, d, st, &splitk));
            k_col2im<<<grid_for((long long)B * H * H * Ci), 256, 0, st>>>(dcols.as<float>(), B, H, H, Ci,
                                                                          dact[l].as<float>());
            T2X_HIP(hipGetLastError());
        }
    }
    return T2X_OK;
}

int Net::loss_grad(int B, bool bf16, float shift_weight, hipStream_t st)
{
    T2X_TRY(forward(B, bf16, st));
    k_cnn_loss<<<B, 256, 0, st>>>(y.as<float>(), n_bins, label.as<int>(), shift_t.as<float>(), shift_weight, B,
                                  dy.as<float>(), loss.as<double>(), correct.as<int>(), rank.as<int>());
    T2X_HIP(hipGetLastError());
    return backward(B, bf16, st);
}

int Net::adam_step(float lr, hipStream_t st)
{
    adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
    k_adam<<<grid_for(n_params), 256, 0, st>>>(params.as<float>(), grad.as<float>(), m1.as<float>(), m2.as<float>(),
                                                n_params, lr, b1, b2, eps, c1, c2);
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

// He-uniform initialisation from a counter-based generator (host), zero biases
int Net::init(int n_bins_, uint64_t seed, hipStream_t st)
{
    layout(n_bins_);
    std::vector<float> h((size_t)n_params, 0.f);
    auto uni = [&](uint64_t i) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (i + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < 4; l++) {
        const long long n = (long long)kChannels[l + 1] * 9 * kChannels[l];
        const double lim = std::sqrt(6.0 / (9.0 * kChannels[l]));
        for (long long i = 0; i < n; i++) h[w_off[l] + i] = (float)((2.0 * uni(w_off[l] + i) - 1.0) * lim);
    }
    const double limh = std::sqrt(6.0 / kFeat);
    for (long long i = 0; i < (long long)(n_bins + 2) * kFeat; i++)
        h[wh_off + i] = (float)((2.0 * uni(wh_off + i) - 1.0) * limh);
    T2X_TRY(params.ensure(sizeof(float) * n_params));
    T2X_TRY(grad.ensure(sizeof(float) * n_params));
    T2X_TRY(m1.ensure(sizeof(float) * n_params));
    T2X_TRY(m2.ensure(sizeof(float) * n_params));
    T2X_HIP(hipMemcpyAsync(params.ptr, h.data(), sizeof(float) * n_params, hipMemcpyHostToDevice, st));
    T2X_HIP(hipMemsetAsync(m1.ptr, 0, sizeof(float) * n_params, st));
    T2X_HIP(hipMemsetAsync(m2.ptr, 0, sizeof(float) * n_params, st));
    adam_t = 0;
    trained = false;
    return T2X_OK;
}

}  // namespace cnn
}  // namespace t2x
