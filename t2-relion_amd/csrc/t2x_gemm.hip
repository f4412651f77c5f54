// t2x_gemm.hip — FP32 MFMA GEMM for the coarse-pass cross term.
//
// C[m][n] = sum_k A[m][k] * B[n][k] (both operands K-contiguous). The coarse diff2 of
// getAllSquaredDifferences (ml_optimiser.cpp:7265-7273) is recast as
//   diff2[b,t,n] = C_b + N1[b,n] - sum_k A[(b,t),k] B[n,k]
// with rows (particle, translation), columns (class, orientation) and k = (pixel, re/im).
// Tiles: 128x128 per workgroup, 4 waves of 64x64, K step 16, v_mfma_f32_16x16x4f32.
// Workgroups are remapped so that each XCD gets a contiguous range of tiles and walks the M
// dimension fastest inside a column group, keeping B tiles hot in that XCD's L2.
#include "t2x_kernels.h"

namespace t2x {

enum { EPI_STORE = 0, EPI_DIFF2 = 1 };

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
};

#define GBM 128
#define GBN 128
#define GBK 16
#define GPAD 4

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_tn(const float* __restrict__ A, long long lda,
                                                 const float* __restrict__ B, long long ldb,
                                                 int K, int tiles_m, int tiles_n, EpiParams ep)
{
    __shared__ float As[GBK][GBM + GPAD];
    __shared__ float Bs[GBK][GBN + GPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    // XCD-aware tile order: workgroup ids are dealt round-robin to 8 XCDs
    const int nwg = tiles_m * tiles_n;
    int pid = blockIdx.x;
    if ((nwg & 7) == 0) {   // bijective only when the grid divides evenly over the 8 XCDs
        const int per = nwg / 8;
        pid = (pid & 7) * per + (pid >> 3);
    }
    // within the linear order: groups of 8 tile-columns, M fastest
    const int group = 8;
    const int gid = pid / (group * tiles_m);
    const int first_n = gid * group;
    const int gsize = min(tiles_n - first_n, group);
    const int in_group = pid % (group * tiles_m);
    const int tm = in_group % tiles_m;
    const int tn = first_n + in_group / tiles_m;
    if (tn >= tiles_n || gsize <= 0) return;
    const long long m0 = (long long)tm * GBM, n0 = (long long)tn * GBN;

    floatx4 acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // loader mapping: 128 rows x 16 k = 512 float4; 2 per thread
    const int lrow = tid >> 2, lk4 = (tid & 3) * 4;
    const float* Ap0 = A + (m0 + lrow) * lda + lk4;
    const float* Ap1 = A + (m0 + lrow + 64) * lda + lk4;
    const float* Bp0 = B + (n0 + lrow) * ldb + lk4;
    const float* Bp1 = B + (n0 + lrow + 64) * ldb + lk4;

    float4 ra0 = *reinterpret_cast<const float4*>(Ap0);
    float4 ra1 = *reinterpret_cast<const float4*>(Ap1);
    float4 rb0 = *reinterpret_cast<const float4*>(Bp0);
    float4 rb1 = *reinterpret_cast<const float4*>(Bp1);

    for (int k0 = 0; k0 < K; k0 += GBK) {
        __syncthreads();
        As[lk4 + 0][lrow] = ra0.x; As[lk4 + 1][lrow] = ra0.y; As[lk4 + 2][lrow] = ra0.z; As[lk4 + 3][lrow] = ra0.w;
        As[lk4 + 0][lrow + 64] = ra1.x; As[lk4 + 1][lrow + 64] = ra1.y; As[lk4 + 2][lrow + 64] = ra1.z; As[lk4 + 3][lrow + 64] = ra1.w;
        Bs[lk4 + 0][lrow] = rb0.x; Bs[lk4 + 1][lrow] = rb0.y; Bs[lk4 + 2][lrow] = rb0.z; Bs[lk4 + 3][lrow] = rb0.w;
        Bs[lk4 + 0][lrow + 64] = rb1.x; Bs[lk4 + 1][lrow + 64] = rb1.y; Bs[lk4 + 2][lrow + 64] = rb1.z; Bs[lk4 + 3][lrow + 64] = rb1.w;
        __syncthreads();
        if (k0 + GBK < K) {   // prefetch the next K slice into registers while the MFMAs run
            ra0 = *reinterpret_cast<const float4*>(Ap0 + k0 + GBK);
            ra1 = *reinterpret_cast<const float4*>(Ap1 + k0 + GBK);
            rb0 = *reinterpret_cast<const float4*>(Bp0 + k0 + GBK);
            rb1 = *reinterpret_cast<const float4*>(Bp1 + k0 + GBK);
        }
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            const int k = kk + (lane >> 4);
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[k][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = Bs[k][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }

    // epilogue: lane holds C[4*(lane/16)+r][lane%16] of each 16x16 sub-tile
    const int col_in = lane & 15, row_grp = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const long long m = m0 + wm * 64 + i * 16 + row_grp + r;
            float rmin = 3.0e38f;
            const bool mvalid = m < ep.M;
            int bidx = 0;
            double cb = 0.;
            if (EPI == EPI_DIFF2 && mvalid) {
                bidx = (int)(m / ep.T);
                cb = ep.Cb[bidx];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const long long n = n0 + wn * 64 + j * 16 + col_in;
                float v = acc[i][j][r];
                if (mvalid && n < ep.N) {
                    if (EPI == EPI_STORE) {
                        ep.C[m * ep.ldc + n] = v;
                    } else {
                        float d = (float)(cb + (double)ep.N1[(long long)bidx * ep.ldn + n] - (double)v);
                        ep.C[m * ep.ldc + n] = d;
                        if (ep.prior_o[n] > 0.) rmin = fminf(rmin, d);
                    }
                }
            }
            if (EPI == EPI_DIFF2) {
                // min over the 16 lanes sharing this row (same lane>>4 group)
                for (int o = 1; o < 16; o <<= 1) rmin = fminf(rmin, __shfl_xor(rmin, o));
                if (col_in == 0 && mvalid)
                    ep.rowmin[m * ep.rowmin_ld + (n0 + wn * 64) / 64] = rmin;
            }
        }
    }
}

template __global__ void k_gemm_tn<EPI_STORE>(const float*, long long, const float*, long long, int, int, int, EpiParams);
template __global__ void k_gemm_tn<EPI_DIFF2>(const float*, long long, const float*, long long, int, int, int, EpiParams);

// Host launcher. A must have tiles_m*128 rows, B tiles_n*128 rows, both K (multiple of 16)
// columns, zero padded.
int launch_gemm(int epi, const float* A, long long lda, const float* B, long long ldb, int K,
                int Mrows_pad, int Nrows_pad, const EpiParams& ep, hipStream_t st)
{
    if (K % GBK != 0 || Mrows_pad % GBM != 0 || Nrows_pad % GBN != 0) {
        set_error("gemm: operands not padded to tile multiples");
        return T2X_EINVAL;
    }
    const int tm = Mrows_pad / GBM, tn = Nrows_pad / GBN;
    const long long nwg = (long long)tm * tn;
    if (nwg > 0x7fffffffLL) { set_error("gemm: grid too large"); return T2X_EINVAL; }
    if (epi == EPI_STORE)
        k_gemm_tn<EPI_STORE><<<(unsigned)nwg, 256, 0, st>>>(A, lda, B, ldb, K, tm, tn, ep);
    else
        k_gemm_tn<EPI_DIFF2><<<(unsigned)nwg, 256, 0, st>>>(A, lda, B, ldb, K, tm, tn, ep);
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

}  // namespace t2x

namespace t2x {

// ------------------------------------------------------------------------------------------
// Split-precision cross term: A = Ah + Al, B = Bh + Bl in BF16 or FP16 (hi = round(x),
// lo = round(x - hi)), C = Ah Bh + Ah Bl + Al Bh on v_mfma_f32_16x16x32_{bf16,f16} with FP32
// accumulation (the Al Bl term is below the kept precision). FP16 operands carry a power-of-two
// scale per row (rowscale), applied back in the epilogue. Tiles 128x128, K step 32.
// ------------------------------------------------------------------------------------------

template <class H> struct Vec8;
template <> struct Vec8<__bf16> { typedef __bf16 type __attribute__((ext_vector_type(8))); };
template <> struct Vec8<_Float16> { typedef _Float16 type __attribute__((ext_vector_type(8))); };

template <class H>
__device__ __forceinline__ floatx4 mfma16x16x32(typename Vec8<H>::type a, typename Vec8<H>::type b, floatx4 c);
template <>
__device__ __forceinline__ floatx4 mfma16x16x32<__bf16>(Vec8<__bf16>::type a, Vec8<__bf16>::type b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ floatx4 mfma16x16x32<_Float16>(Vec8<_Float16>::type a, Vec8<_Float16>::type b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

struct SplitOperand {
    const void* hi;        // [rows][K] 16-bit
    const void* lo;
    const float* inv_scale;   // [rows] 1/scale, or null (scale 1)
    long long ld;          // elements per row
};

#define XK 32
#define XPAD 8   // row stride 40 halfs = 80 bytes: conflict-free 16-byte reads for 16 lanes

template <class H, int EPI>
__global__ __launch_bounds__(256) void k_gemm_x3(SplitOperand A, SplitOperand B, int K, int tiles_m,
                                                 int tiles_n, EpiParams ep)
{
    typedef typename Vec8<H>::type v8;
    __shared__ __attribute__((aligned(16))) H Ash[GBM][XK + XPAD];
    __shared__ __attribute__((aligned(16))) H Asl[GBM][XK + XPAD];
    __shared__ __attribute__((aligned(16))) H Bsh[GBN][XK + XPAD];
    __shared__ __attribute__((aligned(16))) H Bsl[GBN][XK + XPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int nwg = tiles_m * tiles_n;
    int pid = blockIdx.x;
    if ((nwg & 7) == 0) {
        const int per = nwg / 8;
        pid = (pid & 7) * per + (pid >> 3);
    }
    const int group = 8;
    const int gid = pid / (group * tiles_m);
    const int first_n = gid * group;
    const int in_group = pid % (group * tiles_m);
    const int tm = in_group % tiles_m;
    const int tn = first_n + in_group / tiles_m;
    if (tn >= tiles_n) return;
    const long long m0 = (long long)tm * GBM, n0 = (long long)tn * GBN;

    floatx4 acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // loader: each operand tile is 128 rows x 32 halfs = 512 chunks of 16 bytes; 2 per thread
    const int lrow = tid >> 2, lc = (tid & 3) * 8;
    const H* Ah = reinterpret_cast<const H*>(A.hi);
    const H* Al = reinterpret_cast<const H*>(A.lo);
    const H* Bh = reinterpret_cast<const H*>(B.hi);
    const H* Bl = reinterpret_cast<const H*>(B.lo);
    const size_t ra0 = (size_t)(m0 + lrow) * A.ld + lc, ra1 = (size_t)(m0 + lrow + 64) * A.ld + lc;
    const size_t rb0 = (size_t)(n0 + lrow) * B.ld + lc, rb1 = (size_t)(n0 + lrow + 64) * B.ld + lc;
    uint4 pa0h = *reinterpret_cast<const uint4*>(Ah + ra0), pa1h = *reinterpret_cast<const uint4*>(Ah + ra1);
    uint4 pa0l = *reinterpret_cast<const uint4*>(Al + ra0), pa1l = *reinterpret_cast<const uint4*>(Al + ra1);
    uint4 pb0h = *reinterpret_cast<const uint4*>(Bh + rb0), pb1h = *reinterpret_cast<const uint4*>(Bh + rb1);
    uint4 pb0l = *reinterpret_cast<const uint4*>(Bl + rb0), pb1l = *reinterpret_cast<const uint4*>(Bl + rb1);

    for (int k0 = 0; k0 < K; k0 += XK) {
        __syncthreads();
        *reinterpret_cast<uint4*>(&Ash[lrow][lc]) = pa0h;
        *reinterpret_cast<uint4*>(&Ash[lrow + 64][lc]) = pa1h;
        *reinterpret_cast<uint4*>(&Asl[lrow][lc]) = pa0l;
        *reinterpret_cast<uint4*>(&Asl[lrow + 64][lc]) = pa1l;
        *reinterpret_cast<uint4*>(&Bsh[lrow][lc]) = pb0h;
        *reinterpret_cast<uint4*>(&Bsh[lrow + 64][lc]) = pb1h;
        *reinterpret_cast<uint4*>(&Bsl[lrow][lc]) = pb0l;
        *reinterpret_cast<uint4*>(&Bsl[lrow + 64][lc]) = pb1l;
        __syncthreads();
        if (k0 + XK < K) {
            const int kn = k0 + XK;
            pa0h = *reinterpret_cast<const uint4*>(Ah + ra0 + kn);
            pa1h = *reinterpret_cast<const uint4*>(Ah + ra1 + kn);
            pa0l = *reinterpret_cast<const uint4*>(Al + ra0 + kn);
            pa1l = *reinterpret_cast<const uint4*>(Al + ra1 + kn);
            pb0h = *reinterpret_cast<const uint4*>(Bh + rb0 + kn);
            pb1h = *reinterpret_cast<const uint4*>(Bh + rb1 + kn);
            pb0l = *reinterpret_cast<const uint4*>(Bl + rb0 + kn);
            pb1l = *reinterpret_cast<const uint4*>(Bl + rb1 + kn);
        }
        const int kc = (lane >> 4) * 8;
        v8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ah[i] = *reinterpret_cast<const v8*>(&Ash[wm * 64 + i * 16 + (lane & 15)][kc]);
            al[i] = *reinterpret_cast<const v8*>(&Asl[wm * 64 + i * 16 + (lane & 15)][kc]);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            bh[j] = *reinterpret_cast<const v8*>(&Bsh[wn * 64 + j * 16 + (lane & 15)][kc]);
            bl[j] = *reinterpret_cast<const v8*>(&Bsl[wn * 64 + j * 16 + (lane & 15)][kc]);
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[i][j] = mfma16x16x32<H>(al[i], bh[j], acc[i][j]);
                acc[i][j] = mfma16x16x32<H>(ah[i], bl[j], acc[i][j]);
                acc[i][j] = mfma16x16x32<H>(ah[i], bh[j], acc[i][j]);
            }
    }

    const int col_in = lane & 15, row_grp = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const long long m = m0 + wm * 64 + i * 16 + row_grp + r;
            float rmin = 3.0e38f;
            const bool mvalid = m < ep.M;
            int bidx = 0;
            double cb = 0.;
            float sa = 1.f;
            if (mvalid) {
                if (A.inv_scale) sa = A.inv_scale[m];
                if (EPI == EPI_DIFF2) {
                    bidx = (int)(m / ep.T);
                    cb = ep.Cb[bidx];
                }
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const long long n = n0 + wn * 64 + j * 16 + col_in;
                if (mvalid && n < ep.N) {
                    float v = acc[i][j][r] * sa;
                    if (B.inv_scale) v *= B.inv_scale[n];
                    if (EPI == EPI_STORE) {
                        ep.C[m * ep.ldc + n] = v;
                    } else {
                        float d = (float)(cb + (double)ep.N1[(long long)bidx * ep.ldn + n] - (double)v);
                        ep.C[m * ep.ldc + n] = d;
                        if (ep.prior_o[n] > 0.) rmin = fminf(rmin, d);
                    }
                }
            }
            if (EPI == EPI_DIFF2) {
                for (int o = 1; o < 16; o <<= 1) rmin = fminf(rmin, __shfl_xor(rmin, o));
                if (col_in == 0 && mvalid) ep.rowmin[m * ep.rowmin_ld + (n0 + wn * 64) / 64] = rmin;
            }
        }
    }
}

// hi/lo split of a row-major float matrix (rows x K), optionally with a power-of-two scale per
// row chosen so the row maximum maps to [0.5, 1); rows beyond `valid_rows` are written as zero.
template <class H>
__global__ void k_split_rows(const float* __restrict__ X, long long ldx, int K, int valid_rows,
                             H* __restrict__ hi, H* __restrict__ lo, long long ld, float* inv_scale,
                             int use_scale)
{
    const int r = blockIdx.x;
    const float* x = X + (size_t)r * ldx;
    float s = 1.f;
    if (use_scale) {
        float mx = 0.f;
        if (r < valid_rows)
            for (int k = threadIdx.x; k < K; k += blockDim.x) mx = fmaxf(mx, fabsf(x[k]));
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        __shared__ float red[4];
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
        __syncthreads();
        mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        int e = 0;
        if (mx > 0.f) (void)frexpf(mx, &e);
        s = ldexpf(1.f, -e);   // mx * s in [0.5, 1)
        if (threadIdx.x == 0 && inv_scale) inv_scale[r] = ldexpf(1.f, e);
    } else if (threadIdx.x == 0 && inv_scale) {
        inv_scale[r] = 1.f;
    }
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        float v = (r < valid_rows) ? x[k] * s : 0.f;
        H h = (H)v;
        H l = (H)(v - (float)h);
        hi[(size_t)r * ld + k] = h;
        lo[(size_t)r * ld + k] = l;
    }
}

int launch_split(int prec, const float* X, long long ldx, int K, int rows, int valid_rows, void* hi,
                 void* lo, long long ld, float* inv_scale, hipStream_t st)
{
    if (prec == T2X_PREC_BF16X3)
        k_split_rows<__bf16><<<rows, 256, 0, st>>>(X, ldx, K, valid_rows, (__bf16*)hi, (__bf16*)lo, ld, inv_scale, 0);
    else
        k_split_rows<_Float16><<<rows, 256, 0, st>>>(X, ldx, K, valid_rows, (_Float16*)hi, (_Float16*)lo, ld, inv_scale, 1);
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

int launch_gemm_x3(int prec, int epi, const void* Ah, const void* Al, const float* As, long long lda,
                   const void* Bh, const void* Bl, const float* Bs, long long ldb, int K, int Mrows_pad,
                   int Nrows_pad, const EpiParams& ep, hipStream_t st)
{
    if (K % XK != 0 || Mrows_pad % GBM != 0 || Nrows_pad % GBN != 0) {
        set_error("gemm_x3: operands not padded to tile multiples");
        return T2X_EINVAL;
    }
    const int tm = Mrows_pad / GBM, tn = Nrows_pad / GBN;
    const long long nwg = (long long)tm * tn;
    if (nwg > 0x7fffffffLL) { set_error("gemm: grid too large"); return T2X_EINVAL; }
    SplitOperand A{Ah, Al, As, lda}, B{Bh, Bl, Bs, ldb};
    if (prec == T2X_PREC_BF16X3) {
        if (epi == EPI_STORE) k_gemm_x3<__bf16, EPI_STORE><<<(unsigned)nwg, 256, 0, st>>>(A, B, K, tm, tn, ep);
        else k_gemm_x3<__bf16, EPI_DIFF2><<<(unsigned)nwg, 256, 0, st>>>(A, B, K, tm, tn, ep);
    } else {
        if (epi == EPI_STORE) k_gemm_x3<_Float16, EPI_STORE><<<(unsigned)nwg, 256, 0, st>>>(A, B, K, tm, tn, ep);
        else k_gemm_x3<_Float16, EPI_DIFF2><<<(unsigned)nwg, 256, 0, st>>>(A, B, K, tm, tn, ep);
    }
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

}  // namespace t2x
