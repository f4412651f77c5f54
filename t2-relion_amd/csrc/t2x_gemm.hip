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
