// t2x_gemm.hip — FP32 MFMA GEMM for the coarse-pass cross term.
//
// C[m][n] = sum_k A[m][k] * B[n][k] (both operands K-contiguous). The coarse diff2 of
// getAllSquaredDifferences (ml_optimiser.cpp:7265-7273) is recast as
//   diff2[b,t,n] = C_b + N1[b,n] - sum_k A[(b,t),k] B[n,k]
// with rows (particle, translation), columns (class, orientation) and k = (pixel, re/im).
// Tiles: 128x128 per workgroup, 4 waves of 64x64, K step 16, v_mfma_f32_16x16x4f32 (FP32) or
// K step 32, v_mfma_f32_16x16x32_{bf16,f16} on split hi/lo operands (split precision).
#include "t2x_kernels.h"
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

namespace t2x {

enum { EPI_STORE = 0, EPI_DIFF2 = 1 };


#define GBM 128
#define GBN 128

// Tile order. Workgroups are dealt round-robin to the 8 XCDs, and each XCD keeps ~64 of them
// resident. Tiles are walked in groups of `group_m` tile-rows with M fastest, so with
// group_m = 64 the 512 workgroups in flight cover 64 tile-rows x 8 tile-columns and each XCD
// gets 8 rows x 8 columns of them: every A and B tile slice fetched into an XCD's L2 feeds 8
// workgroups there, and all XCDs walk the same B columns at the same time (one HBM read of B
// per row group, shared through the MALL). The map is a bijection for any tiles_m, tiles_n.
// The epilogue's N1 term of particle p, column n: N1 itself, or in a bounds launch of the pruned
// coarse pass N1 - c0 sqrt(N1h) - c1 sqrt(N1) (the Cauchy-Schwarz bound of the cross term beyond
// the low-frequency prefix and the margin for the dropped FP16 product; c0, c1 per particle).
__device__ __forceinline__ float n1_at(const EpiParams& ep, long long p, long long n)
{
    const long long q = p * ep.ldn + n;
    const float v = ep.N1[q];
    if (ep.N1h == nullptr) return v;
    const double c0 = ep.prune_c[2 * p], c1 = ep.prune_c[2 * p + 1];
    return (float)((double)v - c0 * sqrt(fmax((double)ep.N1h[q], 0.)) - c1 * sqrt(fmax((double)v, 0.)));
}

__device__ __forceinline__ void tile_coords(int pid, int tiles_m, int tiles_n, int group_m, int& tm, int& tn)
{
    const int per_group = group_m * tiles_n;
    const int gid = pid / per_group;
    const int first_m = gid * group_m;
    const int gsize = min(tiles_m - first_m, group_m);
    const int in_group = pid - gid * per_group;
    tm = first_m + in_group % gsize;
    tn = in_group / gsize;
}

// Pruned coarse pass, tile plan: one wave per tile of tile_rows x tile_cols (tile_cols a multiple
// of 64). A tile stays uncomputed when every 64-column block of every valid row has its lower
// bound lb more than the negligible-weight cut (60, plus 1 for rounding) above the row's
// particle's seed (an upper bound of its minimum diff2); its blocks then get their bounds as
// minima (already in place when lb is the minima buffer itself). The other tiles are appended to
// `list` (count[0] of them), in no particular order.
__global__ __launch_bounds__(256) void k_prune_plan(const float* lb, const float* __restrict__ seed,
                                                    float* rowmin, int rowmin_ld, long long M, int T,
                                                    int tile_rows, int tile_cols, int tiles_m, int tiles_n,
                                                    unsigned* __restrict__ list, unsigned* __restrict__ count)
{
    const unsigned tile = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (tile >= (unsigned)tiles_m * (unsigned)tiles_n) return;   // uniform over the wave
    const int nblk = tile_cols / 64;
    const long long m0 = (long long)(tile / (unsigned)tiles_n) * tile_rows;
    const long long blk0 = (long long)(tile % (unsigned)tiles_n) * nblk;
    bool keep = false;
    for (int q = lane; q < tile_rows * nblk; q += 64) {
        const long long m = m0 + q / nblk, blk = blk0 + q % nblk;
        if (m >= M || blk >= rowmin_ld) continue;
        if (!(lb[m * rowmin_ld + blk] > seed[m / T] + 61.f)) keep = true;
    }
    if (__any(keep)) {
        if (lane == 0) list[atomicAdd(&count[0], 1u)] = tile;
        return;
    }
    if (rowmin == lb) return;
    for (int q = lane; q < tile_rows * nblk; q += 64) {
        const long long m = m0 + q / nblk, blk = blk0 + q % nblk;
        if (m >= M || blk >= rowmin_ld) continue;
        rowmin[m * rowmin_ld + blk] = lb[m * rowmin_ld + blk];
    }
}

// Shared epilogue of the GEMM kernels for one wave's block of MI x 4 MFMA tiles (lane holds
// C[4*(lane/16)+r][lane%16] of each 16x16 tile). Rows start at mbase, columns at ncol0 (a
// multiple of 64). EPI_STORE writes C = acc * row scale * column scale; EPI_DIFF2 writes
// diff2 = C_b + N1[b][n] - cross (b = row / T, rounded once from FP64) and the min over the
// 64 columns of each row restricted to scored orientations (prior > 0). Each 16-row block's
// inputs are gathered before use so their loads overlap. With a row map, GEMM row m (< ep.M)
// is output row rowmap[m].
template <int MI, int EPI>
__device__ __forceinline__ void gemm_epilogue(const floatx4 (&acc)[MI][4], long long mbase, long long ncol0,
                                              int lane, const float* __restrict__ a_scale,
                                              const float* __restrict__ b_scale, const EpiParams& ep,
                                              const int* __restrict__ rowmap = nullptr)
{
    const int col_in = lane & 15, row_grp = (lane >> 4) * 4;
    float sb[4];
    bool nval[4], scored[4];
    long long ncol[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        ncol[j] = ncol0 + j * 16 + col_in;
        nval[j] = ncol[j] < ep.N;
        sb[j] = (b_scale && nval[j]) ? b_scale[ncol[j]] : 1.f;
        scored[j] = (EPI == EPI_DIFF2) && nval[j] && ep.prior_o[ncol[j]] > 0.;
    }
    const float invT = (EPI == EPI_DIFF2) ? 1.f / (float)ep.T : 0.f;
#pragma unroll
    for (int i = 0; i < MI; i++) {
        int mrow[4], pb[4];
        bool mval[4];
        float sa[4], n1[4][4];
        double cb[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const long long m = mbase + i * 16 + row_grp + r;
            mval[r] = m < ep.M;
            const int mm = mval[r] ? (rowmap ? rowmap[m] : (int)m) : 0;   // output row
            mrow[r] = mm;
            sa[r] = a_scale ? a_scale[mm] : 1.f;
            if (EPI == EPI_DIFF2) {
                int b = (int)((float)mm * invT);   // particle index m / T: float estimate, then exact
                while (b > 0 && b * ep.T > mm) b--;
                while ((b + 1) * ep.T <= mm) b++;
                pb[r] = b;
                cb[r] = ep.Cb[b];
#pragma unroll
                for (int j = 0; j < 4; j++) n1[r][j] = n1_at(ep, b, nval[j] ? ncol[j] : 0);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float rmin = 3.0e38f;
            float* crow = ep.C + (long long)mrow[r] * ep.ldc;
            if (EPI == EPI_STORE) {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (mval[r] && nval[j]) crow[ncol[j]] = acc[i][j][r] * sa[r] * sb[j];
            } else {
                float dv[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float v = acc[i][j][r] * sa[r] * sb[j];
                    dv[j] = (float)(cb[r] + (double)n1[r][j] - (double)v);
                    if (mval[r] && nval[j] && scored[j]) rmin = fminf(rmin, dv[j]);
                }
                for (int o = 1; o < 16; o <<= 1) rmin = fminf(rmin, __shfl_xor(rmin, o));
                bool store = ep.C != nullptr;   // null: a bounds-only launch (block minima alone)
                if (ep.pmin && mval[r]) {
                    // the particle's final minimum is <= est, so a block more than the cut (60,
                    // plus one for rounding) above est is negligible whatever the final minimum
                    const float est = ordered_to_float(ep.pmin[pb[r]]);
                    store = store && !(rmin > est + 61.f);
                    if (rmin < est && col_in == 0) atomicMin(&ep.pmin[pb[r]], float_to_ordered(rmin));
                }
                if (store) {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (mval[r] && nval[j]) crow[ncol[j]] = dv[j];
                }
                if (col_in == 0 && mval[r]) ep.rowmin[(long long)mrow[r] * ep.rowmin_ld + ncol0 / 64] = rmin;
            }
        }
    }
}
#define GBK 16
#define GPAD 4

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_tn(const float* __restrict__ A, long long lda,
                                                 const float* __restrict__ B, long long ldb,
                                                 int K, int tiles_m, int tiles_n, int group_m, EpiParams ep)
{
    __shared__ float As[GBK][GBM + GPAD];
    __shared__ float Bs[GBK][GBN + GPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, group_m, tm, tn);
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
    gemm_epilogue<4, EPI>(acc, m0 + wm * 64, n0 + wn * 64, lane, nullptr, nullptr, ep);
}

template __global__ void k_gemm_tn<EPI_STORE>(const float*, long long, const float*, long long, int, int, int, int, EpiParams);
template __global__ void k_gemm_tn<EPI_DIFF2>(const float*, long long, const float*, long long, int, int, int, int, EpiParams);

// group_m of the tile walk; T2X_GEMM_GROUP_M overrides it for tile-order sweeps
static int gemm_group_m()
{
    static int g = [] {
        const char* e = getenv("T2X_GEMM_GROUP_M");
        int v = e ? atoi(e) : 0;
        return v > 0 ? v : 64;
    }();
    return g;
}

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
        k_gemm_tn<EPI_STORE><<<(unsigned)nwg, 256, 0, st>>>(A, lda, B, ldb, K, tm, tn, gemm_group_m(), ep);
    else
        k_gemm_tn<EPI_DIFF2><<<(unsigned)nwg, 256, 0, st>>>(A, lda, B, ldb, K, tm, tn, gemm_group_m(), ep);
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
                                                 int tiles_n, int group_m, EpiParams ep)
{
    typedef typename Vec8<H>::type v8;
    __shared__ __attribute__((aligned(16))) H Ash[GBM][XK + XPAD];
    __shared__ __attribute__((aligned(16))) H Asl[GBM][XK + XPAD];
    __shared__ __attribute__((aligned(16))) H Bsh[GBN][XK + XPAD];
    __shared__ __attribute__((aligned(16))) H Bsl[GBN][XK + XPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, group_m, tm, tn);
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
        for (int i = 0; i < 4; i++) {
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(al[i], bh[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah[i], bl[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah[i], bh[j], acc[i][j]);
        }
    }

    gemm_epilogue<4, EPI>(acc, m0 + wm * 64, n0 + wn * 64, lane, A.inv_scale, B.inv_scale, ep);
}

// Wide variant: 256x256 tile per workgroup of 8 waves, each wave 128x64 (8x4 MFMA tiles), so
// each wave issues 24 LDS fragment reads per 96 MFMAs (the 128x128 kernel: 16 per 48) and each
// operand byte fetched from L2 feeds twice the flops. LDS rows are 32 halfs (64 B, 4 chunks of
// 16 B) with chunk c of row r stored at c ^ ((r >> 1) & 3): conflict-free 16-byte reads for the
// MFMA fragment pattern (16 rows x 4 chunks per wave) and for the row-major tile stores.
#define WBM 256
#define WBN 256

__device__ __forceinline__ int swz(int r, int c) { return r * 32 + ((c ^ ((r >> 1) & 3)) << 3); }

template <class H, int EPI>
__global__ __launch_bounds__(512) void k_gemm_x3w(SplitOperand A, SplitOperand B, int K, int tiles_m,
                                                  int tiles_n, int group_m, EpiParams ep,
                                                  const int* __restrict__ rowmap)
{
    typedef typename Vec8<H>::type v8;
    // two LDS stages (2 x 64 KiB): the tile slices for step k+1 are stored while step k's MFMAs
    // run, so each K step needs a single barrier
    __shared__ __attribute__((aligned(16))) H lds[2][4][WBM * 32];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;   // 2 waves along M (128 rows), 4 along N (64 cols)
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, group_m, tm, tn);
    const long long m0 = (long long)tm * WBM, n0 = (long long)tn * WBN;

    floatx4 acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // loader: rows lrow and lrow + 128, 16-byte chunk lc of the 64-byte K slice
    const int lrow = tid >> 2, lc = tid & 3;
    const H* Ah = reinterpret_cast<const H*>(A.hi);
    const H* Al = reinterpret_cast<const H*>(A.lo);
    const H* Bh = reinterpret_cast<const H*>(B.hi);
    const H* Bl = reinterpret_cast<const H*>(B.lo);
    // A rows: through the row map when given (rows beyond ep.M read row 0, their output is dropped)
    const long long am0 = rowmap ? ((m0 + lrow < ep.M) ? rowmap[m0 + lrow] : 0) : m0 + lrow;
    const long long am1 = rowmap ? ((m0 + lrow + 128 < ep.M) ? rowmap[m0 + lrow + 128] : 0) : m0 + lrow + 128;
    const size_t ra0 = (size_t)am0 * A.ld + lc * 8, ra1 = (size_t)am1 * A.ld + lc * 8;
    const size_t rb0 = (size_t)(n0 + lrow) * B.ld + lc * 8, rb1 = (size_t)(n0 + lrow + 128) * B.ld + lc * 8;
    const int s0 = swz(lrow, lc), s1 = swz(lrow + 128, lc);
    uint4 p0, p1, p2, p3, p4, p5, p6, p7;
#define X3W_GLOAD(k)                                               \
    do {                                                           \
        p0 = *reinterpret_cast<const uint4*>(Ah + ra0 + (k));      \
        p1 = *reinterpret_cast<const uint4*>(Ah + ra1 + (k));      \
        p2 = *reinterpret_cast<const uint4*>(Al + ra0 + (k));      \
        p3 = *reinterpret_cast<const uint4*>(Al + ra1 + (k));      \
        p4 = *reinterpret_cast<const uint4*>(Bh + rb0 + (k));      \
        p5 = *reinterpret_cast<const uint4*>(Bh + rb1 + (k));      \
        p6 = *reinterpret_cast<const uint4*>(Bl + rb0 + (k));      \
        p7 = *reinterpret_cast<const uint4*>(Bl + rb1 + (k));      \
    } while (0)
#define X3W_LSTORE(b)                                              \
    do {                                                           \
        *reinterpret_cast<uint4*>(&lds[b][0][s0]) = p0;            \
        *reinterpret_cast<uint4*>(&lds[b][0][s1]) = p1;            \
        *reinterpret_cast<uint4*>(&lds[b][1][s0]) = p2;            \
        *reinterpret_cast<uint4*>(&lds[b][1][s1]) = p3;            \
        *reinterpret_cast<uint4*>(&lds[b][2][s0]) = p4;            \
        *reinterpret_cast<uint4*>(&lds[b][2][s1]) = p5;            \
        *reinterpret_cast<uint4*>(&lds[b][3][s0]) = p6;            \
        *reinterpret_cast<uint4*>(&lds[b][3][s1]) = p7;            \
    } while (0)
    X3W_GLOAD(0);
    X3W_LSTORE(0);
    __syncthreads();

    const int fr = lane & 15, fc = lane >> 4;   // fragment row within 16, 16-byte chunk 0..3
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += XK) {
        const bool more = k0 + XK < K;
        if (more) X3W_GLOAD(k0 + XK);
        const H* Ash = lds[buf][0];
        const H* Asl = lds[buf][1];
        const H* Bsh = lds[buf][2];
        const H* Bsl = lds[buf][3];
        v8 bh[4], bl[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = swz(wn * 64 + j * 16 + fr, fc);
            bh[j] = *reinterpret_cast<const v8*>(&Bsh[o]);
            bl[j] = *reinterpret_cast<const v8*>(&Bsl[o]);
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int o = swz(wm * 128 + i * 16 + fr, fc);
            const v8 ah = *reinterpret_cast<const v8*>(&Ash[o]);
            const v8 al = *reinterpret_cast<const v8*>(&Asl[o]);
            // consecutive MFMAs go to 4 different accumulators; each element still sums
            // lo.hi, hi.lo, hi.hi in that order
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(al, bh[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah, bl[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah, bh[j], acc[i][j]);
        }
        if (more) X3W_LSTORE(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

#undef X3W_GLOAD
#undef X3W_LSTORE
    gemm_epilogue<8, EPI>(acc, m0 + wm * 128, n0 + wn * 64, lane, A.inv_scale, B.inv_scale, ep, rowmap);
}

// Direct-load variant: 256x128 tile per workgroup of 4 waves (each 128x64), operand slices
// moved from global memory straight into LDS with global_load_lds_dwordx4 (no register
// staging), single LDS stage of 48 KiB, so two workgroups share a CU and each one's load and
// barrier phases overlap the other's MFMAs. One wave instruction fills 1 KiB of LDS (16 rows x
// 4 chunks); lane q writes physical chunk q & 3 of row q >> 2 and therefore fetches logical
// chunk (q & 3) ^ ((row >> 1) & 3), which yields the same swizzled layout as swz().
#define DBM 256
#define DBN 128

typedef const void __attribute__((address_space(1)))* gptr_t;
typedef void __attribute__((address_space(3)))* lptr_t;

template <class H>
__device__ __forceinline__ void load_rows_lds(const H* __restrict__ g, long long ld, long long row0, int k0,
                                              H* lds_rows, int lane)
{
    // 16 rows starting at row0 (global) -> lds_rows (16 x 32 halfs, swizzled)
    const int r = lane >> 2, pc = lane & 3;
    const int c = pc ^ ((r >> 1) & 3);
    const H* src = g + (size_t)(row0 + r) * ld + k0 + c * 8;
    __builtin_amdgcn_global_load_lds((gptr_t)src, (lptr_t)lds_rows, 16, 0, 0);
}

template <class H, int EPI>
__global__ __launch_bounds__(256, 2) void k_gemm_x3d(SplitOperand A, SplitOperand B, int K, int tiles_m,
                                                     int tiles_n, int group_m, EpiParams ep)
{
    typedef typename Vec8<H>::type v8;
    __shared__ __attribute__((aligned(16))) H Ash[DBM * 32];
    __shared__ __attribute__((aligned(16))) H Asl[DBM * 32];
    __shared__ __attribute__((aligned(16))) H Bsh[DBN * 32];
    __shared__ __attribute__((aligned(16))) H Bsl[DBN * 32];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1;   // 2 waves along M (128 rows), 2 along N (64 cols)
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, group_m, tm, tn);
    const long long m0 = (long long)tm * DBM, n0 = (long long)tn * DBN;
    const H* Ah = reinterpret_cast<const H*>(A.hi);
    const H* Al = reinterpret_cast<const H*>(A.lo);
    const H* Bh = reinterpret_cast<const H*>(B.hi);
    const H* Bl = reinterpret_cast<const H*>(B.lo);

    floatx4 acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int fr = lane & 15, fc = lane >> 4;
    for (int k0 = 0; k0 < K; k0 += XK) {
        // each wave moves 4 A row blocks (hi and lo) and 2 B row blocks (hi and lo) of 16 rows
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int rb = (wave * 4 + q) * 16;
            load_rows_lds<H>(Ah, A.ld, m0 + rb, k0, &Ash[rb * 32], lane);
            load_rows_lds<H>(Al, A.ld, m0 + rb, k0, &Asl[rb * 32], lane);
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int rb = (wave * 2 + q) * 16;
            load_rows_lds<H>(Bh, B.ld, n0 + rb, k0, &Bsh[rb * 32], lane);
            load_rows_lds<H>(Bl, B.ld, n0 + rb, k0, &Bsl[rb * 32], lane);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        v8 bh[4], bl[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = swz(wn * 64 + j * 16 + fr, fc);
            bh[j] = *reinterpret_cast<const v8*>(&Bsh[o]);
            bl[j] = *reinterpret_cast<const v8*>(&Bsl[o]);
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int o = swz(wm * 128 + i * 16 + fr, fc);
            const v8 ah = *reinterpret_cast<const v8*>(&Ash[o]);
            const v8 al = *reinterpret_cast<const v8*>(&Asl[o]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(al, bh[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah, bl[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah, bh[j], acc[i][j]);
        }
        __syncthreads();
    }

    gemm_epilogue<8, EPI>(acc, m0 + wm * 128, n0 + wn * 64, lane, A.inv_scale, B.inv_scale, ep);
}

// Pipelined wide variant: 256x256 tile, 8 waves of 128x64 as k_gemm_x3w, but the operand slices
// go from global memory straight into LDS (global_load_lds_dwordx4, no register staging), one
// K step ahead of use into the stage that was just read, and the single barrier of a K step
// sits before its last two MFMA groups: the fragment reads of the next step (its B fragments and
// first A fragment) are issued right after the barrier and overlap those MFMAs, instead of every
// wave waiting on LDS after the barrier. The two stages are separate LDS arrays and the K loop is
// unrolled by two, so every fragment read names a stage the compiler can tell apart from the
// stage being filled (it must not wait for those loads before reading the other stage). Each
// element sees the MFMA chain of k_gemm_x3w (lo.hi, hi.lo, hi.hi per K step, K steps in order),
// so both give identical results.
template <class H>
__device__ __forceinline__ void lds_dma16(const H* src, H* lds_dst)
{
    __builtin_amdgcn_global_load_lds((gptr_t)src, (lptr_t)lds_dst, 16, 0, 0);
}

template <class H>
__device__ __forceinline__ void x3_group(floatx4 (&acc)[4], typename Vec8<H>::type al, typename Vec8<H>::type ah,
                                         const typename Vec8<H>::type (&bh)[4], const typename Vec8<H>::type (&bl)[4],
                                         bool with_bl = true, bool with_al = true)
{
    if (with_al) {
#pragma unroll
        for (int j = 0; j < 4; j++) acc[j] = mfma16x16x32<H>(al, bh[j], acc[j]);
    }
    if (with_bl) {
#pragma unroll
        for (int j = 0; j < 4; j++) acc[j] = mfma16x16x32<H>(ah, bl[j], acc[j]);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) acc[j] = mfma16x16x32<H>(ah, bh[j], acc[j]);
}

// NP = 4 (one product over two K chunks per stage): the hi halves of chunk 0 (ah, bh) and of
// chunk 1 (al, bl, held in the lo slots).
template <class H>
__device__ __forceinline__ void x3_group2(floatx4 (&acc)[4], typename Vec8<H>::type a1, typename Vec8<H>::type a0,
                                          const typename Vec8<H>::type (&b0)[4], const typename Vec8<H>::type (&b1)[4])
{
#pragma unroll
    for (int j = 0; j < 4; j++) acc[j] = mfma16x16x32<H>(a1, b1[j], acc[j]);
#pragma unroll
    for (int j = 0; j < 4; j++) acc[j] = mfma16x16x32<H>(a0, b0[j], acc[j]);
}

// Epilogue inputs of one 256x256 tile, staged in LDS before the K loop so that their global
// latency hides behind it (read per 16-row block after the loop, they left every wave waiting
// on memory eight times per tile): per tile row its output row (-1: padding), particle, slot
// (the tile's particles in order) and row scale; per tile column its scale and whether it is
// scored; per slot the particle and C_b + N1 (in FP64, as the epilogue adds them first) over
// the tile's columns. A tile spans at most ceil(255 / T) + 1 particles, so EPI_SLOTS = 8 covers
// T >= 37; for fewer translations the kernel takes gemm_epilogue. Row indices stay below 2^31.
#define EPI_SLOTS 8
struct EpiStage {
    int4 row[256];
    float sb[256];
    int scored[256];
    int pslot[EPI_SLOTS];
    double cbn1[EPI_SLOTS][256];
};

// Fills es (all 512 threads; published by the caller's next barrier). Returns false (uniformly)
// when the tile spans more than EPI_SLOTS particles.
template <int EPI>
__device__ __forceinline__ bool stage_epilogue(EpiStage& es, long long m0l, long long n0, const float* a_scale,
                                               const float* b_scale, const EpiParams& ep, const int* rowmap, int tid)
{
    const unsigned T = (EPI == EPI_DIFF2) ? (unsigned)ep.T : 1u;
    const unsigned m0 = (unsigned)m0l;
    const unsigned p0 = m0 / T;   // slot 0 (GEMM rows of slot s: T (p0 + s) .. T (p0 + s + 1) - 1)
    const int mlast = min((int)m0 + 255, ep.M - 1);
    const int nslots = (EPI == EPI_DIFF2 && mlast >= (int)m0) ? (int)((unsigned)mlast / T - p0) + 1 : 0;
    if (nslots > EPI_SLOTS) return false;
    if (tid < 256) {
        const unsigned m = m0 + tid;
        int4 e = make_int4(-1, 0, 0, __float_as_int(1.f));
        if ((int)m < ep.M) {
            const unsigned mm = rowmap ? (unsigned)rowmap[m] : m;
            e.x = (int)mm;
            e.y = (int)(mm / T);
            e.z = (int)(m / T - p0);
            if (a_scale) e.w = __float_as_int(a_scale[mm]);
        }
        es.row[tid] = e;
    } else {
        const int c = tid - 256;
        const long long n = n0 + c;
        const bool nv = n < ep.N;
        es.sb[c] = (b_scale && nv) ? b_scale[n] : 1.f;
        es.scored[c] = (EPI == EPI_DIFF2) && nv && ep.prior_o[n] > 0.;
    }
    if (EPI == EPI_DIFF2) {
        for (int q = tid; q < nslots * 256; q += 512) {
            const int s = q >> 8, c = q & 255;
            const unsigned ms = max(m0, (p0 + s) * T);   // a valid row of slot s
            const unsigned p = rowmap ? (unsigned)rowmap[ms] / T : p0 + s;
            const long long n = n0 + c;
            es.cbn1[s][c] = ep.Cb[p] + (double)(n < ep.N ? n1_at(ep, p, n) : 0.f);
        }
        if (tid < EPI_SLOTS) {
            int p = -1;   // no particle: slot beyond the tile's valid rows
            if (tid < nslots) {
                const unsigned ms = max(m0, (p0 + tid) * T);
                p = rowmap ? (int)((unsigned)rowmap[ms] / T) : (int)(p0 + tid);
            }
            es.pslot[tid] = p;
        }
    }
    return true;
}

// Minimum over the 16 lanes of each DPP row (lanes 16r .. 16r+15): quad_perm xor 1 and xor 2,
// then the half-row mirror and the row mirror.
__device__ __forceinline__ float min_row16(float v)
{
    v = fminf(v, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, false)));
    return v;
}

// gemm_epilogue for one wave's 128x64 block of a staged tile (tile rows trow0.., columns
// tcol0..); same arithmetic, same outputs. The running minimum of each of the wave's particles
// is read once, and lowered by the wave's own block minima as it goes (any value taken by the
// running minimum bounds the final minimum from above, which is all the store test needs).
// C_b + N1 of slot z at tile column c, in FP64 (the order gemm_epilogue adds them in)
__device__ __forceinline__ double es_base(const EpiStage& es, int z, int c) { return es.cbn1[z][c]; }

template <int EPI, int MI, class ES>
__device__ __forceinline__ void gemm_epilogue_staged(const floatx4 (&acc)[MI][4], const ES& es, int trow0,
                                                     int tcol0, long long n0, int lane, const EpiParams& ep)
{
    const int col_in = lane & 15, row_grp = (lane >> 4) * 4;
    float sb[4];
    bool nval[4], scored[4];
    long long ncol[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int c = tcol0 + j * 16 + col_in;
        ncol[j] = n0 + c;
        nval[j] = ncol[j] < ep.N;
        sb[j] = es.sb[c];
        scored[j] = es.scored[c] != 0;
    }
    // running minima of the first 4 particles of the wave's rows (slots z0 .. z0 + 3; with
    // T >= 43 that is all of them, rows of further slots are stored without the test)
    const int z0 = max(0, es.row[trow0].z);
    float est[4] = {3.4e38f, 3.4e38f, 3.4e38f, 3.4e38f};
    const bool lazy = (EPI == EPI_DIFF2) && ep.pmin != nullptr;
    if (lazy) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int p = z0 + q < EPI_SLOTS ? es.pslot[z0 + q] : -1;
            if (p >= 0) est[q] = ordered_to_float(ep.pmin[p]);
        }
    }
    const long long blk = (n0 + tcol0) / 64;
    // branch-light: padding rows (output row -1, uniform over the 16 lanes of a row) are
    // masked, the running minima are selected and updated with conditional moves
#pragma unroll
    for (int i = 0; i < MI; i++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int4 e = es.row[trow0 + i * 16 + row_grp + r];
            const bool rv = e.x >= 0;
            const float sa = __int_as_float(e.w);
            float* crow = ep.C + (long long)max(e.x, 0) * ep.ldc;
            if (EPI == EPI_STORE) {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (rv && nval[j]) crow[ncol[j]] = acc[i][j][r] * sa * sb[j];
            } else {
                float rmin = 3.0e38f, dv[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float v = acc[i][j][r] * sa * sb[j];
                    dv[j] = (float)(es_base(es, e.z, tcol0 + j * 16 + col_in) - (double)v);
                    rmin = (nval[j] && scored[j]) ? fminf(rmin, dv[j]) : rmin;
                }
                rmin = min_row16(rmin);
                bool store = rv && ep.C != nullptr;   // null: a bounds-only launch (block minima alone)
                if (lazy) {
                    // the particle's final minimum is <= es_q, so a block more than the cut (60,
                    // plus one for rounding) above es_q is negligible whatever the final minimum
                    const int q = e.z - z0;
                    float es_q = 3.4e38f;
                    es_q = q == 0 ? est[0] : es_q;
                    es_q = q == 1 ? est[1] : es_q;
                    es_q = q == 2 ? est[2] : es_q;
                    es_q = q == 3 ? est[3] : es_q;
                    store = store && !(rmin > es_q + 61.f);
                    const bool lower = rv && rmin < es_q;
                    if (lower && col_in == 0) atomicMin(&ep.pmin[e.y], float_to_ordered(rmin));
                    est[0] = (lower && q == 0) ? rmin : est[0];
                    est[1] = (lower && q == 1) ? rmin : est[1];
                    est[2] = (lower && q == 2) ? rmin : est[2];
                    est[3] = (lower && q == 3) ? rmin : est[3];
                }
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (store && nval[j]) crow[ncol[j]] = dv[j];
                if (rv && col_in == 0) ep.rowmin[(long long)e.x * ep.rowmin_ld + blk] = rmin;
            }
        }
    }
}

// NP = 3: Al Bh + Ah Bl + Ah Bh; NP = 2 (FP16X2): Al Bh + Ah Bh, the B operand's hi half alone
// (its lo half neither loaded nor read); NP = 1 (a bounds form of the pruned coarse pass):
// Ah Bh alone; NP = 4 (the other bounds form): Ah Bh alone over two K chunks per stage, the lo
// slots holding the hi halves of the second chunk (K a multiple of 64).
// One tile, rows [m0, m0 + WBM) x columns [n0, n0 + WBN): stage s of the K loop lives in st0
// (s even) or st1 (s odd); [operand][row * 32 + chunk] with operands Ah, Al, Bh, Bl and 16-byte
// chunks swizzled as swz().
template <class H, int EPI, int NP>
__device__ __forceinline__ void x3p_tile(H (&st0)[4][WBM * 32], H (&st1)[4][WBM * 32], EpiStage& es,
                                         const SplitOperand& A, const SplitOperand& B, int K, long long m0,
                                         long long n0, const EpiParams& ep, const int* __restrict__ rowmap)
{
    typedef typename Vec8<H>::type v8;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    constexpr int KS = NP == 4 ? 2 * XK : XK;   // K columns per stage

    floatx4 acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // loader: wave w moves row blocks [32w, 32w+16) and [32w+16, 32w+32) of each operand; one
    // instruction fills 16 rows x 64 bytes, lane q writing physical chunk q & 3 of row q >> 2,
    // so it fetches logical chunk (q & 3) ^ ((row >> 1) & 3)
    const int lr = lane >> 2;
    const int lch = (lane & 3) ^ ((lr >> 1) & 3);
    const int rb0 = wave * 32, rb1 = wave * 32 + 16;
    const long long ga0 = m0 + rb0 + lr, ga1 = m0 + rb1 + lr;
    // A rows through the row map when given (rows beyond ep.M read row 0, their output is dropped)
    const long long am0 = rowmap ? (ga0 < ep.M ? rowmap[ga0] : 0) : ga0;
    const long long am1 = rowmap ? (ga1 < ep.M ? rowmap[ga1] : 0) : ga1;
    // per-lane byte offsets below 4 GiB (checked by the launcher), bases wave-uniform, so each
    // load is a scalar base plus one 32-bit lane offset
    const char* Ah = reinterpret_cast<const char*>(A.hi);
    const char* Al = reinterpret_cast<const char*>(A.lo);
    const char* Bh = reinterpret_cast<const char*>(B.hi);
    const char* Bl = reinterpret_cast<const char*>(B.lo);
    const unsigned oa0 = (unsigned)((am0 * A.ld + lch * 8) * (long long)sizeof(H));
    const unsigned oa1 = (unsigned)((am1 * A.ld + lch * 8) * (long long)sizeof(H));
    const unsigned ob0 = (unsigned)(((n0 + rb0 + lr) * B.ld + lch * 8) * (long long)sizeof(H));
    const unsigned ob1 = (unsigned)(((n0 + rb1 + lr) * B.ld + lch * 8) * (long long)sizeof(H));
#define X3P_LOAD(ST, k)                                                                           \
    do {                                                                                          \
        const long long kb = (long long)(k) * (long long)sizeof(H);                               \
        lds_dma16<H>(reinterpret_cast<const H*>(Ah + kb + oa0), &ST[0][rb0 * 32]);               \
        lds_dma16<H>(reinterpret_cast<const H*>(Ah + kb + oa1), &ST[0][rb1 * 32]);               \
        if (NP == 2 || NP == 3) {                                                                 \
            lds_dma16<H>(reinterpret_cast<const H*>(Al + kb + oa0), &ST[1][rb0 * 32]);           \
            lds_dma16<H>(reinterpret_cast<const H*>(Al + kb + oa1), &ST[1][rb1 * 32]);           \
        }                                                                                         \
        lds_dma16<H>(reinterpret_cast<const H*>(Bh + kb + ob0), &ST[2][rb0 * 32]);               \
        lds_dma16<H>(reinterpret_cast<const H*>(Bh + kb + ob1), &ST[2][rb1 * 32]);               \
        if (NP == 3) {                                                                            \
            lds_dma16<H>(reinterpret_cast<const H*>(Bl + kb + ob0), &ST[3][rb0 * 32]);           \
            lds_dma16<H>(reinterpret_cast<const H*>(Bl + kb + ob1), &ST[3][rb1 * 32]);           \
        }                                                                                         \
        if (NP == 4) {   /* the hi halves' next K chunk into the lo slots */                      \
            const long long kc = kb + XK * (long long)sizeof(H);                                  \
            lds_dma16<H>(reinterpret_cast<const H*>(Ah + kc + oa0), &ST[1][rb0 * 32]);           \
            lds_dma16<H>(reinterpret_cast<const H*>(Ah + kc + oa1), &ST[1][rb1 * 32]);           \
            lds_dma16<H>(reinterpret_cast<const H*>(Bh + kc + ob0), &ST[3][rb0 * 32]);           \
            lds_dma16<H>(reinterpret_cast<const H*>(Bh + kc + ob1), &ST[3][rb1 * 32]);           \
        }                                                                                         \
    } while (0)

    // fragment offsets: A block i of this wave at oa + i * 512, B block j at ob + j * 512 (the
    // swizzle depends on the row only through (row >> 1) & 3, the same for every 16-row block)
    const int fr = lane & 15, fc = lane >> 4;
    const int oa = swz(wm * 128 + fr, fc), ob = swz(wn * 64 + fr, fc);
#define X3P_READ_B(ST, BH, BL)                                                   \
    do {                                                                         \
        _Pragma("unroll") for (int j = 0; j < 4; j++) {                          \
            BH[j] = *reinterpret_cast<const v8*>(&ST[2][ob + j * 512]);          \
            if (NP >= 3) BL[j] = *reinterpret_cast<const v8*>(&ST[3][ob + j * 512]); \
        }                                                                        \
    } while (0)
#define X3P_READ_A(ST, i, AH, AL)                                                \
    do {                                                                         \
        AH = *reinterpret_cast<const v8*>(&ST[0][oa + (i) * 512]);               \
        if (NP >= 2) AL = *reinterpret_cast<const v8*>(&ST[1][oa + (i) * 512]);  \
        else AL = v8{};                                                          \
    } while (0)

    // One K step on stage CUR (fragments bh/bl and A block 0 already in registers): A blocks
    // 1..7 are read one group ahead (6 and 7 together), then the barrier; after it the next
    // stage NXT is complete and CUR free, so stage s+2 starts loading into CUR and the first
    // fragments of NXT are read while groups 6 and 7 run.
#define X3P_STEP(CUR, NXT, s)                                                    \
    do {                                                                         \
        v8 a6h, a6l, a7h, a7l;                                                   \
        _Pragma("unroll") for (int i = 0; i < 6; i++) {                          \
            v8 nh, nl;                                                           \
            if (i < 5) {                                                         \
                X3P_READ_A(CUR, i + 1, nh, nl);                                  \
            } else {                                                             \
                X3P_READ_A(CUR, 6, a6h, a6l);                                    \
                X3P_READ_A(CUR, 7, a7h, a7l);                                    \
            }                                                                    \
            if (NP == 4) x3_group2<H>(acc[i], al, ah, bh, bl);                    \
            else x3_group<H>(acc[i], al, ah, bh, bl, NP == 3, NP >= 2);          \
            if (i < 5) { ah = nh; al = nl; }                                     \
        }                                                                        \
        __builtin_amdgcn_s_waitcnt(0);                                           \
        __syncthreads();                                                         \
        if ((s) + 2 < nk) X3P_LOAD(CUR, ((s) + 2) * KS);                         \
        const bool nxt = (s) + 1 < nk;                                           \
        if (nxt) X3P_READ_A(NXT, 0, ah, al);                                     \
        /* groups 6 and 7 column by column: B fragment j of this step is dead */ \
        /* after column j, and fragment j of the next step takes its registers */ \
        _Pragma("unroll") for (int j = 0; j < 4; j++) {                          \
            if (NP == 2 || NP == 3) acc[6][j] = mfma16x16x32<H>(a6l, bh[j], acc[6][j]); \
            if (NP == 3) acc[6][j] = mfma16x16x32<H>(a6h, bl[j], acc[6][j]);     \
            if (NP == 4) acc[6][j] = mfma16x16x32<H>(a6l, bl[j], acc[6][j]);     \
            acc[6][j] = mfma16x16x32<H>(a6h, bh[j], acc[6][j]);                  \
            if (NP == 2 || NP == 3) acc[7][j] = mfma16x16x32<H>(a7l, bh[j], acc[7][j]); \
            if (NP == 3) acc[7][j] = mfma16x16x32<H>(a7h, bl[j], acc[7][j]);     \
            if (NP == 4) acc[7][j] = mfma16x16x32<H>(a7l, bl[j], acc[7][j]);     \
            acc[7][j] = mfma16x16x32<H>(a7h, bh[j], acc[7][j]);                  \
            if (nxt) {                                                           \
                bh[j] = *reinterpret_cast<const v8*>(&NXT[2][ob + j * 512]);     \
                if (NP >= 3) bl[j] = *reinterpret_cast<const v8*>(&NXT[3][ob + j * 512]); \
            }                                                                    \
        }                                                                        \
    } while (0)

    const int nk = K / KS;
    X3P_LOAD(st0, 0);
    const bool staged = stage_epilogue<EPI>(es, m0, n0, A.inv_scale, B.inv_scale, ep, rowmap, tid);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (nk > 1) X3P_LOAD(st1, KS);
    v8 bh[4], bl[4], ah, al;
    X3P_READ_B(st0, bh, bl);
    X3P_READ_A(st0, 0, ah, al);
    for (int s = 0; s < nk; s += 2) {
        X3P_STEP(st0, st1, s);
        if (s + 1 < nk) X3P_STEP(st1, st0, s + 1);
    }
#undef X3P_STEP
#undef X3P_READ_A
#undef X3P_READ_B
#undef X3P_LOAD
    if (staged)
        gemm_epilogue_staged<EPI>(acc, es, wm * 128, wn * 64, n0, lane, ep);
    else
        gemm_epilogue<8, EPI>(acc, m0 + wm * 128, n0 + wn * 64, lane, A.inv_scale, B.inv_scale, ep, rowmap);
}

// LIST (the pruned coarse pass): a persistent grid over the tiles listed in ep.tile_list
// (*ep.tile_count of them, indices tm * tiles_n + tn), no row map.
template <class H, int EPI, int NP = 3, bool LIST = false>
__global__ __launch_bounds__(512) void k_gemm_x3p(SplitOperand A, SplitOperand B, int K, int tiles_m,
                                                  int tiles_n, int group_m, EpiParams ep,
                                                  const int* __restrict__ rowmap)
{
    __shared__ __attribute__((aligned(16))) H st0[4][WBM * 32];
    __shared__ __attribute__((aligned(16))) H st1[4][WBM * 32];
    __shared__ EpiStage es;
    if constexpr (LIST) {
        const unsigned n = *ep.tile_count;
        for (unsigned it = blockIdx.x; it < n; it += gridDim.x) {
            const unsigned t = ep.tile_list[it];
            x3p_tile<H, EPI, NP>(st0, st1, es, A, B, K, (long long)(t / (unsigned)tiles_n) * WBM,
                                 (long long)(t % (unsigned)tiles_n) * WBN, ep, nullptr);
            __syncthreads();   // the next tile reuses the LDS stages
        }
    } else {
        int tm, tn;
        tile_coords(blockIdx.x, tiles_m, tiles_n, group_m, tm, tn);
        x3p_tile<H, EPI, NP>(st0, st1, es, A, B, K, (long long)tm * WBM, (long long)tn * WBN, ep, rowmap);
    }
}

// Two-product half-height variant (FP16X2 only): tile 128 (M) x 256 (N), 4 waves of 64 x 128,
// 2 stages of 32 KB (A hi, A lo: 128 rows; B hi: 256 rows) + a 9 KB epilogue stage, so two
// workgroups share a compute unit: one workgroup's K-step barriers and epilogue overlap the
// other's MFMAs, where the 256x256 kernel's single workgroup stalls all 8 waves at each barrier.
// Per K step a wave reads 8 A and 8 B fragments for 64 MFMAs (k_gemm_x3p<NP = 2>: 16 + 4 for 64).
// Each element sees the MFMA chain of k_gemm_x3p<f16, EPI, 2> (lo.hi, hi.hi per K step, K steps
// in order) and the same FP64 epilogue: the two give identical results.
struct EpiStageH {
    int4 row[128];
    float sb[256];
    int scored[256];
    int pslot[EPI_SLOTS];
    double cb[EPI_SLOTS];
    float n1[EPI_SLOTS][256];
};
__device__ __forceinline__ double es_base(const EpiStageH& es, int z, int c) { return es.cb[z] + (double)es.n1[z][c]; }

// stage_epilogue for a 128-row tile with 256 threads; false (uniformly) beyond EPI_SLOTS particles
// per tile (T < 19)
__device__ __forceinline__ bool stage_epilogue_h(EpiStageH& es, long long m0l, long long n0, const float* a_scale,
                                                 const float* b_scale, const EpiParams& ep, const int* rowmap, int tid)
{
    const unsigned T = (unsigned)ep.T;
    const unsigned m0 = (unsigned)m0l;
    const unsigned p0 = m0 / T;
    const int mlast = min((int)m0 + 127, ep.M - 1);
    const int nslots = (mlast >= (int)m0) ? (int)((unsigned)mlast / T - p0) + 1 : 0;
    if (nslots > EPI_SLOTS) return false;
    if (tid < 128) {
        const unsigned m = m0 + tid;
        int4 e = make_int4(-1, 0, 0, __float_as_int(1.f));
        if ((int)m < ep.M) {
            const unsigned mm = rowmap ? (unsigned)rowmap[m] : m;
            e.x = (int)mm;
            e.y = (int)(mm / T);
            e.z = (int)(m / T - p0);
            if (a_scale) e.w = __float_as_int(a_scale[mm]);
        }
        es.row[tid] = e;
    }
    {
        const int c = tid;
        const long long n = n0 + c;
        const bool nv = n < ep.N;
        es.sb[c] = (b_scale && nv) ? b_scale[n] : 1.f;
        es.scored[c] = nv && ep.prior_o[n] > 0.;
    }
    for (int q = tid; q < nslots * 256; q += 256) {
        const int sl = q >> 8, c = q & 255;
        const unsigned ms = max(m0, (p0 + sl) * T);   // a valid row of slot sl
        const unsigned p = rowmap ? (unsigned)rowmap[ms] / T : p0 + sl;
        const long long n = n0 + c;
        es.n1[sl][c] = n < ep.N ? n1_at(ep, p, n) : 0.f;
    }
    if (tid < EPI_SLOTS) {
        int p = -1;
        if (tid < nslots) {
            const unsigned ms = max(m0, (p0 + tid) * T);
            p = rowmap ? (int)((unsigned)rowmap[ms] / T) : (int)(p0 + tid);
        }
        es.pslot[tid] = p;
        es.cb[tid] = p >= 0 ? ep.Cb[p] : 0.0;
    }
    return true;
}

// ONE: the bounds form of the pruned coarse pass, one product (A hi . B hi) per element; A lo is
// neither loaded nor multiplied (the caller allows for the dropped term in its bound).
template <bool ONE>
__device__ __forceinline__ void x2h_tile(_Float16 (&sa0)[2][128 * 32], _Float16 (&sb0)[256 * 32],
                                         _Float16 (&sa1)[2][128 * 32], _Float16 (&sb1)[256 * 32], EpiStageH& es,
                                         const SplitOperand& A, const SplitOperand& B, int K, long long m0,
                                         long long n0, const EpiParams& ep, const int* __restrict__ rowmap)
{
    typedef _Float16 H;
    typedef Vec8<H>::type v8;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1;

    // acc[h][i][j]: rows 16 i of the wave's 64, columns 64 h + 16 j of its 128
    floatx4 acc[2][4][4];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[h][i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // loader: wave w moves A row blocks [32w, 32w+16), [32w+16, 32w+32) of hi and lo, and B row
    // blocks [64w + 16q, 64w + 16q + 16), q < 4; lane l fetches logical chunk (l & 3) ^ ((row >> 1) & 3)
    // of row l >> 2 into physical chunk l & 3 (the swizzle of swz())
    const int lr = lane >> 2;
    const int lch = (lane & 3) ^ ((lr >> 1) & 3);
    const int ra0 = wave * 32, ra1 = wave * 32 + 16, rb = wave * 64;
    const long long ga0 = m0 + ra0 + lr, ga1 = m0 + ra1 + lr;
    const long long am0 = rowmap ? (ga0 < ep.M ? rowmap[ga0] : 0) : ga0;
    const long long am1 = rowmap ? (ga1 < ep.M ? rowmap[ga1] : 0) : ga1;
    const char* Ah = reinterpret_cast<const char*>(A.hi);
    const char* Al = reinterpret_cast<const char*>(A.lo);
    const char* Bh = reinterpret_cast<const char*>(B.hi);
    const unsigned oa0 = (unsigned)((am0 * A.ld + lch * 8) * (long long)sizeof(H));
    const unsigned oa1 = (unsigned)((am1 * A.ld + lch * 8) * (long long)sizeof(H));
    const unsigned obq = (unsigned)(((n0 + rb + lr) * B.ld + lch * 8) * (long long)sizeof(H));
    const unsigned ob16 = (unsigned)(16 * B.ld * (long long)sizeof(H));   // 16 rows of B
#define X2H_LOAD(SA, SB, k)                                                                  \
    do {                                                                                     \
        const long long kb = (long long)(k) * (long long)sizeof(H);                          \
        lds_dma16<H>(reinterpret_cast<const H*>(Ah + kb + oa0), &SA[0][ra0 * 32]);          \
        lds_dma16<H>(reinterpret_cast<const H*>(Ah + kb + oa1), &SA[0][ra1 * 32]);          \
        if constexpr (!ONE) {                                                                \
            lds_dma16<H>(reinterpret_cast<const H*>(Al + kb + oa0), &SA[1][ra0 * 32]);      \
            lds_dma16<H>(reinterpret_cast<const H*>(Al + kb + oa1), &SA[1][ra1 * 32]);      \
        }                                                                                    \
        _Pragma("unroll") for (int q = 0; q < 4; q++)                                        \
            lds_dma16<H>(reinterpret_cast<const H*>(Bh + kb + obq + q * ob16), &SB[(rb + 16 * q) * 32]); \
    } while (0)

    const int fr = lane & 15, fc = lane >> 4;
    const int oa = swz(wm * 64 + fr, fc), ob = swz(wn * 128 + fr, fc);
#define X2H_READ_A(SA, i, AH, AL)                                   \
    do {                                                            \
        AH = *reinterpret_cast<const v8*>(&SA[0][oa + (i) * 512]);  \
        if constexpr (ONE) AL = v8{};                               \
        else AL = *reinterpret_cast<const v8*>(&SA[1][oa + (i) * 512]); \
    } while (0)
    // one A block against the 8 B fragments: lo.hi then hi.hi per element
#define X2H_GROUP(i, AL, AH)                                                                        \
    do {                                                                                            \
        if constexpr (!ONE) {                                                                       \
            _Pragma("unroll") for (int j = 0; j < 8; j++)                                           \
                acc[j >> 2][i][j & 3] = mfma16x16x32<H>(AL, bh[j], acc[j >> 2][i][j & 3]);          \
        }                                                                                           \
        _Pragma("unroll") for (int j = 0; j < 8; j++)                                               \
            acc[j >> 2][i][j & 3] = mfma16x16x32<H>(AH, bh[j], acc[j >> 2][i][j & 3]);              \
    } while (0)
    // One K step on stage (CA, CB) (B fragments and A block 0 in registers): A blocks 1..3 read one
    // group ahead, the barrier before the last group; after it stage s+2 loads into (CA, CB) and
    // the next stage's first fragments are read while group 3 runs, B fragment j after column j.
#define X2H_STEP(CA, CB, NA, NB, s)                                                \
    do {                                                                           \
        v8 nh, nl, a3h, a3l;                                                       \
        X2H_READ_A(CA, 1, nh, nl);                                                 \
        X2H_GROUP(0, al, ah);                                                      \
        ah = nh; al = nl;                                                          \
        X2H_READ_A(CA, 2, nh, nl);                                                 \
        X2H_GROUP(1, al, ah);                                                      \
        ah = nh; al = nl;                                                          \
        X2H_READ_A(CA, 3, a3h, a3l);                                               \
        X2H_GROUP(2, al, ah);                                                      \
        __builtin_amdgcn_s_waitcnt(0);                                             \
        __syncthreads();                                                           \
        if ((s) + 2 < nk) X2H_LOAD(CA, CB, ((s) + 2) * XK);                        \
        const bool nxt = (s) + 1 < nk;                                             \
        if (nxt) X2H_READ_A(NA, 0, ah, al);                                        \
        _Pragma("unroll") for (int j = 0; j < 8; j++) {                            \
            if constexpr (!ONE) acc[j >> 2][3][j & 3] = mfma16x16x32<H>(a3l, bh[j], acc[j >> 2][3][j & 3]); \
            acc[j >> 2][3][j & 3] = mfma16x16x32<H>(a3h, bh[j], acc[j >> 2][3][j & 3]); \
            if (nxt) bh[j] = *reinterpret_cast<const v8*>(&NB[ob + j * 512]);      \
        }                                                                          \
    } while (0)

    const int nk = K / XK;
    X2H_LOAD(sa0, sb0, 0);
    const bool staged = stage_epilogue_h(es, m0, n0, A.inv_scale, B.inv_scale, ep, rowmap, tid);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (nk > 1) X2H_LOAD(sa1, sb1, XK);
    v8 bh[8], ah, al;
#pragma unroll
    for (int j = 0; j < 8; j++) bh[j] = *reinterpret_cast<const v8*>(&sb0[ob + j * 512]);
    X2H_READ_A(sa0, 0, ah, al);
    for (int s = 0; s < nk; s += 2) {
        X2H_STEP(sa0, sb0, sa1, sb1, s);
        if (s + 1 < nk) X2H_STEP(sa1, sb1, sa0, sb0, s + 1);
    }
#undef X2H_STEP
#undef X2H_GROUP
#undef X2H_READ_A
#undef X2H_LOAD
    if (staged) {
        gemm_epilogue_staged<EPI_DIFF2>(acc[0], es, wm * 64, wn * 128, n0, lane, ep);
        gemm_epilogue_staged<EPI_DIFF2>(acc[1], es, wm * 64, wn * 128 + 64, n0, lane, ep);
    } else {
        gemm_epilogue<4, EPI_DIFF2>(acc[0], m0 + wm * 64, n0 + wn * 128, lane, A.inv_scale, B.inv_scale, ep, rowmap);
        gemm_epilogue<4, EPI_DIFF2>(acc[1], m0 + wm * 64, n0 + wn * 128 + 64, lane, A.inv_scale, B.inv_scale, ep,
                                    rowmap);
    }
}

template <bool ONE>
__global__ __launch_bounds__(256, 2) void k_gemm_x2h(SplitOperand A, SplitOperand B, int K, int tiles_m, int tiles_n,
                                                     int group_m, EpiParams ep, const int* __restrict__ rowmap)
{
    __shared__ __attribute__((aligned(16))) _Float16 sa0[2][128 * 32];   // stage 0: A hi, A lo
    __shared__ __attribute__((aligned(16))) _Float16 sb0[256 * 32];      // stage 0: B hi
    __shared__ __attribute__((aligned(16))) _Float16 sa1[2][128 * 32];
    __shared__ __attribute__((aligned(16))) _Float16 sb1[256 * 32];
    __shared__ EpiStageH es;
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, group_m, tm, tn);
    x2h_tile<ONE>(sa0, sb0, sa1, sb1, es, A, B, K, (long long)tm * 128, (long long)tn * 256, ep, rowmap);
}

// Learned-proposal variant: one workgroup per work item = (bin, row tile, column tile). The
// rows of a bin are the T translation rows of each particle that scores the bin (particle list
// plist[pstart .. pstart+cnt)), the columns its contiguous block of orientations [col0,
// col0+ncols) (ncols a multiple of 64, so each 64-column row minimum belongs to one work item).
// 128x128 tiles, 4 waves of 64x64, swizzled LDS, split-precision operands as k_gemm_x3w.
struct BinWork {
    int pstart, cnt, rt, col0, ncols;
};

template <class H>
__global__ __launch_bounds__(256) void k_gemm_x3_bins(SplitOperand A, SplitOperand B, int K, const int* __restrict__ plist,
                                                      const BinWork* __restrict__ work, int T, long long n_clamp,
                                                      EpiParams ep)
{
    typedef typename Vec8<H>::type v8;
    __shared__ __attribute__((aligned(16))) H lds[4][128 * 32];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const BinWork wk = work[blockIdx.x];
    const int nrows = wk.cnt * T;
    // A row of local row r (clamped to a valid row when beyond the bin's rows)
    auto arow = [&](int r) -> long long {
        const int ml = wk.rt * 128 + r;
        if (ml >= nrows) return 0;
        return (long long)plist[wk.pstart + ml / T] * T + ml % T;
    };
    const int lrow = tid >> 2, lc = tid & 3;
    const long long ma0 = arow(lrow), ma1 = arow(lrow + 64);
    const long long nb0 = min((long long)wk.col0 + lrow, n_clamp), nb1 = min((long long)wk.col0 + lrow + 64, n_clamp);
    const H* Ah = reinterpret_cast<const H*>(A.hi);
    const H* Al = reinterpret_cast<const H*>(A.lo);
    const H* Bh = reinterpret_cast<const H*>(B.hi);
    const H* Bl = reinterpret_cast<const H*>(B.lo);
    const size_t ra0 = (size_t)ma0 * A.ld + lc * 8, ra1 = (size_t)ma1 * A.ld + lc * 8;
    const size_t rb0 = (size_t)nb0 * B.ld + lc * 8, rb1 = (size_t)nb1 * B.ld + lc * 8;
    const int s0 = swz(lrow, lc), s1 = swz(lrow + 64, lc);
    floatx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int fr = lane & 15, fc = lane >> 4;
    for (int k0 = 0; k0 < K; k0 += XK) {
        const uint4 p0 = *reinterpret_cast<const uint4*>(Ah + ra0 + k0), p1 = *reinterpret_cast<const uint4*>(Ah + ra1 + k0);
        const uint4 p2 = *reinterpret_cast<const uint4*>(Al + ra0 + k0), p3 = *reinterpret_cast<const uint4*>(Al + ra1 + k0);
        const uint4 p4 = *reinterpret_cast<const uint4*>(Bh + rb0 + k0), p5 = *reinterpret_cast<const uint4*>(Bh + rb1 + k0);
        const uint4 p6 = *reinterpret_cast<const uint4*>(Bl + rb0 + k0), p7 = *reinterpret_cast<const uint4*>(Bl + rb1 + k0);
        __syncthreads();
        *reinterpret_cast<uint4*>(&lds[0][s0]) = p0;
        *reinterpret_cast<uint4*>(&lds[0][s1]) = p1;
        *reinterpret_cast<uint4*>(&lds[1][s0]) = p2;
        *reinterpret_cast<uint4*>(&lds[1][s1]) = p3;
        *reinterpret_cast<uint4*>(&lds[2][s0]) = p4;
        *reinterpret_cast<uint4*>(&lds[2][s1]) = p5;
        *reinterpret_cast<uint4*>(&lds[3][s0]) = p6;
        *reinterpret_cast<uint4*>(&lds[3][s1]) = p7;
        __syncthreads();
        v8 bh[4], bl[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = swz(wn * 64 + j * 16 + fr, fc);
            bh[j] = *reinterpret_cast<const v8*>(&lds[2][o]);
            bl[j] = *reinterpret_cast<const v8*>(&lds[3][o]);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = swz(wm * 64 + i * 16 + fr, fc);
            const v8 ah = *reinterpret_cast<const v8*>(&lds[0][o]);
            const v8 al = *reinterpret_cast<const v8*>(&lds[1][o]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(al, bh[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah, bl[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah, bh[j], acc[i][j]);
        }
    }
    // epilogue (diff2): rows gathered through the particle list
    const int col_in = lane & 15, row_grp = (lane >> 4) * 4;
    const int cbase = wn * 64;
    if (cbase >= wk.ncols) return;   // this wave's 64 columns are outside the bin
    const long long ncol0 = (long long)wk.col0 + cbase;
    float sb[4];
    bool scored[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const long long n = ncol0 + j * 16 + col_in;
        sb[j] = B.inv_scale ? B.inv_scale[n] : 1.f;
        scored[j] = ep.prior_o[n] > 0.;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ml = wk.rt * 128 + wm * 64 + i * 16 + row_grp + r;
            const bool mval = ml < nrows;
            int p = 0;
            long long m = 0;
            if (mval) {
                p = plist[wk.pstart + ml / T];
                m = (long long)p * T + ml % T;
            }
            const float sa = (mval && A.inv_scale) ? A.inv_scale[m] : 1.f;
            const double cb = mval ? ep.Cb[p] : 0.;
            float rmin = 3.0e38f;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const long long n = ncol0 + j * 16 + col_in;
                if (mval) {
                    const float v = acc[i][j][r] * sa * sb[j];
                    const float d = (float)(cb + (double)ep.N1[(long long)p * ep.ldn + n] - (double)v);
                    ep.C[m * ep.ldc + n] = d;
                    if (scored[j]) rmin = fminf(rmin, d);
                }
            }
            for (int o = 1; o < 16; o <<= 1) rmin = fminf(rmin, __shfl_xor(rmin, o));
            if (col_in == 0 && mval) ep.rowmin[m * ep.rowmin_ld + ncol0 / 64] = rmin;
        }
    }
}

// Local searches: grouped split-precision GEMM. Work item (b, s0) multiplies 128 of particle b's
// in-range slices (rows s0.., gathered from the batch's slice operand through rowu) by the
// particle's T translation rows of the image-side operand (columns; T <= 128, rows b*T..b*T+127,
// the operand holding at least 128 readable rows past the last particle), so no per-(particle,
// orientation) operand is formed. Epilogue: diff2 = C_b + n1 - G in FP64 rounded once, stored in
// the sample list (orientation-major, T per orientation), and the per-particle minimum.
template <class H>
__global__ __launch_bounds__(256) void k_gemm_local_x3(SplitOperand S, SplitOperand W, int K, const int2* __restrict__ items,
                                                       int T, const long long* __restrict__ orient_off,
                                                       const int* __restrict__ rowu, const double* __restrict__ cb,
                                                       const double* __restrict__ n1, float* __restrict__ d2,
                                                       unsigned int* __restrict__ dmin)
{
    typedef typename Vec8<H>::type v8;
    __shared__ __attribute__((aligned(16))) H Ash[GBM][XK + XPAD];
    __shared__ __attribute__((aligned(16))) H Asl[GBM][XK + XPAD];
    __shared__ __attribute__((aligned(16))) H Bsh[GBN][XK + XPAD];
    __shared__ __attribute__((aligned(16))) H Bsl[GBN][XK + XPAD];
    __shared__ float wmin[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int2 it = items[blockIdx.x];
    const int b = it.x, s0 = it.y;
    const long long o0 = orient_off[b];
    const int ncols = (int)(orient_off[b + 1] - o0);
    const long long w0 = (long long)b * T;

    floatx4 acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int lrow = tid >> 2, lc = (tid & 3) * 8;
    const H* Sh = reinterpret_cast<const H*>(S.hi);
    const H* Sl = reinterpret_cast<const H*>(S.lo);
    const H* Wh = reinterpret_cast<const H*>(W.hi);
    const H* Wl = reinterpret_cast<const H*>(W.lo);
    const int sa = s0 + lrow, sb = s0 + lrow + 64;
    const long long ua = sa < ncols ? rowu[o0 + sa] : 0, ub = sb < ncols ? rowu[o0 + sb] : 0;
    const size_t ra0 = (size_t)ua * S.ld + lc, ra1 = (size_t)ub * S.ld + lc;
    const size_t rb0 = (size_t)(w0 + lrow) * W.ld + lc, rb1 = (size_t)(w0 + lrow + 64) * W.ld + lc;
    uint4 pa0h = *reinterpret_cast<const uint4*>(Sh + ra0), pa1h = *reinterpret_cast<const uint4*>(Sh + ra1);
    uint4 pa0l = *reinterpret_cast<const uint4*>(Sl + ra0), pa1l = *reinterpret_cast<const uint4*>(Sl + ra1);
    uint4 pb0h = *reinterpret_cast<const uint4*>(Wh + rb0), pb1h = *reinterpret_cast<const uint4*>(Wh + rb1);
    uint4 pb0l = *reinterpret_cast<const uint4*>(Wl + rb0), pb1l = *reinterpret_cast<const uint4*>(Wl + rb1);

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
            pa0h = *reinterpret_cast<const uint4*>(Sh + ra0 + kn);
            pa1h = *reinterpret_cast<const uint4*>(Sh + ra1 + kn);
            pa0l = *reinterpret_cast<const uint4*>(Sl + ra0 + kn);
            pa1l = *reinterpret_cast<const uint4*>(Sl + ra1 + kn);
            pb0h = *reinterpret_cast<const uint4*>(Wh + rb0 + kn);
            pb1h = *reinterpret_cast<const uint4*>(Wh + rb1 + kn);
            pb0l = *reinterpret_cast<const uint4*>(Wl + rb0 + kn);
            pb1l = *reinterpret_cast<const uint4*>(Wl + rb1 + kn);
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
        for (int i = 0; i < 4; i++) {
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(al[i], bh[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah[i], bl[j], acc[i][j]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma16x16x32<H>(ah[i], bh[j], acc[i][j]);
        }
    }

    // epilogue: lane holds C[4*(lane/16)+r][lane%16] of each 16x16 tile; rows = slices,
    // columns = translations (16 consecutive translations per 16 lanes: coalesced stores)
    const double c_b = cb[b];
    float m = 3.0e38f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int s = s0 + wm * 64 + i * 16 + (lane >> 4) * 4 + r;
            if (s >= ncols) continue;
            const long long row = o0 + s;
            const float ss = S.inv_scale ? S.inv_scale[rowu[row]] : 1.f;
            const double base = c_b + n1[row];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int t = wn * 64 + j * 16 + (lane & 15);
                if (t >= T) continue;
                const float ws = W.inv_scale ? W.inv_scale[w0 + t] : 1.f;
                const float g = acc[i][j][r] * ss * ws;
                const float d = (float)(base - (double)g);
                d2[row * T + t] = d;
                m = fminf(m, d);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o));
    if (lane == 0) wmin[wave] = m;
    __syncthreads();
    if (tid == 0) {
        const float mm = fminf(fminf(wmin[0], wmin[1]), fminf(wmin[2], wmin[3]));
        if (mm < 3.0e38f) atomicMin(&dmin[b], float_to_ordered(mm));
    }
}

int launch_gemm_local_x3(int prec, const void* Sh, const void* Sl, const float* Ss, const void* Wh, const void* Wl,
                         const float* Ws, long long ld, int K, const int2* items, int n_items, int T,
                         const long long* orient_off, const int* rowu, const double* cb, const double* n1,
                         float* d2, unsigned int* dmin, hipStream_t st)
{
    if (K % XK != 0 || T > GBN || n_items <= 0) {
        set_error("gemm_local: K not a multiple of 32, more than 128 translations, or no work");
        return T2X_EINVAL;
    }
    SplitOperand S{Sh, Sl, Ss, ld}, W{Wh, Wl, Ws, ld};
    if (prec == T2X_PREC_BF16X3)
        k_gemm_local_x3<__bf16><<<n_items, 256, 0, st>>>(S, W, K, items, T, orient_off, rowu, cb, n1, d2, dmin);
    else
        k_gemm_local_x3<_Float16><<<n_items, 256, 0, st>>>(S, W, K, items, T, orient_off, rowu, cb, n1, d2, dmin);
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

int launch_gemm_x3_bins(int prec, const void* Ah, const void* Al, const float* As, long long lda, const void* Bh,
                        const void* Bl, const float* Bs, long long ldb, int K, const int* plist, const void* work,
                        int n_work, int T, long long n_clamp, const EpiParams& ep, hipStream_t st)
{
    if (K % XK != 0) { set_error("gemm_x3_bins: K not padded"); return T2X_EINVAL; }
    if (n_work <= 0) return T2X_OK;
    SplitOperand A{Ah, Al, As, lda}, B{Bh, Bl, Bs, ldb};
    const BinWork* w = reinterpret_cast<const BinWork*>(work);
    if (prec == T2X_PREC_BF16X3)
        k_gemm_x3_bins<__bf16><<<(unsigned)n_work, 256, 0, st>>>(A, B, K, plist, w, T, n_clamp, ep);
    else
        k_gemm_x3_bins<_Float16><<<(unsigned)n_work, 256, 0, st>>>(A, B, K, plist, w, T, n_clamp, ep);
    T2X_HIP(hipGetLastError());
    return T2X_OK;
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

// Variants of the split-precision GEMM (all give identical results: every output element sees
// the same MFMA chain and the same epilogue arithmetic): tile 128 (k_gemm_x3), 256 (k_gemm_x3w,
// 256x256), 257 (k_gemm_x3d, 256x128 with direct-to-LDS loads) or 258 (k_gemm_x3p, 256x256
// pipelined), each with a tile-walk group of `group` tile rows.
static bool valid_tile(int t) { return t == 128 || t == 256 || t == 257 || t == 258 || t == 259; }

static int launch_x3_variant(int prec, int epi, const SplitOperand& A, const SplitOperand& B, int K, int Mrows_pad,
                             int Nrows_pad, const EpiParams& ep, hipStream_t st, int tile, int g, const int* rowmap)
{
    const int tile_m = tile == 257 ? DBM : (tile == 258 ? 256 : (tile == 259 ? 128 : tile));
    const int tile_n = tile == 257 ? DBN : (tile == 258 || tile == 259 ? 256 : tile);
    if (K % XK != 0 || Mrows_pad % tile_m != 0 || Nrows_pad % tile_n != 0) {
        set_error("gemm_x3: operands not padded to tile multiples");
        return T2X_EINVAL;
    }
    const int tm = Mrows_pad / tile_m, tn = Nrows_pad / tile_n;
    const long long nwg = (long long)tm * tn;
    if (nwg > 0x7fffffffLL) { set_error("gemm: grid too large"); return T2X_EINVAL; }
    const bool bf = prec == T2X_PREC_BF16X3, store = epi == EPI_STORE;
    const unsigned grid = (unsigned)nwg;
    if (tile == 259) {   // two-product diff2 kernel (FP16X2) of 128x256 tiles
        if (prec != T2X_PREC_FP16X2 || store) {
            set_error("gemm_x3: tile 259 serves the FP16X2 diff2 GEMM only");
            return T2X_EINVAL;
        }
        k_gemm_x2h<false><<<grid, 256, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap);
    } else if (tile == 257) {
        if (bf) { if (store) k_gemm_x3d<__bf16, EPI_STORE><<<grid, 256, 0, st>>>(A, B, K, tm, tn, g, ep);
                  else k_gemm_x3d<__bf16, EPI_DIFF2><<<grid, 256, 0, st>>>(A, B, K, tm, tn, g, ep); }
        else    { if (store) k_gemm_x3d<_Float16, EPI_STORE><<<grid, 256, 0, st>>>(A, B, K, tm, tn, g, ep);
                  else k_gemm_x3d<_Float16, EPI_DIFF2><<<grid, 256, 0, st>>>(A, B, K, tm, tn, g, ep); }
    } else if (tile == 258) {
        if (bf) { if (store) k_gemm_x3p<__bf16, EPI_STORE><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap);
                  else k_gemm_x3p<__bf16, EPI_DIFF2><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap); }
        else if (prec == T2X_PREC_FP16X2) {
                  if (store) k_gemm_x3p<_Float16, EPI_STORE, 2><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap);
                  else k_gemm_x3p<_Float16, EPI_DIFF2, 2><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap); }
        else    { if (store) k_gemm_x3p<_Float16, EPI_STORE><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap);
                  else k_gemm_x3p<_Float16, EPI_DIFF2><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap); }
    } else if (tile == 256) {
        if (bf) { if (store) k_gemm_x3w<__bf16, EPI_STORE><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap);
                  else k_gemm_x3w<__bf16, EPI_DIFF2><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap); }
        else    { if (store) k_gemm_x3w<_Float16, EPI_STORE><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap);
                  else k_gemm_x3w<_Float16, EPI_DIFF2><<<grid, 512, 0, st>>>(A, B, K, tm, tn, g, ep, rowmap); }
    } else {
        if (bf) { if (store) k_gemm_x3<__bf16, EPI_STORE><<<grid, 256, 0, st>>>(A, B, K, tm, tn, g, ep);
                  else k_gemm_x3<__bf16, EPI_DIFF2><<<grid, 256, 0, st>>>(A, B, K, tm, tn, g, ep); }
        else    { if (store) k_gemm_x3<_Float16, EPI_STORE><<<grid, 256, 0, st>>>(A, B, K, tm, tn, g, ep);
                  else k_gemm_x3<_Float16, EPI_DIFF2><<<grid, 256, 0, st>>>(A, B, K, tm, tn, g, ep); }
    }
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

// Bounds form of the pruned coarse pass: block minima of the diff2 with one FP16 product per
// element (A hi . B hi), no values stored (ep.C null): k_gemm_x3p with two K chunks per LDS stage
// (the stage's lo slots hold the second chunk's hi halves), so each barrier and operand wait
// covers twice the MFMA work of a one-chunk stage.
int launch_gemm_h1_bounds(const void* Ah, const float* As, long long lda, long long a_rows, const void* Bh,
                          const float* Bs, long long ldb, int K, int Mrows_pad, int Nrows_pad, const EpiParams& ep,
                          hipStream_t st)
{
    const long long lim = 1LL << 32;
    if (a_rows * lda * 2 >= lim || (long long)Nrows_pad * ldb * 2 >= lim) {
        set_error("gemm_h1_bounds: operands beyond 4 GiB (32-bit offsets)");
        return T2X_EINVAL;
    }
    if (ep.C != nullptr || ep.pmin != nullptr) {
        set_error("gemm_h1_bounds: a bounds-only launch (no stores, no lazy minima)");
        return T2X_EINVAL;
    }
    if (K % XK != 0 || Mrows_pad % 128 != 0 || Nrows_pad % 256 != 0) {
        set_error("gemm_h1_bounds: operands not padded to tile multiples");
        return T2X_EINVAL;
    }
    // 260 (default): k_gemm_x3p NP = 4 (256x256, two K chunks per stage; K a multiple of 64);
    // 258: k_gemm_x3p NP = 1; 259 (and the fallback): k_gemm_x2h<true> (128x256)
    const char* env_t = getenv("T2X_BOUNDS_TILE");
    const int bt = env_t ? atoi(env_t) : 260;
    const bool wide = (bt == 258 || (bt == 260 && K % (2 * XK) == 0)) && Mrows_pad % 256 == 0;
    const int tm = Mrows_pad / (wide ? 256 : 128), tn = Nrows_pad / 256;
    const long long nwg = (long long)tm * tn;
    if (nwg > 0x7fffffffLL) { set_error("gemm: grid too large"); return T2X_EINVAL; }
    SplitOperand A{Ah, nullptr, As, lda}, B{Bh, nullptr, Bs, ldb};
    const char* env_g = getenv("T2X_BOUNDS_GROUP");   // tile-walk group (tile rows per group)
    const int g = env_g && atoi(env_g) > 0 ? atoi(env_g) : 32;
    if (wide && bt == 260)
        k_gemm_x3p<_Float16, EPI_DIFF2, 4><<<(unsigned)nwg, 512, 0, st>>>(A, B, K, tm, tn, g, ep, nullptr);
    else if (wide)
        k_gemm_x3p<_Float16, EPI_DIFF2, 1><<<(unsigned)nwg, 512, 0, st>>>(A, B, K, tm, tn, g, ep, nullptr);
    else
        k_gemm_x2h<true><<<(unsigned)nwg, 256, 0, st>>>(A, B, K, tm, tn, g, ep, nullptr);
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

// Pruned coarse pass, full GEMM: the tile plan (k_prune_plan over the tiles of the kernel that
// follows, from the block bounds lb and the per-particle seeds) and then the diff2 GEMM over the
// kept tiles: k_gemm_x3p (256x256 tiles; 2 products under FP16X2, else 3) on a persistent grid of
// one workgroup per compute unit. list holds list_cap entries; count (one word) must be zero on
// entry and holds the kept tiles' number on return.
int launch_gemm_pruned(int prec, const void* Ah, const void* Al, const float* As, long long lda, long long a_rows,
                       const void* Bh, const void* Bl, const float* Bs, long long ldb, int K, int Mrows_pad,
                       int Nrows_pad, const EpiParams& ep, const float* lb, const float* seed, unsigned* list,
                       long long list_cap, unsigned* count, int n_cu, hipStream_t st)
{
    const long long lim = 1LL << 32;
    if (a_rows * lda * 2 >= lim || (long long)Nrows_pad * ldb * 2 >= lim) {
        set_error("gemm_pruned: operands beyond 4 GiB (32-bit offsets)");
        return T2X_EINVAL;
    }
    const int tr = WBM, tc = WBN;
    if (K % XK != 0 || Mrows_pad % tr != 0 || Nrows_pad % tc != 0) {
        set_error("gemm_pruned: operands not padded to tile multiples");
        return T2X_EINVAL;
    }
    if (ep.C == nullptr || ep.rowmin == nullptr || ep.T <= 0 || n_cu <= 0) {
        set_error("gemm_pruned: a diff2 launch with block minima");
        return T2X_EINVAL;
    }
    const int tm = Mrows_pad / tr, tn = Nrows_pad / tc;
    const long long ntiles = (long long)tm * tn;
    if (ntiles > list_cap || ntiles > 0x7fffffffLL) { set_error("gemm_pruned: tile list too small"); return T2X_EINVAL; }
    k_prune_plan<<<(unsigned)((ntiles + 3) / 4), 256, 0, st>>>(lb, seed, ep.rowmin, ep.rowmin_ld, ep.M, ep.T, tr, tc,
                                                               tm, tn, list, count);
    T2X_HIP(hipGetLastError());
    EpiParams el = ep;
    el.tile_list = list;
    el.tile_count = count;
    SplitOperand A{Ah, Al, As, lda}, B{Bh, Bl, Bs, ldb};
    if (prec == T2X_PREC_FP16X2)
        k_gemm_x3p<_Float16, EPI_DIFF2, 2, true><<<(unsigned)n_cu, 512, 0, st>>>(A, B, K, tm, tn, 0, el, nullptr);
    else if (prec == T2X_PREC_BF16X3)
        k_gemm_x3p<__bf16, EPI_DIFF2, 3, true><<<(unsigned)n_cu, 512, 0, st>>>(A, B, K, tm, tn, 0, el, nullptr);
    else
        k_gemm_x3p<_Float16, EPI_DIFF2, 3, true><<<(unsigned)n_cu, 512, 0, st>>>(A, B, K, tm, tn, 0, el, nullptr);
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

// Autotuner. For each (device, precision, epilogue, N, K, row map or not) the first launch runs
// once untimed (so that lazy-store state has settled) and then times every eligible candidate
// twice (keeping the better run) with HIP events on the caller's stream, on the real operands
// (each run writes the same outputs); the fastest is kept for the rest of the process. N and K
// fix the per-tile work; M only scales the number of tiles, so batches of other particle counts
// reuse the choice.
// The best tile walk depends on the number of tile rows as well, so the key holds the row count
// rounded to the nearest power of two (row-mapped launches cover a varying subset of a batch).
// Whether the operands fit 32-bit byte offsets is part of the key: the row count of the key is
// not the operand's (a row-mapped launch reads a whole batch), so a choice made for small operands
// must never be reused for operands of 4 GiB or more.
struct TuneKey {
    int dev, prec, epi, N, K, mapped, mlog2, off32;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(dev, prec, epi, N, K, mapped, mlog2, off32) <
               std::tie(o.dev, o.prec, o.epi, o.N, o.K, o.mapped, o.mlog2, o.off32);
    }
};
static int rows_log2(long long m) { return m > 0 ? (int)std::lround(std::log2((double)m)) : 0; }
static std::mutex g_tune_mtx;
static std::map<TuneKey, GemmChoice> g_tune;

static const GemmChoice kCandidates[] = {{258, 32}, {258, 16}, {258, 64}, {259, 32}, {259, 16}, {259, 64},
                                         {256, 32}, {257, 32}, {128, 64}};

static int tile_rows(int tile) { return tile == 257 ? DBM : (tile == 258 ? 256 : (tile == 259 ? 128 : tile)); }
static int tile_cols(int tile) { return tile == 257 ? DBN : (tile == 258 || tile == 259 ? 256 : tile); }

static bool eligible(const GemmChoice& c, bool mapped, bool off32, int prec = 0, int epi = EPI_DIFF2)
{
    if (c.tile == 259 && (prec != T2X_PREC_FP16X2 || epi != EPI_DIFF2)) return false;   // FP16X2 diff2 only
    if (prec == T2X_PREC_FP16X2 && c.tile != 258 && c.tile != 259) return false;       // 2-product kernels
    if (c.tile == 259 && !off32) return false;                                         // 32-bit operand offsets
    if (mapped && c.tile != 258 && c.tile != 256 && c.tile != 259) return false;   // row maps: pipelined / 256 kernels
    if (c.tile == 258 && !off32) return false;                      // 32-bit operand offsets
    return true;
}

int launch_gemm_x3(int prec, int epi, const void* Ah, const void* Al, const float* As, long long lda,
                   long long a_rows, const void* Bh, const void* Bl, const float* Bs, long long ldb, int K,
                   int Mrows_pad, int Nrows_pad, const EpiParams& ep, hipStream_t st, GemmChoice want,
                   const int* rowmap, GemmChoice* used)
{
    const long long lim = 1LL << 32;   // the pipelined kernel addresses operands with 32-bit byte offsets
    const bool off32 = a_rows * lda * 2 < lim && (long long)Nrows_pad * ldb * 2 < lim;
    const bool mapped = rowmap != nullptr;
    if (prec == T2X_PREC_FP16X2 && !off32) {
        set_error("FP16X2: operands beyond 4 GiB (the 2-product kernel addresses them with 32-bit offsets)");
        return T2X_EINVAL;
    }
    SplitOperand A{Ah, Al, As, lda}, B{Bh, Bl, Bs, ldb};
    GemmChoice c = want;
    if (c.tile == 0) {   // environment override (tile sweeps), else the autotuner
        const char* e = getenv("T2X_GEMM_TILE");
        const int v = e ? atoi(e) : 0;
        if (valid_tile(v)) c = {v, 0};
    }
    if (c.tile != 0) {
        if (!valid_tile(c.tile)) { set_error("gemm_x3: unknown tile variant"); return T2X_EINVAL; }
        if (!eligible(c, mapped, off32, prec, epi)) c.tile = prec == T2X_PREC_FP16X2 ? 258 : 256;
        if (c.group <= 0) c.group = c.tile == 128 ? 64 : 32;
        if (used) *used = c;
        return launch_x3_variant(prec, epi, A, B, K, Mrows_pad, Nrows_pad, ep, st, c.tile, c.group, rowmap);
    }
    int dev = 0;
    T2X_HIP(hipGetDevice(&dev));
    const TuneKey key{dev, prec, epi, Nrows_pad, K, mapped ? 1 : 0, rows_log2(ep.M), off32 ? 1 : 0};
    {
        GemmChoice hit{0, 0};
        {
            std::lock_guard<std::mutex> lk(g_tune_mtx);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) hit = it->second;
        }
        // a cached choice is re-checked against this launch (offsets, tile multiples) before use
        if (hit.tile != 0 && eligible(hit, mapped, off32, prec, epi) && Mrows_pad % tile_rows(hit.tile) == 0 &&
            Nrows_pad % tile_cols(hit.tile) == 0) {
            if (used) *used = hit;
            return launch_x3_variant(prec, epi, A, B, K, Mrows_pad, Nrows_pad, ep, st, hit.tile, hit.group, rowmap);
        }
    }
    GemmChoice best{0, 0};
    float best_ms = 3.0e38f;
    hipEvent_t e0, e1;
    T2X_HIP(hipEventCreate(&e0));
    T2X_HIP(hipEventCreate(&e1));
    int rc = launch_x3_variant(prec, epi, A, B, K, Mrows_pad, Nrows_pad, ep, st, 258 * off32 + 256 * !off32,
                               32, rowmap);
    for (const GemmChoice& cand : kCandidates) {
        if (rc != T2X_OK) break;
        if (!eligible(cand, mapped, off32, prec, epi)) continue;
        if (Mrows_pad % tile_rows(cand.tile) || Nrows_pad % tile_cols(cand.tile)) continue;
        for (int rep = 0; rep < 2 && rc == T2X_OK; rep++) {   // best of two runs per candidate
            if (hipEventRecord(e0, st) != hipSuccess) {
                set_error("gemm_x3 autotune: event record failed");
                rc = T2X_EDEVICE;
                break;
            }
            rc = launch_x3_variant(prec, epi, A, B, K, Mrows_pad, Nrows_pad, ep, st, cand.tile, cand.group, rowmap);
            if (rc != T2X_OK) break;
            float ms = 0.f;
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                set_error("gemm_x3 autotune: event timing failed");
                rc = T2X_EDEVICE;
                break;
            }
            if (ms < best_ms) { best_ms = ms; best = cand; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != T2X_OK) return rc;
    if (best.tile == 0) { set_error("gemm_x3: no eligible tile variant"); return T2X_EINVAL; }
    {
        std::lock_guard<std::mutex> lk(g_tune_mtx);
        g_tune[key] = best;
    }
    if (used) *used = best;
    return T2X_OK;   // the candidates' runs have written the outputs
}

}  // namespace t2x
