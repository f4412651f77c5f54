// t2x_internal.h — device state and shared helpers of libt2x (MI355X / gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <atomic>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/t2x.h"

namespace t2x {

void set_error(const std::string& msg);

#define T2X_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::t2x::set_error(std::string(#call) + ": " + hipGetErrorString(e_) + " at " +  \
                             __FILE__ + ":" + std::to_string(__LINE__));                   \
            return T2X_EDEVICE;                                                            \
        }                                                                                  \
    } while (0)

#define T2X_TRY(expr)                     \
    do {                                  \
        int s_ = (expr);                  \
        if (s_ != T2X_OK) return s_;      \
    } while (0)

// A device allocation that grows on demand and is reused across calls.
// Counts every (re)allocation and release of a DevBuf: a captured graph holds raw device
// pointers, so it is replayed only while this count is unchanged.
inline std::atomic<unsigned long long> g_alloc_gen{0};
inline unsigned long long alloc_generation() { return g_alloc_gen.load(); }

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int ensure(size_t need)
    {
        if (need <= bytes) return T2X_OK;
        g_alloc_gen++;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        size_t b = need + need / 8 + 256;
        if (hipMalloc(&ptr, b) != hipSuccess) {
            set_error("hipMalloc of " + std::to_string(b) + " bytes failed");
            return T2X_ENOMEM;
        }
        bytes = b;
        return T2X_OK;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(ptr); }
    void release()
    {
        if (ptr) g_alloc_gen++;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// Page-locked host memory (hipHostMalloc): the staging area of the batch's one readback, so
// that it is a single DMA copy rather than a blit into a driver staging buffer per vector.
struct HostBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    int ensure(size_t need)
    {
        if (need <= bytes) return T2X_OK;
        release();
        size_t b = need + need / 8 + 256;
        if (hipHostMalloc(&ptr, b, hipHostMallocDefault) != hipSuccess) {
            ptr = nullptr;
            set_error("hipHostMalloc of " + std::to_string(b) + " bytes failed");
            return T2X_ENOMEM;
        }
        bytes = b;
        return T2X_OK;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(ptr); }
    void release()
    {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// One class of the model: the projector (PPref) and the backprojector (BPref).
struct ClassVolumes {
    // projector
    int p_ref_dim = 0, p_pad = 0, p_rmax = 0;
    double p_pf = 0;
    // 3D references: pad x pad x (pad/2+1) complex values in 4x4x4 bricks (vol_brick_index);
    // 2D references: the pad x (pad/2+1) plane, row-major
    float2* proj = nullptr;
    uint64_t proj_version = 0;
    // backprojector
    int b_ref_dim = 0, b_pad = 0, b_rmax = 0;
    double b_pf = 0;
    float2* bp_data = nullptr;
    float* bp_weight = nullptr;
};

// Host-side sampling grid (HealpixSampling), computed in double precision.
struct HostSampling {
    long nr_dir = 0, nr_psi = 0, nr_trans = 0;
    int over_rot = 1, over_trans = 1;     // fine-pass oversampling factors
    std::vector<double> rot, tilt, psi;   // coarse directions / psi
    std::vector<double> tx, ty;           // coarse translations (A)
    std::vector<long> ipix;
    // fine pass: sub-directions per coarse direction (over_rot/over_psi of them) and sub-psi
    int n_subdir = 1, n_subpsi = 1;
    std::vector<double> sub_rot, sub_tilt;  // [nr_dir * n_subdir]
    std::vector<double> sub_psi;            // [nr_psi * n_subpsi]
    std::vector<double> sub_dir_sc;         // [nr_dir * n_subdir][4]: sin/cos of rot, tilt
    std::vector<double> sub_psi_sc;         // [nr_psi * n_subpsi][2]: sin/cos of psi
    std::vector<double> fine_tx_px, fine_ty_px;  // [nr_trans * over_trans]
    std::vector<double> coarse_tx_px, coarse_ty_px;
};

int build_sampling(const t2x_estep_config* cfg, HostSampling& s);
int check_config(const t2x_estep_config* c);

// Cached per-configuration device tables.
struct PassTables {
    int size = 0;          // image size of the pass (coarse or current)
    int npix = 0;          // compact pixels with Minvsigma2 != 0
    int npix_pad = 0;      // padded to a multiple of 16
    int n_og = 1;          // optics groups: rows of the Minvsigma2 table
    DevBuf d_pix;          // int2 (x, y) logical frequencies, y signed
    DevBuf d_win;          // int: index into the current-size window of each compact pixel
    DevBuf d_minv;         // float Minvsigma2 of each compact pixel, one row per optics group
    DevBuf d_ires;         // int shell index of each compact pixel
    DevBuf d_inproj;       // uchar: pixel inside the projected disc x^2+y^2 <= (size/2)^2
    int px_w = 0, py0 = 0, py_h = 0;   // frequency window (x, y + py0) of the pixels
    std::vector<int> h_x, h_y, h_win, h_ires;
    std::vector<float> h_minv;
    std::vector<unsigned char> h_inproj;
};

struct Stats {
    t2x_stats s{};
};

namespace cnn { struct Net; }

struct Device;
void destroy_graph(Device* d);

struct Device {
    int id = 0;
    hipStream_t stream = nullptr;
    // this device's copy of the static staging array of t2x_projector_set_noalloc (its address
    // from hipGetSymbolAddress with the device current; the symbol itself is not a device address)
    float* noalloc_scratch = nullptr;
    hipEvent_t ev[16];
    std::map<int, ClassVolumes> classes;
    uint64_t proj_counter = 0;

    // FFT plan for batched N x N R2C
    hipfftHandle fft_plan = 0;
    int fft_n = 0, fft_batch = 0;
    // batched N x N C2R of the mask-background noise (do_zero_mask 0)
    hipfftHandle noise_plan = 0;
    int noise_n = 0, noise_batch = 0;
    DevBuf noise_spec, noise_img, sigma2_tab, dvp_tab;
    // fine pass: tiles per particle, significant-translation masks, the phase rows as the GEMM
    // operand (FP32 or hi / lo halves)
    DevBuf fine_tile_off, fine_tmask, ab_fk, ab_fh, ab_fl;
    // the batch plan (k_batch_plan: list totals, error status, work-queue counters), the
    // capacity of the significant-pair lists grown by earlier batches, and the number of compute
    // units (the persistent kernels' grids)
    DevBuf plan;
    long long plan_cap = 0;
    int n_cu = 0;
    // the batch's readback: its vectors packed into rb_dev by one kernel, then one copy to rb_host
    DevBuf rb_dev;
    HostBuf rb_host;
    // brick backprojection: item planes, frames and pixel values, the call's bricks, the batch's
    // items by class and its units of work
    DevBuf bp_filt, bp_geom, bp_val, bp_units, bp_cls, bp_boff;
    // the phase tables' sizes and translations (ab_coarse / ab_fine hold them while unchanged)
    std::vector<double> ab_key;
    DevBuf prep_part;   // per particle and row band: background sums of the mask
    // pruned coarse pass: N1 over the low-frequency prefix, its constant terms, the block lower
    // bounds, the per-particle seeds and the tile counters
    DevBuf prune_n1, prune_c, prune_seed, prune_cnt, prune_list;
    bool graph_prune = false;   // the batch graph was captured with the pruned coarse pass
    DevBuf coarse_Ai;   // projection frames of the coarse orientations, for the angles in coarse_Ai_key
    std::vector<double> coarse_Ai_key;
    // launch_graph: the batch graph captured in call graph_call for batches of graph_B
    // particles (valid while alloc_generation() == graph_gen), and the batch sizes the current
    // call has run eagerly
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_B = 0;
    unsigned long long graph_gen = 0, graph_call = 0, call_counter = 0;
    std::set<int> eager_B;

    // cached coarse projections
    std::string coarse_key;
    DevBuf coarse_B;      // [O_all_pad][K_pad] float: re/im of the coarse slices
    DevBuf coarse_B2;     // [O_all_pad][Kn_pad] float: |F|^2 of the coarse slices
    DevBuf coarse_Bh, coarse_Bl, coarse_Bs;   // 16-bit hi/lo split of coarse_B + row scales
    DevBuf coarse_B2h, coarse_B2l, coarse_B2s;   // the same for coarse_B2 (|F|^2 of the N1 term)
    DevBuf A1h, A1l, A1s;                        // the same for A1 (per batch)
    int coarse_O_all = 0, coarse_O_pad = 0;

    // scratch
    DevBuf img, img2, fft_out, Fimg_cur, Fimg_nm_cur, Fctf_cur, pscal, power;
    DevBuf A, A1, N1, diff2, minb, sel_out, prior_o, prior_t, cls_prior, sig_pairs, sig_count, sig_orient,
        sig_orient_count, orient_pairs_off, fine_items, fine_diff2, fine_min, fine_sel, ab_fine,
        ab_coarse, wsum_dev, store_items, best, tables, Y_fine, dbg_tmp, cand_w, cand_e, cand_n,
        pairs_strided, Ah, Al, As, sum_part, log_prior_o, pmin;
    DevBuf loc_dir, loc_psi, loc_mean, loc_ang, loc_tab, loc_list, loc_nsel;   // local searches
    // coarse pass of a local search on per-particle sample lists
    DevBuf lc_orients, lc_pairs, lc_d2, lc_min, lc_pair_off, lc_orient_off, lc_cls, lc_dirsc, lc_psisc,
        lc_n1, lc_G, lc_z, lc_ab, lc_ulist, lc_rowu, lc_F, lc_F2,
        lc_fh, lc_fl, lc_fs, lc_items, lc_flags, lc_pos, lc_scan_tmp;
    // per-call tables and per-batch offsets (grow-only, so no allocation per call or batch)
    DevBuf minv_fd, dirsc, psisc, ftx, fty, vols_tab, pads_tab, bpd_tab, bpw_tab, bpads_tab, bprmax_tab,
        ctx, cty, ctfc, pair_off_d, orient_off_d, item_off_d;
    PassTables coarse_tab, fine_tab;
    std::string tab_key;

    Stats stats;
    std::mutex mtx;
    std::shared_ptr<cnn::Net> net;   // pose-proposal network (t2x_proposal_*)
    DevBuf prop_crop, prop_img, prop_out, prop_mask, prop_plist, prop_work, prop_rows;
    hipfftHandle prop_plan = 0;
    int prop_plan_b = 0;
    std::vector<float> prop_gate_top1, prop_gate_mass;   // per particle of the last E-step call
};

// ---------------- device helpers ----------------

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Counter-based Gaussian sample keyed by (seed, particle, index): Box-Muller on two 53-bit
// uniforms from the splitmix64 finaliser.
__device__ __forceinline__ double noise_sample(uint64_t seed, long particle, long index)
{
    uint64_t k = mix64(seed ^ mix64((uint64_t)particle * 0x100000001B3ULL + 0x51ED27ULL));
    uint64_t a = mix64(k + 2 * (uint64_t)index);
    uint64_t b = mix64(k + 2 * (uint64_t)index + 1);
    double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740992.0);
    double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

// Order-preserving map of a float to uint32 for atomicMin.
__device__ __forceinline__ unsigned int float_to_ordered(float f)
{
    unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float ordered_to_float(unsigned int u)
{
    unsigned int v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

// Euler_angles2matrix (euler.cpp:52-91) from precomputed sines and cosines, with the
// reference's operation order and no fused multiply-add, so that host and device produce the
// same bits as the CPU path (whose x86-64 build does not contract a*b+c).
__host__ __device__ inline void euler_from_sincos(double sa, double ca, double sb, double cb,
                                                  double sg, double cg, double* A)
{
#pragma clang fp contract(off)
    double cc = cb * ca, cs = cb * sa, sc = sb * ca, ss = sb * sa;
    A[0] = cg * cc - sg * sa;  A[1] = cg * cs + sg * ca;  A[2] = -cg * sb;
    A[3] = -sg * cc - cg * sa; A[4] = -sg * cs + cg * ca; A[5] = sg * sb;
    A[6] = sc;                 A[7] = ss;                 A[8] = cb;
}

// Host only: angles in degrees, DEG2RAD(d) = ((d) * PI / 180) (macros.h:275).
inline void euler_matrix_host(double rot, double tilt, double psi, double* A)
{
    const double kP = 3.14159265358979323846;
    double a = rot * kP / 180, b = tilt * kP / 180, g = psi * kP / 180;
    euler_from_sincos(std::sin(a), std::cos(a), std::sin(b), std::cos(b), std::sin(g), std::cos(g), A);
}

// Matrix2D::inv for 3x3 matrices (matrix2d.h:1108-1123): adjugate over the determinant,
// reference operation order, no contraction.
__host__ __device__ inline void inv3(const double* A, double* R)
{
#pragma clang fp contract(off)
    R[0] = A[8] * A[4] - A[7] * A[5];
    R[1] = -(A[8] * A[1] - A[7] * A[2]);
    R[2] = A[5] * A[1] - A[4] * A[2];
    R[3] = -(A[8] * A[3] - A[6] * A[5]);
    R[4] = A[8] * A[0] - A[6] * A[2];
    R[5] = -(A[5] * A[0] - A[3] * A[2]);
    R[6] = A[7] * A[3] - A[6] * A[4];
    R[7] = -(A[7] * A[0] - A[6] * A[1]);
    R[8] = A[4] * A[0] - A[3] * A[1];
    double tmp = A[0] * R[0] + A[3] * R[1] + A[6] * R[2];
    for (int i = 0; i < 9; i++) R[i] /= tmp;
}

// Trilinear sample of the half-complex projector volume at logical (xp, yp, zp) (already
// multiplied by the padding factor), with the Friedel flip of projector.cpp:583-640.
// Returns false where the reference skips the pixel (outside the box).
// Element (x, y, z) of a projector volume stored in 4x4x4 bricks (x fastest inside a brick,
// bricks x fastest): the 8 corners of a trilinear sample lie in at most 8 bricks of 512 bytes
// whatever the slice orientation, instead of 4 rows and 2 planes far apart.
__host__ __device__ __forceinline__ size_t vol_brick_index(int x, int y, int z, int nbx, int nby)
{
    return ((((size_t)(z >> 2) * nby + (y >> 2)) * nbx + (x >> 2)) << 6) + (((z & 3) * 4 + (y & 3)) * 4 + (x & 3));
}

__host__ __device__ __forceinline__ size_t vol_brick_elems(int pad)
{
    const size_t nbx = (pad / 2 + 1 + 3) / 4, nb = (pad + 3) / 4;
    return nbx * nb * nb * 64;
}

__device__ __forceinline__ bool trilinear_sample(const float2* __restrict__ vol, int pad, float xp,
                                                 float yp, float zp, float2* out)
{
    const int xdim = pad / 2 + 1, h = pad / 2;
    bool neg = xp < 0.f;
    if (neg) { xp = -xp; yp = -yp; zp = -zp; }
    float fx0 = floorf(xp), fy0 = floorf(yp), fz0 = floorf(zp);
    int x0 = (int)fx0, y0 = (int)fy0 + h, z0 = (int)fz0 + h;
    float fx = xp - fx0, fy = yp - fy0, fz = zp - fz0;
    if (x0 < 0 || x0 + 1 >= xdim || y0 < 0 || y0 + 1 >= pad || z0 < 0 || z0 + 1 >= pad) return false;
    const int nbx = (xdim + 3) / 4, nby = (pad + 3) / 4;
    float2 d000 = vol[vol_brick_index(x0, y0, z0, nbx, nby)], d001 = vol[vol_brick_index(x0 + 1, y0, z0, nbx, nby)];
    float2 d010 = vol[vol_brick_index(x0, y0 + 1, z0, nbx, nby)];
    float2 d011 = vol[vol_brick_index(x0 + 1, y0 + 1, z0, nbx, nby)];
    float2 d100 = vol[vol_brick_index(x0, y0, z0 + 1, nbx, nby)];
    float2 d101 = vol[vol_brick_index(x0 + 1, y0, z0 + 1, nbx, nby)];
    float2 d110 = vol[vol_brick_index(x0, y0 + 1, z0 + 1, nbx, nby)];
    float2 d111 = vol[vol_brick_index(x0 + 1, y0 + 1, z0 + 1, nbx, nby)];
    float2 dx00 = make_float2(d000.x + (d001.x - d000.x) * fx, d000.y + (d001.y - d000.y) * fx);
    float2 dx01 = make_float2(d100.x + (d101.x - d100.x) * fx, d100.y + (d101.y - d100.y) * fx);
    float2 dx10 = make_float2(d010.x + (d011.x - d010.x) * fx, d010.y + (d011.y - d010.y) * fx);
    float2 dx11 = make_float2(d110.x + (d111.x - d110.x) * fx, d110.y + (d111.y - d110.y) * fx);
    float2 dxy0 = make_float2(dx00.x + (dx10.x - dx00.x) * fy, dx00.y + (dx10.y - dx00.y) * fy);
    float2 dxy1 = make_float2(dx01.x + (dx11.x - dx01.x) * fy, dx01.y + (dx11.y - dx01.y) * fy);
    float2 v = make_float2(dxy0.x + (dxy1.x - dxy0.x) * fz, dxy0.y + (dxy1.y - dxy0.y) * fz);
    if (neg) v.y = -v.y;
    *out = v;
    return true;
}

// Bilinear sample of a half-complex 2D reference ([pad][pad/2+1], y centred) at logical
// (xp, yp) (already multiplied by the padding factor), with the Friedel flip of
// Projector::rotate2D (projector.cpp:955-995).
__device__ __forceinline__ bool bilinear_sample(const float2* __restrict__ ref, int pad, float xp, float yp,
                                                float2* out)
{
    const int xdim = pad / 2 + 1, h = pad / 2;
    bool neg = xp < 0.f;
    if (neg) { xp = -xp; yp = -yp; }
    float fx0 = floorf(xp), fy0 = floorf(yp);
    int x0 = (int)fx0, y0 = (int)fy0 + h;
    float fx = xp - fx0, fy = yp - fy0;
    if (x0 < 0 || x0 + 1 >= xdim || y0 < 0 || y0 + 1 >= pad) return false;
    const float2* p = ref + (size_t)y0 * xdim + x0;
    float2 d00 = p[0], d01 = p[1], d10 = p[xdim], d11 = p[xdim + 1];
    float2 dx0 = make_float2(d00.x + (d01.x - d00.x) * fx, d00.y + (d01.y - d00.y) * fx);
    float2 dx1 = make_float2(d10.x + (d11.x - d10.x) * fx, d10.y + (d11.y - d10.y) * fx);
    float2 v = make_float2(dx0.x + (dx1.x - dx0.x) * fy, dx0.y + (dx1.y - dx0.y) * fy);
    if (neg) v.y = -v.y;
    *out = v;
    return true;
}

// Slice value of a 3D reference (trilinear) or of a 2D reference (bilinear, zp unused).
__device__ __forceinline__ bool ref_sample(const float2* __restrict__ vol, int pad, int ref2d, float xp, float yp,
                                           float zp, float2* out)
{
    return ref2d ? bilinear_sample(vol, pad, xp, yp, out) : trilinear_sample(vol, pad, xp, yp, zp, out);
}

// CTF::getCTF (ctf.h:181-230) with precomputed constants; double precision.
struct CtfConst {
    double K1, K2, K3, K4, K5, Axx, Axy, Ayy, scale;
};

inline CtfConst ctf_const(const t2x_ctf_params& p)
{
    CtfConst c;
    double local_Cs = p.cs_mm * 1e7, local_kV = p.voltage_kv * 1e3;
    double rad = p.defocus_angle * 3.14159265358979323846 / 180.;
    double lambda = 12.2643247 / std::sqrt(local_kV * (1. + local_kV * 0.978466e-6));
    c.K1 = 3.14159265358979323846 / 2 * 2 * lambda;
    c.K2 = 3.14159265358979323846 / 2 * local_Cs * lambda * lambda * lambda;
    c.K3 = std::atan(p.q0 / std::sqrt(1 - p.q0 * p.q0));
    c.K4 = -p.bfactor / 4.;
    c.K5 = p.phase_shift * 3.14159265358979323846 / 180.;
    double s = std::sin(rad), co = std::cos(rad);
    double dU = -p.defocus_u, dV = -p.defocus_v;
    c.Axx = co * co * dU + s * s * dV;
    c.Axy = co * s * dU - s * co * dV;
    c.Ayy = s * s * dU + co * co * dV;
    c.scale = p.scale;
    return c;
}

__device__ __forceinline__ double ctf_value(const CtfConst& c, double X, double Y)
{
    double u2 = X * X + Y * Y;
    double gamma = c.K1 * (c.Axx * X * X + 2.0 * c.Axy * X * Y + c.Ayy * Y * Y) + c.K2 * u2 * u2 - c.K5 - c.K3;
    double r = -sin(gamma) * exp(c.K4 * u2) * c.scale;
    if (fabs(r) < 1e-8) r = (r >= 0 ? 1e-8 : -1e-8);
    return r;
}

}  // namespace t2x
