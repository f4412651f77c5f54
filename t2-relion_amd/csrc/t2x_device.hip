// t2x_device.hip — device bundle, projector and backprojector volumes, memory helpers.
//
// Replaces MlDeviceBundle (src/acc/cuda/cuda_ml_optimiser.h:18-81, .cu:1153-1313),
// AccProjector::setMdlDim/initMdl (src/acc/acc_projector_impl.h:10-223) and
// AccBackprojector::setMdlDim/initMdl/getMdlData/clear (src/acc/acc_backprojector_impl.h:8-143).
// Volumes live in plain HBM arrays (no textures): the E-step kernels do their own exact
// trilinear interpolation, as the CPU path does (projector.cpp:585-640).
#include "t2x_internal.h"
#include <algorithm>
#include <thread>

namespace t2x {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace t2x

using namespace t2x;

namespace t2x {

// Staging array of t2x_projector_set_noalloc: the re and im planes of one chunk of
// kNoallocChunk elements each (2 x 32 MiB per GPU, held by every GPU the library runs on until
// the process ends). Host code never passes the symbol itself to a kernel or a copy (on the host
// it names a stand-in, not device memory): each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr size_t kNoallocChunk = size_t(8) << 20;
static __device__ __align__(256) float scratch[2 * kNoallocChunk];

static std::mutex g_scratch_mtx;
static std::map<int, float*> g_scratch_addr;   // device id -> address of that device's copy

// Called with device_id current (t2x_device_open).
static int noalloc_scratch_address(int device_id, float** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    auto it = g_scratch_addr.find(device_id);
    if (it != g_scratch_addr.end()) {
        *out = it->second;
        return T2X_OK;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("hipGetSymbolAddress of the staging array failed: ") + hipGetErrorString(e));
        return T2X_EDEVICE;
    }
    g_scratch_addr[device_id] = static_cast<float*>(p);
    *out = static_cast<float*>(p);
    return T2X_OK;
}

}  // namespace t2x

extern "C" const char* t2x_last_error(void) { return g_last_error.c_str(); }
extern "C" int t2x_abi_version(void) { return T2X_ABI_VERSION; }

extern "C" int t2x_device_open(int device_id, t2x_device** out)
{
    if (!out) { set_error("null output handle"); return T2X_EINVAL; }
    *out = nullptr;
    int n = 0;
    T2X_HIP(hipGetDeviceCount(&n));
    if (device_id < 0 || device_id >= n) {
        set_error("device id " + std::to_string(device_id) + " out of range (" + std::to_string(n) + " devices)");
        return T2X_EINVAL;
    }
    T2X_HIP(hipSetDevice(device_id));
    float* scratch_ptr = nullptr;
    T2X_TRY(noalloc_scratch_address(device_id, &scratch_ptr));
    int n_cu = 0;
    T2X_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device_id));
    auto* d = new Device();
    d->id = device_id;
    d->noalloc_scratch = scratch_ptr;
    d->n_cu = n_cu > 0 ? n_cu : 256;
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        delete d;
        set_error("hipStreamCreate failed");
        return T2X_EDEVICE;
    }
    for (auto& e : d->ev) (void)hipEventCreate(&e);
    *out = reinterpret_cast<t2x_device*>(d);
    return T2X_OK;
}

static void release_class(ClassVolumes& c)
{
    if (c.proj) (void)hipFree(c.proj);
    if (c.bp_data) (void)hipFree(c.bp_data);
    if (c.bp_weight) (void)hipFree(c.bp_weight);
    c.proj = nullptr;
    c.bp_data = nullptr;
    c.bp_weight = nullptr;
}

void t2x::destroy_graph(Device* d)
{
    if (d->graph_exec) (void)hipGraphExecDestroy(d->graph_exec);
    if (d->graph) (void)hipGraphDestroy(d->graph);
    d->graph_exec = nullptr;
    d->graph = nullptr;
    d->graph_B = 0;
}

extern "C" int t2x_device_close(t2x_device* h)
{
    if (!h) return T2X_OK;
    Device* d = reinterpret_cast<Device*>(h);
    (void)hipSetDevice(d->id);
    (void)hipStreamSynchronize(d->stream);
    destroy_graph(d);
    for (auto& kv : d->classes) release_class(kv.second);
    DevBuf* bufs[] = {&d->img, &d->img2, &d->fft_out, &d->Fimg_cur, &d->Fimg_nm_cur, &d->Fctf_cur,
                      &d->pscal, &d->power, &d->A, &d->A1, &d->N1, &d->diff2, &d->minb,
                      &d->sel_out, &d->prior_o, &d->prior_t, &d->cls_prior, &d->sig_pairs, &d->sig_count,
                      &d->sig_orient, &d->sig_orient_count, &d->orient_pairs_off, &d->fine_items,
                      &d->fine_diff2, &d->fine_min, &d->fine_sel, &d->ab_fine, &d->ab_coarse,
                      &d->wsum_dev, &d->store_items, &d->best, &d->tables, &d->Y_fine,
                      &d->dbg_tmp, &d->coarse_B, &d->coarse_B2, &d->cand_w, &d->cand_e,
                      &d->cand_n, &d->pairs_strided, &d->loc_dir, &d->loc_psi, &d->loc_mean,
                      &d->loc_ang, &d->loc_tab, &d->loc_list};
    for (DevBuf* b : bufs) b->release();
    PassTables* tabs[] = {&d->coarse_tab, &d->fine_tab};
    for (PassTables* t : tabs) {
        t->d_pix.release(); t->d_win.release(); t->d_minv.release(); t->d_ires.release(); t->d_inproj.release();
    }
    if (d->fft_plan) (void)hipfftDestroy(d->fft_plan);
    if (d->noise_plan) (void)hipfftDestroy(d->noise_plan);
    if (d->prop_plan) (void)hipfftDestroy(d->prop_plan);
    d->net.reset();
    for (auto& e : d->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(d->stream);
    delete d;
    return T2X_OK;
}

static size_t vol_elems(int ref_dim, int pad)
{
    size_t xdim = pad / 2 + 1;
    return ref_dim == 3 ? (size_t)pad * pad * xdim : (size_t)pad * xdim;
}

static int check_dims(int iclass, int ref_dim, int pad, int r_max, double pf)
{
    if (iclass < 0) { set_error("negative class index"); return T2X_EINVAL; }
    if (ref_dim != 3 && ref_dim != 2) { set_error("ref_dim must be 3 or 2"); return T2X_EINVAL; }
    if (pad < 3 || !(pad & 1)) { set_error("pad_size must be odd"); return T2X_EINVAL; }
    if (r_max < 1 || pf <= 0) { set_error("bad r_max or padding factor"); return T2X_EINVAL; }
    long need = 2L * ((long)std::lround(pf * r_max) + 1) + 1;
    if (pad < need) { set_error("pad_size smaller than 2*(ROUND(pf*r_max)+1)+1"); return T2X_EINVAL; }
    return T2X_OK;
}

// Interleaves re/im into the projector layout: a 2D plane row-major, a 3D volume in 4x4x4 bricks.
// Elements [base, base + n) of the caller's layout, re[k]/im[k] holding element base + k.
__global__ void k_interleave(const float* __restrict__ re, const float* __restrict__ im,
                             float2* __restrict__ out, size_t base, size_t n, int ref_dim, int pad)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t i = base + k;
    if (ref_dim == 2) { out[i] = make_float2(re[k], im[k]); return; }
    const int xdim = pad / 2 + 1;
    const int x = (int)(i % xdim), y = (int)((i / xdim) % pad), z = (int)(i / ((size_t)xdim * pad));
    out[vol_brick_index(x, y, z, (xdim + 3) / 4, (pad + 3) / 4)] = make_float2(re[k], im[k]);
}

extern "C" int t2x_projector_set(t2x_device* h, int iclass, int ref_dim, int pad_size, int r_max,
                                 double pf, const float* re, const float* im)
{
    if (!h || !re || !im) { set_error("null argument"); return T2X_EINVAL; }
    T2X_TRY(check_dims(iclass, ref_dim, pad_size, r_max, pf));
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    T2X_HIP(hipSetDevice(d->id));
    ClassVolumes& c = d->classes[iclass];
    size_t n = vol_elems(ref_dim, pad_size);
    if (c.proj && (c.p_pad != pad_size || c.p_ref_dim != ref_dim)) {
        (void)hipFree(c.proj);
        c.proj = nullptr;
    }
    if (!c.proj) {
        const size_t ne = ref_dim == 3 ? vol_brick_elems(pad_size) : n;
        if (hipMalloc(&c.proj, ne * sizeof(float2)) != hipSuccess) {
            set_error("hipMalloc projector volume failed");
            return T2X_ENOMEM;
        }
    }
    c.p_ref_dim = ref_dim;
    c.p_pad = pad_size;
    c.p_rmax = r_max;
    c.p_pf = pf;
    c.proj_version = ++d->proj_counter;
    d->coarse_key.clear();
    // Stage the two planes and interleave on the device. Pointers may be host or device memory.
    float *dre = nullptr, *dim = nullptr;
    if (hipMalloc(&dre, n * sizeof(float)) != hipSuccess || hipMalloc(&dim, n * sizeof(float)) != hipSuccess) {
        if (dre) (void)hipFree(dre);
        set_error("hipMalloc staging failed");
        return T2X_ENOMEM;
    }
    T2X_HIP(hipMemcpyAsync(dre, re, n * sizeof(float), hipMemcpyDefault, d->stream));
    T2X_HIP(hipMemcpyAsync(dim, im, n * sizeof(float), hipMemcpyDefault, d->stream));
    k_interleave<<<(unsigned)((n + 255) / 256), 256, 0, d->stream>>>(dre, dim, c.proj, 0, n, ref_dim, pad_size);
    T2X_HIP(hipGetLastError());
    T2X_HIP(hipStreamSynchronize(d->stream));
    (void)hipFree(dre);
    (void)hipFree(dim);
    return T2X_OK;
}

// The same upload without any allocation: the volume must exist (an earlier t2x_projector_set of
// this class with the same ref_dim and pad_size) and the planes pass through the device's static
// staging array in chunks of kNoallocChunk elements, so every volume size is served.
extern "C" int t2x_projector_set_noalloc(t2x_device* h, int iclass, int ref_dim, int pad_size, int r_max,
                                         double pf, const float* re, const float* im)
{
    if (!h || !re || !im) { set_error("null argument"); return T2X_EINVAL; }
    T2X_TRY(check_dims(iclass, ref_dim, pad_size, r_max, pf));
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    auto it = d->classes.find(iclass);
    if (it == d->classes.end() || !it->second.proj) {
        set_error("t2x_projector_set_noalloc: class " + std::to_string(iclass) +
                  " has no projector volume (it would need an allocation: call t2x_projector_set once first)");
        return T2X_ESTATE;
    }
    ClassVolumes& c = it->second;
    if (c.p_pad != pad_size || c.p_ref_dim != ref_dim) {
        set_error("t2x_projector_set_noalloc: class " + std::to_string(iclass) + " holds a volume of ref_dim " +
                  std::to_string(c.p_ref_dim) + " and pad_size " + std::to_string(c.p_pad) +
                  "; another shape would need a new allocation");
        return T2X_EINVAL;
    }
    if (!d->noalloc_scratch) { set_error("staging array address not set on this handle"); return T2X_ESTATE; }
    T2X_HIP(hipSetDevice(d->id));
    c.p_rmax = r_max;
    c.p_pf = pf;
    c.proj_version = ++d->proj_counter;
    d->coarse_key.clear();
    const size_t n = vol_elems(ref_dim, pad_size);
    float* sre = d->noalloc_scratch;
    float* sim = d->noalloc_scratch + kNoallocChunk;
    // stream order makes the next chunk's copies wait for the previous chunk's kernel
    for (size_t i0 = 0; i0 < n; i0 += kNoallocChunk) {
        const size_t m = std::min(kNoallocChunk, n - i0);
        T2X_HIP(hipMemcpyAsync(sre, re + i0, m * sizeof(float), hipMemcpyDefault, d->stream));
        T2X_HIP(hipMemcpyAsync(sim, im + i0, m * sizeof(float), hipMemcpyDefault, d->stream));
        k_interleave<<<(unsigned)((m + 255) / 256), 256, 0, d->stream>>>(sre, sim, c.proj, i0, m, ref_dim, pad_size);
        T2X_HIP(hipGetLastError());
    }
    T2X_HIP(hipStreamSynchronize(d->stream));
    return T2X_OK;
}

// Projector volume of class iclass in the caller's layout (tests only).
extern "C" int t2x_debug_projector_get(t2x_device* h, int iclass, float* re, float* im)
{
    if (!h || !re || !im) { set_error("null argument"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    auto it = d->classes.find(iclass);
    if (it == d->classes.end() || !it->second.proj) { set_error("projector not set"); return T2X_ESTATE; }
    const ClassVolumes& c = it->second;
    const size_t n = vol_elems(c.p_ref_dim, c.p_pad);
    const size_t ne = c.p_ref_dim == 3 ? vol_brick_elems(c.p_pad) : n;
    T2X_HIP(hipSetDevice(d->id));
    T2X_HIP(hipStreamSynchronize(d->stream));
    std::vector<float2> v(ne);
    T2X_HIP(hipMemcpy(v.data(), c.proj, ne * sizeof(float2), hipMemcpyDeviceToHost));
    const int pad = c.p_pad, xdim = pad / 2 + 1;
    for (size_t i = 0; i < n; i++) {
        size_t src = i;
        if (c.p_ref_dim == 3) {
            const int x = (int)(i % xdim), y = (int)((i / xdim) % pad), z = (int)(i / ((size_t)xdim * pad));
            src = vol_brick_index(x, y, z, (xdim + 3) / 4, (pad + 3) / 4);
        }
        re[i] = v[src].x;
        im[i] = v[src].y;
    }
    return T2X_OK;
}

extern "C" int t2x_backprojector_init(t2x_device* h, int iclass, int ref_dim, int pad_size,
                                      int r_max, double pf)
{
    if (!h) { set_error("null handle"); return T2X_EINVAL; }
    T2X_TRY(check_dims(iclass, ref_dim, pad_size, r_max, pf));
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    T2X_HIP(hipSetDevice(d->id));
    ClassVolumes& c = d->classes[iclass];
    size_t n = vol_elems(ref_dim, pad_size);
    if (c.bp_data && (c.b_pad != pad_size || c.b_ref_dim != ref_dim)) {
        (void)hipFree(c.bp_data);
        (void)hipFree(c.bp_weight);
        c.bp_data = nullptr;
        c.bp_weight = nullptr;
    }
    if (!c.bp_data) {
        if (hipMalloc(&c.bp_data, n * sizeof(float2)) != hipSuccess ||
            hipMalloc(&c.bp_weight, n * sizeof(float)) != hipSuccess) {
            set_error("hipMalloc backprojector volumes failed");
            return T2X_ENOMEM;
        }
    }
    c.b_ref_dim = ref_dim;
    c.b_pad = pad_size;
    c.b_rmax = r_max;
    c.b_pf = pf;
    T2X_HIP(hipMemsetAsync(c.bp_data, 0, n * sizeof(float2), d->stream));
    T2X_HIP(hipMemsetAsync(c.bp_weight, 0, n * sizeof(float), d->stream));
    T2X_HIP(hipStreamSynchronize(d->stream));
    return T2X_OK;
}

extern "C" int t2x_backprojector_clear(t2x_device* h, int iclass)
{
    if (!h) { set_error("null handle"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    auto it = d->classes.find(iclass);
    if (it == d->classes.end() || !it->second.bp_data) { set_error("backprojector not initialised"); return T2X_ESTATE; }
    ClassVolumes& c = it->second;
    size_t n = vol_elems(c.b_ref_dim, c.b_pad);
    T2X_HIP(hipSetDevice(d->id));
    T2X_HIP(hipMemsetAsync(c.bp_data, 0, n * sizeof(float2), d->stream));
    T2X_HIP(hipMemsetAsync(c.bp_weight, 0, n * sizeof(float), d->stream));
    T2X_HIP(hipStreamSynchronize(d->stream));
    return T2X_OK;
}

extern "C" int t2x_backprojector_get(t2x_device* h, int iclass, double* re, double* im, double* w)
{
    if (!h) { set_error("null handle"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    std::lock_guard<std::mutex> lk(d->mtx);
    auto it = d->classes.find(iclass);
    if (it == d->classes.end() || !it->second.bp_data) { set_error("backprojector not initialised"); return T2X_ESTATE; }
    ClassVolumes& c = it->second;
    size_t n = vol_elems(c.b_ref_dim, c.b_pad);
    T2X_HIP(hipSetDevice(d->id));
    // both volumes in one page-locked staging area (two DMA copies, no driver staging)
    T2X_TRY(d->rb_host.ensure(n * (sizeof(float2) + sizeof(float))));
    float2* hd = d->rb_host.as<float2>();
    float* hw = reinterpret_cast<float*>(hd + n);
    T2X_HIP(hipMemcpyAsync(hd, c.bp_data, n * sizeof(float2), hipMemcpyDeviceToHost, d->stream));
    T2X_HIP(hipMemcpyAsync(hw, c.bp_weight, n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    T2X_HIP(hipStreamSynchronize(d->stream));
    // FP32 -> FP64 into the caller's arrays, split over host threads (a 3D volume holds millions
    // of voxels: one thread alone is held by memory latency)
    auto convert = [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; i++) {
            if (re) re[i] = hd[i].x;
            if (im) im[i] = hd[i].y;
            if (w) w[i] = hw[i];
        }
    };
    const size_t nthr = std::min<size_t>(8, std::max<size_t>(1, n >> 18));
    std::vector<std::thread> th;
    for (size_t t = 1; t < nthr; t++) th.emplace_back(convert, n * t / nthr, n * (t + 1) / nthr);
    convert(0, n / nthr);
    for (auto& x : th) x.join();
    return T2X_OK;
}

extern "C" int t2x_stats_get(t2x_device* h, t2x_stats* st)
{
    if (!h || !st) { set_error("null argument"); return T2X_EINVAL; }
    *st = reinterpret_cast<Device*>(h)->stats.s;
    return T2X_OK;
}

extern "C" int t2x_stats_reset(t2x_device* h)
{
    if (!h) { set_error("null handle"); return T2X_EINVAL; }
    reinterpret_cast<Device*>(h)->stats.s = t2x_stats{};
    return T2X_OK;
}

extern "C" int t2x_device_alloc(t2x_device* h, size_t bytes, void** p)
{
    if (!h || !p) { set_error("null argument"); return T2X_EINVAL; }
    T2X_HIP(hipSetDevice(reinterpret_cast<Device*>(h)->id));
    if (hipMalloc(p, bytes) != hipSuccess) { set_error("hipMalloc failed"); return T2X_ENOMEM; }
    return T2X_OK;
}

extern "C" int t2x_device_free(t2x_device* h, void* p)
{
    if (!h) { set_error("null handle"); return T2X_EINVAL; }
    T2X_HIP(hipFree(p));
    return T2X_OK;
}

extern "C" int t2x_memcpy_h2d(t2x_device* h, void* dst, const void* src, size_t bytes)
{
    if (!h) { set_error("null handle"); return T2X_EINVAL; }
    T2X_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return T2X_OK;
}

extern "C" int t2x_memcpy_d2h(t2x_device* h, void* dst, const void* src, size_t bytes)
{
    if (!h) { set_error("null handle"); return T2X_EINVAL; }
    T2X_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return T2X_OK;
}

__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" int t2x_device_copy(t2x_device* h, void* dst, const void* src, size_t bytes)
{
    if (!h || !dst || !src) { set_error("null argument"); return T2X_EINVAL; }
    if (bytes % 16 || (reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16) {
        set_error("t2x_device_copy: bytes and pointers must be multiples of 16");
        return T2X_EINVAL;
    }
    Device* d = reinterpret_cast<Device*>(h);
    const size_t n = bytes / 16;
    if (n == 0) return T2X_OK;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
    k_copy16<<<grid, 256, 0, d->stream>>>(reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst), n);
    T2X_HIP(hipGetLastError());
    return T2X_OK;
}

extern "C" int t2x_device_synchronize(t2x_device* h)
{
    if (!h) { set_error("null handle"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    T2X_HIP(hipStreamSynchronize(d->stream));
    return T2X_OK;
}

// Diagnostic copy-out of internal buffers (tests only): 0 coarse slices (re/im), 1 coarse |F|^2,
// 2 N1, 3 image-side A, 4 A1, 5 coarse diff2. Offsets and counts in floats.
extern "C" int t2x_debug_copy(t2x_device* h, int which, size_t offset, size_t count, float* dst)
{
    if (!h || !dst) { set_error("null argument"); return T2X_EINVAL; }
    Device* d = reinterpret_cast<Device*>(h);
    DevBuf* bufs[] = {&d->coarse_B, &d->coarse_B2, &d->N1, &d->A, &d->A1, &d->diff2};
    if (which < 0 || which > 5) { set_error("bad buffer id"); return T2X_EINVAL; }
    DevBuf* b = bufs[which];
    if ((offset + count) * sizeof(float) > b->bytes) { set_error("range outside buffer"); return T2X_EINVAL; }
    T2X_HIP(hipStreamSynchronize(d->stream));
    T2X_HIP(hipMemcpy(dst, b->as<float>() + offset, count * sizeof(float), hipMemcpyDeviceToHost));
    return T2X_OK;
}
