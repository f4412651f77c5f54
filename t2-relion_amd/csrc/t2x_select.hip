// t2x_select.hip — FP64 weights and significance selection of one E-step pass.
//
// Restates convertAllSquaredDifferencesToWeights (ml_optimiser.cpp:7510-8041):
//   w = pdf_orientation/mean * pdf_offset/mean * exp(-(diff2 - min_diff2)), 0 if diff2-min > 700;
//   the significant weight is found by accumulating the positive weights in descending order
//   until the running sum exceeds adaptive_fraction * sum (ml_optimiser.cpp:7947-7967).
// Instead of sorting every weight (the CUB radix sort of acc_ml_optimiser_impl.h:1430-1469),
// one workgroup per particle runs an exact radix *select* on the IEEE-754 bit pattern of the
// positive FP64 weights: per-digit counts and weight sums locate the bin where the descending
// cumulative sum crosses the target; the final bin (<= SEL_CAP elements) is sorted in LDS and
// walked sequentially, so the threshold value and count equal those of the sorted walk.
#include "t2x_kernels.h"

namespace t2x {

#define SEL_THREADS 1024
#define SEL_BINS 4096
#define SEL_CAP 1024

__device__ __forceinline__ double pass_weight(const WeightSrc& s, int b, long long e, double mind2,
                                              const long long* pair_off_b_cache)
{
    float d;
    double prior;
    if (s.kind == SRC_COARSE) {
        long long t = e / s.O_all, n = e - t * s.O_all;
        if (!coarse_allowed(s, b, n)) return 0.;
        // blocks beyond the negligible-weight cut (their values may not have been stored)
        if ((double)s.rowmin[((size_t)b * s.T + t) * s.rowmin_ld + n / 64] - mind2 > 60.0) return 0.;
        prior = orient_prior(s.prior_o, s.lp, b, n) * trans_prior(s.prior_t, s.T, s.tK, s.tcols, b, n, (int)t);
        if (!(prior > 0.)) return 0.;
        d = s.d2[((size_t)b * s.T + t) * s.ld + n];
    } else {
        long long base = pair_off_b_cache[0];
        long long k = base + e / s.nover;
        int2 nt = s.pairs[k];
        prior = orient_prior(s.prior_o, s.lp, b, nt.x) * trans_prior(s.prior_t, s.T, s.tK, s.tcols, b, nt.x, nt.y);
        if (!(prior > 0.)) return 0.;
        d = s.d2[base * s.nover + e];
    }
    double dd = (double)d - mind2;
    if (dd > 700.) return 0.;
    return prior * exp(-dd);
}

__device__ __forceinline__ long long pass_count(const WeightSrc& s, int b)
{
    if (s.kind == SRC_COARSE) return (long long)s.T * s.O_all;
    return (s.pair_off[b + 1] - s.pair_off[b]) * (long long)s.nover;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

// digit layout of a positive double: 12+12+12+12+12+4 bits from the top
__device__ __forceinline__ int digit_shift(int L) { return L < 5 ? 52 - 12 * L : 0; }
__device__ __forceinline__ int digit_width(int L) { return L < 5 ? 12 : 4; }

// Minimum diff2 of the pass: coarse from the GEMM row minima, fine from the atomic minimum.
__global__ void k_pass_min(WeightSrc s, int B, SelResult* res)
{
    int b = blockIdx.x;
    if (b >= B) return;
    __shared__ float red[SEL_THREADS / 64];
    float m = 3.0e38f;
    if (s.kind == SRC_COARSE) {
        long long nr = (long long)s.T * s.rowmin_ld;
        const float* rm = s.rowmin + (size_t)b * s.T * s.rowmin_ld;
        for (long long i = threadIdx.x; i < nr; i += blockDim.x) m = fminf(m, rm[i]);
    } else if (threadIdx.x == 0) {
        m = ordered_to_float(s.fine_min[b]);
    }
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float mm = red[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) mm = fminf(mm, red[i]);
        res[b].min_d2 = (double)mm;
    }
}

struct SelShared {
    unsigned int cnt[SEL_BINS];
    double sum[SEL_BINS];
    double buf[SEL_CAP];
    double red[SEL_THREADS / 64];
    unsigned long long redu[SEL_THREADS / 64];
    double acc;            // weight sum of elements in digits above the chosen bin
    long long cnt_above;   // number of elements above the chosen bin
    unsigned long long prefix;
    int digit;
    int found;
    int mode;              // 0: cumulative-sum crossing, 1: rank
    long long rank;        // target rank (1-based, descending) in rank mode
    long long bin_count;
};

// Chooses the digit of level L from the LDS histogram (one wave; the highest digit first).
__device__ void choose_digit(SelShared& S, double target, int nbins)
{
    // 64 lanes x (nbins/64) bins; lane 0 owns the highest bins
    const int lane = threadIdx.x;  // called by wave 0 only
    const int per = nbins / 64;
    const int hi = nbins - 1 - lane * per;   // lane's highest bin
    double s = 0.;
    unsigned long long c = 0;
    for (int i = 0; i < per; i++) {
        s += S.sum[hi - i];
        c += S.cnt[hi - i];
    }
    // exclusive prefix over lanes (descending bins)
    double ps = s;
    unsigned long long pc = c;
    for (int o = 1; o < 64; o <<= 1) {
        double ts = __shfl_up(ps, o);
        unsigned long long tc = __shfl_up(pc, o);
        if (lane >= o) { ps += ts; pc += tc; }
    }
    double excl_s = ps - s;
    unsigned long long excl_c = pc - c;
    bool hit;
    if (S.mode == 0) hit = (S.acc + excl_s <= target) && (S.acc + ps > target);
    else hit = (S.cnt_above + (long long)excl_c < S.rank) && (S.cnt_above + (long long)pc >= S.rank);
    unsigned long long ball = __ballot(hit);
    int src = ball ? (__ffsll((long long)ball) - 1) : -1;
    if (src < 0) {
        // no crossing inside this prefix bin (summation-order rounding): take the lowest
        // non-empty digit, i.e. the last element of the bin in descending order
        unsigned long long nz = __ballot(c > 0);
        int last = nz ? (63 - __clzll((long long)nz)) : 0;
        src = last;
        if (lane == src) {
            int dsel = -1;
            double a = S.acc + excl_s;
            long long ca = S.cnt_above + (long long)excl_c;
            for (int i = 0; i < per; i++) {
                int d = hi - i;
                if (S.cnt[d] > 0) dsel = d;
            }
            // accumulate everything above dsel
            for (int i = 0; i < per; i++) {
                int d = hi - i;
                if (d == dsel) break;
                a += S.sum[d];
                ca += S.cnt[d];
            }
            S.acc = a;
            S.cnt_above = ca;
            S.digit = dsel;
            S.bin_count = S.cnt[dsel];
            S.found = 0;
        }
        return;
    }
    if (lane == src) {
        double a = S.acc + excl_s;
        long long ca = S.cnt_above + (long long)excl_c;
        int dsel = hi - per + 1;
        for (int i = 0; i < per; i++) {
            int d = hi - i;
            bool h = (S.mode == 0) ? (a + S.sum[d] > target) : (ca + (long long)S.cnt[d] >= S.rank);
            if (h) { dsel = d; break; }
            a += S.sum[d];
            ca += S.cnt[d];
        }
        S.acc = a;
        S.cnt_above = ca;
        S.digit = dsel;
        S.bin_count = S.cnt[dsel];
        S.found = 1;
    }
}

// Bitonic sort (descending) of n <= SEL_CAP doubles in LDS; pads with -1.
__device__ void sort_desc(double* a, int n)
{
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = n + threadIdx.x; i < m; i += blockDim.x) a[i] = -1.0;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    bool desc = ((i & k) == 0);
                    double x = a[i], y = a[ixj];
                    if (desc ? (x < y) : (x > y)) { a[i] = y; a[ixj] = x; }
                }
            }
            __syncthreads();
        }
}

// One selection run (crossing or rank mode) over the positive weights of particle b.
// Returns the threshold in S.buf[0] slot via out params (thread 0).
__device__ void select_run(const WeightSrc& s, int b, double mind2, long long n_elem,
                           const long long* poff, SelShared& S, double target, int mode,
                           long long rank, bool have_level0, double* out_w, long long* out_n)
{
    if (threadIdx.x == 0) {
        S.acc = 0.;
        S.cnt_above = 0;
        S.prefix = 0;
        S.mode = mode;
        S.rank = rank;
    }
    __syncthreads();
    int L = 0;
    for (;; L++) {
        const int shift = digit_shift(L), width = digit_width(L);
        const int nb = 1 << width;
        if (!(L == 0 && have_level0)) {
            for (int i = threadIdx.x; i < SEL_BINS; i += blockDim.x) { S.cnt[i] = 0; S.sum[i] = 0.; }
            __syncthreads();
            const unsigned long long pref = S.prefix;
            const int pshift = shift + width;
            for (long long e = threadIdx.x; e < n_elem; e += blockDim.x) {
                double w = pass_weight(s, b, e, mind2, poff);
                if (!(w > 0.)) continue;
                unsigned long long bits = (unsigned long long)__double_as_longlong(w);
                if (L > 0 && (bits >> pshift) != pref) continue;
                int d = (int)((bits >> shift) & (unsigned long long)(nb - 1));
                atomicAdd(&S.cnt[d], 1u);
                atomicAdd(&S.sum[d], w);
            }
            __syncthreads();
        }
        if (threadIdx.x < 64) choose_digit(S, target, nb < 64 ? 64 : nb);
        __syncthreads();
        if (threadIdx.x == 0) S.prefix = (S.prefix << width) | (unsigned long long)S.digit;
        __syncthreads();
        if (S.bin_count <= SEL_CAP || L == 5) break;
    }
    // gather the chosen bin
    const int shift = digit_shift(L);
    const unsigned long long pref = S.prefix;
    const long long bc = S.bin_count;
    if (bc <= SEL_CAP) {
        __shared__ unsigned int gcount;
        if (threadIdx.x == 0) gcount = 0;
        __syncthreads();
        for (long long e = threadIdx.x; e < n_elem; e += blockDim.x) {
            double w = pass_weight(s, b, e, mind2, poff);
            if (!(w > 0.)) continue;
            unsigned long long bits = (unsigned long long)__double_as_longlong(w);
            if ((bits >> shift) != pref) continue;
            unsigned int slot = atomicAdd(&gcount, 1u);
            if (slot < SEL_CAP) S.buf[slot] = w;
        }
        __syncthreads();
        int n = (int)min((unsigned int)SEL_CAP, gcount);
        sort_desc(S.buf, n);
        if (threadIdx.x == 0) {
            if (mode == 0) {
                double frac = S.acc;
                int i = 0;
                for (; i < n; i++) {
                    frac += S.buf[i];
                    if (frac > target) break;
                }
                if (i >= n) i = n - 1;
                *out_w = S.buf[i];
                *out_n = S.cnt_above + i + 1;
            } else {
                long long r = S.rank - S.cnt_above - 1;
                if (r < 0) r = 0;
                if (r >= n) r = n - 1;
                *out_w = S.buf[r];
                *out_n = S.rank;
            }
        }
    } else if (threadIdx.x == 0) {
        // every element of the bin has the same value (all 64 bits fixed)
        double v = __longlong_as_double((long long)pref);
        if (mode == 0) {
            double frac = S.acc;
            long long i = 0;
            for (; i < bc; i++) {
                frac += v;
                if (frac > target) break;
            }
            if (i >= bc) i = bc - 1;
            *out_w = v;
            *out_n = S.cnt_above + i + 1;
        } else {
            *out_w = v;
            *out_n = S.rank;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(SEL_THREADS) void k_select(WeightSrc s, SelParams prm, int B, SelResult* res,
                                                         const int* only)
{
    const int b = blockIdx.x;
    if (b >= B || (only && !only[b])) return;
    __shared__ SelShared S;
    __shared__ long long poff[2];
    if (threadIdx.x == 0 && s.kind == SRC_FINE) { poff[0] = s.pair_off[b]; poff[1] = s.pair_off[b + 1]; }
    __syncthreads();
    const double mind2 = res[b].min_d2;
    const long long n_elem = pass_count(s, b);
    // level-0 histogram together with the total sum and the positive count
    for (int i = threadIdx.x; i < SEL_BINS; i += blockDim.x) { S.cnt[i] = 0; S.sum[i] = 0.; }
    __syncthreads();
    double lsum = 0.;
    unsigned long long lpos = 0;
    double lmax = 0.;
    for (long long e = threadIdx.x; e < n_elem; e += blockDim.x) {
        double w = pass_weight(s, b, e, mind2, poff);
        if (!(w > 0.)) continue;
        lsum += w;
        lpos++;
        lmax = fmax(lmax, w);
        unsigned long long bits = (unsigned long long)__double_as_longlong(w);
        int d = (int)(bits >> 52);
        atomicAdd(&S.cnt[d], 1u);
        atomicAdd(&S.sum[d], w);
    }
    lsum = wave_sum(lsum);
    lpos = wave_sum_u64(lpos);
    for (int o = 32; o > 0; o >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, o));
    __shared__ double smax[SEL_THREADS / 64];
    if ((threadIdx.x & 63) == 0) {
        S.red[threadIdx.x >> 6] = lsum;
        S.redu[threadIdx.x >> 6] = lpos;
        smax[threadIdx.x >> 6] = lmax;
    }
    __syncthreads();
    __shared__ double total;
    __shared__ long long npos;
    __shared__ double gmax;
    if (threadIdx.x == 0) {
        double t = 0.;
        unsigned long long c = 0;
        double m = 0.;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) { t += S.red[i]; c += S.redu[i]; m = fmax(m, smax[i]); }
        total = t;
        npos = (long long)c;
        gmax = m;
    }
    __syncthreads();
    if (npos == 0) {
        if (threadIdx.x == 0) {
            res[b].sum_w = total;
            res[b].sig_w = 0.;
            res[b].nsig = 0;
            res[b].npos = 0;
            res[b].max_w = 0.;
        }
        return;
    }
    const double target = prm.fraction * total;
    __shared__ double thr_w;
    __shared__ long long thr_n;
    // crossing mode; when the total never exceeds the target the walk ends at the smallest weight
    bool crossing_possible = total > target;
    if (crossing_possible)
        select_run(s, b, mind2, n_elem, poff, S, target, 0, 0, true, &thr_w, &thr_n);
    else
        select_run(s, b, mind2, n_elem, poff, S, target, 1, npos, true, &thr_w, &thr_n);
    __syncthreads();
    if (prm.count_sig && prm.max_sig > 0 && thr_n > prm.max_sig) {
        select_run(s, b, mind2, n_elem, poff, S, target, 1, prm.max_sig, false, &thr_w, &thr_n);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        res[b].sum_w = total;
        res[b].sig_w = thr_w;
        res[b].nsig = thr_n;
        res[b].npos = npos;
        res[b].max_w = gmax;
    }
}

// Counts, per particle, the coarse (orientation, translation) pairs with w >= sig_w and the
// orientations having at least one such pair (Mcoarse_significant, ml_optimiser.cpp:7996-8004;
// isSignificantAnyImageAnyTranslation, :6898-6919).
__global__ __launch_bounds__(1024) void k_sig_count(WeightSrc s, int B, const SelResult* res,
                                                    long long* counts /* [B][2] */, const int* only)
{
    const int b = blockIdx.x;
    if (b >= B || (only && !only[b])) return;
    const double mind2 = res[b].min_d2, sig = res[b].sig_w;
    // dense coarse layout (columns = orientations), or the pair list of a local search's coarse
    // pass (SRC_FINE with one sample per pair: every orientation followed by its T translations)
    const bool plist = s.kind == SRC_FINE;
    const long long* pob = plist ? s.pair_off + b : nullptr;
    const int ncol = plist ? (int)((s.pair_off[b + 1] - s.pair_off[b]) / s.T) : s.O_all;
    unsigned long long np = 0, no = 0;
    for (int n = threadIdx.x; n < ncol; n += blockDim.x) {
        int c = 0;
        for (int t = 0; t < s.T; t++) {
            const long long e = plist ? (long long)n * s.T + t : (long long)t * s.O_all + n;
            double w = pass_weight(s, b, e, mind2, pob);
            if (w >= sig && w > 0.) c++;
        }
        np += c;
        no += (c > 0);
    }
    np = wave_sum_u64(np);
    no = wave_sum_u64(no);
    __shared__ unsigned long long r1[16], r2[16];
    if ((threadIdx.x & 63) == 0) { r1[threadIdx.x >> 6] = np; r2[threadIdx.x >> 6] = no; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, o = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) { a += r1[i]; o += r2[i]; }
        counts[2 * b] = (long long)a;
        counts[2 * b + 1] = (long long)o;
    }
}

// Writes the ordered (orientation-major, then translation) significant pairs and orientations.
// A particle given no room by the batch plan (an error, or a plan beyond the lists' capacity)
// writes nothing.
__global__ __launch_bounds__(1024) void k_sig_write(WeightSrc s, int B, const SelResult* res,
                                                    const long long* pair_off, const long long* orient_off,
                                                    int2* pairs, int* orients, long long* orient_pair_start,
                                                    const int* only)
{
    const int b = blockIdx.x;
    if (b >= B || (only && !only[b]) || pair_off[b + 1] == pair_off[b]) return;
    const double mind2 = res[b].min_d2, sig = res[b].sig_w;
    const bool plist = s.kind == SRC_FINE;   // as in k_sig_count
    const long long* pob = plist ? s.pair_off + b : nullptr;
    const int ncol = plist ? (int)((s.pair_off[b + 1] - s.pair_off[b]) / s.T) : s.O_all;
    __shared__ int scan_p[1024], scan_o[1024];
    __shared__ long long run_p, run_o;
    if (threadIdx.x == 0) { run_p = pair_off[b]; run_o = orient_off[b]; }
    __syncthreads();
    for (int base = 0; base < ncol; base += blockDim.x) {
        const int col = base + threadIdx.x;
        unsigned long long mask[MAX_COARSE_T / 64] = {};   // bit t: translation t is significant
        int c = 0, n = col;
        if (col < ncol) {
            if (plist) n = s.pairs[s.pair_off[b] + (long long)col * s.T].x;
            for (int t = 0; t < s.T; t++) {
                const long long e = plist ? (long long)col * s.T + t : (long long)t * s.O_all + col;
                double w = pass_weight(s, b, e, mind2, pob);
                if (w >= sig && w > 0.) { mask[t >> 6] |= 1ull << (t & 63); c++; }
            }
        }
        scan_p[threadIdx.x] = c;
        scan_o[threadIdx.x] = c > 0;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            int vp = threadIdx.x >= o ? scan_p[threadIdx.x - o] : 0;
            int vo = threadIdx.x >= o ? scan_o[threadIdx.x - o] : 0;
            __syncthreads();
            scan_p[threadIdx.x] += vp;
            scan_o[threadIdx.x] += vo;
            __syncthreads();
        }
        long long p0 = run_p + scan_p[threadIdx.x] - c;
        long long o0 = run_o + scan_o[threadIdx.x] - (c > 0);
        if (c > 0) {
            orients[o0] = n;
            orient_pair_start[o0] = p0;
            int k = 0;
            for (int t = 0; t < s.T; t++)
                if (mask[t >> 6] & (1ull << (t & 63))) pairs[p0 + (k++)] = make_int2(n, t);
        }
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) { run_p = p0 + c; run_o = o0 + (c > 0); }
        __syncthreads();
    }
}

// Max weight of the fine pass and its CPU store-order key (storeWeightedSums walks
// class, direction, psi, oversampled rotation, translation, oversampled translation and keeps
// the first strict maximum, ml_optimiser.cpp:8643-8700); also counts weights >= sig_w.
__global__ __launch_bounds__(1024) void k_fine_argmax(WeightSrc s, int B, SelResult* res,
                                                      long long* nfine_sig)
{
    const int b = blockIdx.x;
    if (b >= B) return;
    __shared__ long long poff[2];
    if (threadIdx.x == 0) { poff[0] = s.pair_off[b]; poff[1] = s.pair_off[b + 1]; }
    __syncthreads();
    const double mind2 = res[b].min_d2, sig = res[b].sig_w;
    const long long n_elem = (poff[1] - poff[0]) * (long long)s.nover;
    double bw = -1.;
    long long bk = LLONG_MAX, be = -1;
    unsigned long long cnt = 0;
    const int ot = s.over_trans;
    for (long long e = threadIdx.x; e < n_elem; e += blockDim.x) {
        double w = pass_weight(s, b, e, mind2, poff);
        if (w >= sig && w > 0.) cnt++;
        int2 nt = s.pairs[poff[0] + e / s.nover];
        int q = (int)(e % s.nover);
        int ir = q / ot, iot = q % ot;
        long long key = (((long long)nt.x * (s.nover / ot) + ir) * s.T + nt.y) * ot + iot;
        if (w > bw || (w == bw && key < bk)) { bw = w; bk = key; be = e; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        double ow = __shfl_xor(bw, o);
        long long ok = __shfl_xor(bk, o), oe = __shfl_xor(be, o);
        if (ow > bw || (ow == bw && ok < bk)) { bw = ow; bk = ok; be = oe; }
    }
    cnt = wave_sum_u64(cnt);
    __shared__ double sw[16];
    __shared__ long long sk[16], se[16];
    __shared__ unsigned long long sc[16];
    if ((threadIdx.x & 63) == 0) { sw[threadIdx.x >> 6] = bw; sk[threadIdx.x >> 6] = bk; se[threadIdx.x >> 6] = be; sc[threadIdx.x >> 6] = cnt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double w = sw[0];
        long long k = sk[0], ee = se[0];
        unsigned long long c = sc[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
            if (sw[i] > w || (sw[i] == w && sk[i] < k)) { w = sw[i]; k = sk[i]; ee = se[i]; }
            c += sc[i];
        }
        res[b].max_w = w;
        res[b].best_key = k;
        res[b].best_e = ee;
        nfine_sig[b] = (long long)c;
    }
}

}  // namespace t2x

namespace t2x {

// ------------------------------------------------------------------------------------------
// Coarse selection by candidate compaction.
//
// With the positive weights sorted descending u_1 >= u_2 >= ... >= u_n, the walk of
// convertAllSquaredDifferencesToWeights stops at the first k with u_1+...+u_k > f*S. Then
// u_k+...+u_n >= (1-f)*S and so u_k >= (1-f)*S/n: every weight the walk can stop at, and
// every weight above it, is at least (1-f)*S/n. Pass 1 computes S, pass 2 keeps the weights
// above half that bound (margin for summation order), and one workgroup per particle sorts
// them in LDS and walks them exactly. A particle whose candidates overflow the buffer sends
// the batch to the radix-select path above.
// ------------------------------------------------------------------------------------------

#define CAND_CAP 4096

// Weights with diff2 - min > kTinyDD are below prior * e^-60 and are neither summed nor
// exponentiated: their total is at most e^-60 * (sum of all priors), below FP64 rounding of S
// unless the prior at the minimum is itself below ~1e-8 of the prior mass.
#define kTinyDD 60.0

struct SumPart {
    double sum, max;
    long long npos, nbound;
};

size_t sum_part_bytes() { return sizeof(SumPart); }

__global__ void k_log_prior(const double* __restrict__ p, long long n, double* __restrict__ lp)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        lp[i] = p[i] > 0. ? log(p[i]) : -INFINITY;
}

// One workgroup per (translation row, particle): partial sum, max and counts of the row. The
// GEMM epilogue left the minimum of every 64-column block of the row (scored columns only, +inf
// for blocks not computed); a block whose minimum is more than kTinyDD above the particle's
// minimum holds only negligible weights and is skipped without reading it. Each wave takes 64
// blocks at a time, one lane per block minimum, and walks the surviving blocks with one lane
// per column.
// The translation prior of row (b, t): its value, or with per-class priors the largest over
// the classes (a bound used to skip rows; every weight uses its own class's value).
__device__ __forceinline__ double row_prior_max(const WeightSrc& s, int b, int t)
{
    if (s.tK == 0) return s.prior_t[(size_t)b * s.T + t];
    double m = 0.;
    const double* r = s.prior_t + ((size_t)b * s.T + t) * s.tK;
    for (int c = 0; c < s.tK; c++) m = fmax(m, r[c]);
    return m;
}

// Per-class translation priors of row (b, t) staged in LDS when there are at most kRowPriorLds
// classes (every column of the row reads one of them).
constexpr int kRowPriorLds = 1024;
__device__ __forceinline__ bool stage_row_priors(const WeightSrc& s, int b, int t, double* sp)
{
    const bool use = s.tK > 0 && s.tK <= kRowPriorLds;
    if (use) {
        const double* r = s.prior_t + ((size_t)b * s.T + t) * s.tK;
        for (int c = threadIdx.x; c < s.tK; c += blockDim.x) sp[c] = r[c];
    }
    __syncthreads();
    return use;
}
__device__ __forceinline__ double col_prior(const WeightSrc& s, const double* sp, bool staged, int b, int n, int t,
                                            double pt_row)
{
    if (s.tK == 0) return pt_row;
    if (staged) return sp[(unsigned)n / (unsigned)s.tcols];
    return trans_prior(s.prior_t, s.T, s.tK, s.tcols, b, n, t);
}

__global__ __launch_bounds__(256) void k_coarse_sum_rows(WeightSrc s, int B, const SelResult* res, SumPart* part)
{
    const int t = blockIdx.x, b = blockIdx.y;
    if (b >= B || t >= s.T) return;
    const double mind2 = res[b].min_d2;
    const double pt = row_prior_max(s, b, t);
    __shared__ double s_pt[kRowPriorLds];
    const bool staged = stage_row_priors(s, b, t, s_pt);
    const size_t row_id = (size_t)b * s.T + t;
    const float* row = s.d2 + row_id * s.ld;
    const float* rmin = s.rowmin + row_id * s.rowmin_ld;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nblk = (s.O_all + 63) / 64;
    double sum = 0., mx = 0.;
    long long npos = 0;
    if (pt > 0.) {
        for (int j0 = wave * 64; j0 < nblk; j0 += nw * 64) {
            const int j = j0 + lane;
            const bool live = j < nblk && (double)rmin[j] - mind2 <= kTinyDD;
            unsigned long long m = __ballot(live);
            while (m) {
                const int jb = j0 + __ffsll((long long)m) - 1;
                m &= m - 1;
                const int n = jb * 64 + lane;
                if (n < s.O_all) {
                    const double po = s.prior_o[n];
                    const double dd = (double)row[n] - mind2;
                    if (po > 0. && dd <= kTinyDD && coarse_allowed(s, b, n)) {
                        const double ptn = col_prior(s, s_pt, staged, b, n, t, pt);
                        const double w = (po * ptn) * exp(-dd);
                        if (w > 0.) {
                            npos++;
                            sum += w;
                            mx = fmax(mx, w);
                        }
                    }
                }
            }
        }
    }
    sum = wave_sum(sum);
    npos = (long long)wave_sum_u64((unsigned long long)npos);
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    __shared__ SumPart red[4];
    if (lane == 0) red[wave] = SumPart{sum, mx, npos, npos};
    __syncthreads();
    if (threadIdx.x == 0) {
        SumPart r{0., 0., 0, 0};
        for (int i = 0; i < nw; i++) {
            r.sum += red[i].sum;
            r.max = fmax(r.max, red[i].max);
            r.npos += red[i].npos;
            r.nbound += red[i].nbound;
        }
        part[row_id] = r;
    }
}

// Fixed-order combination of the row partials (deterministic S). npos and nbound (kept in nsig
// until the walk overwrites it) count the positive non-negligible weights (diff2 - min <= kTinyDD);
// npos only serves the "no positive weight" check, for which the minimum itself always counts.
__global__ void k_coarse_sum_final(WeightSrc s, int B, const SumPart* part, SelResult* res)
{
    const int b = blockIdx.x;
    if (b >= B || threadIdx.x != 0) return;
    SumPart r{0., 0., 0, 0};
    for (int t = 0; t < s.T; t++) {
        const SumPart q = part[(size_t)b * s.T + t];
        r.sum += q.sum;
        r.max = fmax(r.max, q.max);
        r.npos += q.npos;
        r.nbound += q.nbound;
    }
    res[b].sum_w = r.sum;
    res[b].max_w = r.max;
    res[b].npos = r.npos;
    res[b].nsig = r.nbound;
}

// Candidates of the exact walk: weights >= 0.5 (1-f) S / nbound. Every weight the walk can stop
// at satisfies u_k * nbound >= (1-f) S - (negligible mass) (see above), so none is lost.
// One workgroup per (translation row, particle); slots are reserved per wave with one atomic.
__global__ __launch_bounds__(256) void k_coarse_candidates(WeightSrc s, int B, SelParams prm,
                                                           const SelResult* res, double* cand_w,
                                                           int* cand_e, int* cand_count)
{
    const int t = blockIdx.x, b = blockIdx.y;
    if (b >= B || t >= s.T) return;
    const double mind2 = res[b].min_d2;
    const long long nb = res[b].nsig;
    const double thr = (nb > 0) ? 0.5 * (1.0 - prm.fraction) * res[b].sum_w / (double)nb : 0.;
    const double pt = row_prior_max(s, b, t);
    __shared__ double s_pt[kRowPriorLds];
    const bool staged = stage_row_priors(s, b, t, s_pt);
    const size_t row_id = (size_t)b * s.T + t;
    const float* row = s.d2 + row_id * s.ld;
    const float* rmin = s.rowmin + row_id * s.rowmin_ld;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (!(pt > 0.) || !(thr > 0.)) return;
    // w >= thr  <=>  dd <= log(prior_o) + log(pt) - log(thr): compared in the log domain (relaxed by
    // 1e-9, far below the factor-2 margin of thr; with per-class translation priors pt is the
    // row's largest, so the filter only relaxes), exact test on the few that pass
    const double lbias = log(pt) - log(thr) + 1e-9;
    const int nblk = (s.O_all + 63) / 64;
    for (int j0 = wave * 64; j0 < nblk; j0 += nw * 64) {
        const int j = j0 + lane;
        const bool live = j < nblk && (double)rmin[j] - mind2 <= kTinyDD;
        unsigned long long mb = __ballot(live);
        while (mb) {
            const int jb = j0 + __ffsll((long long)mb) - 1;
            mb &= mb - 1;
            const int n = jb * 64 + lane;
            bool keep = false;
            double w = 0.;
            if (n < s.O_all) {
                const double dd = (double)row[n] - mind2;
                if (dd <= kTinyDD && dd <= s.log_prior_o[n] + lbias && coarse_allowed(s, b, n)) {
                    const double po = s.prior_o[n];
                    const double ptn = col_prior(s, s_pt, staged, b, n, t, pt);
                    w = (po * ptn) * exp(-dd);
                    keep = (po > 0.) && (w > 0.) && (w >= thr);
                }
            }
            const unsigned long long m = __ballot(keep);
            if (m) {
                const int leader = __ffsll((long long)m) - 1;
                int first = 0;
                if (lane == leader) first = atomicAdd(&cand_count[b], __popcll(m));
                first = __shfl(first, leader);
                if (keep) {
                    const int slot = first + __popcll(m & ((1ull << lane) - 1));
                    if (slot < CAND_CAP) {
                        cand_w[(size_t)b * CAND_CAP + slot] = w;
                        cand_e[(size_t)b * CAND_CAP + slot] = t * s.O_all + n;
                    }
                }
            }
        }
    }
}

// Bitonic sort in LDS of n <= CAND_CAP (key, payload) pairs; desc on doubles or asc on ints.
__device__ void sort_pairs_desc(double* w, int* e, int n)
{
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = n + threadIdx.x; i < m; i += blockDim.x) { w[i] = -1.0; e[i] = 0x7fffffff; }
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    bool desc = ((i & k) == 0);
                    double x = w[i], y = w[ixj];
                    int ex = e[i], ey = e[ixj];
                    // order: weight descending, then sample index ascending (deterministic ties)
                    bool x_after_y = (x < y) || (x == y && ex > ey);
                    bool sw = desc ? x_after_y : !x_after_y && !(x == y && ex == ey);
                    if (sw) { w[i] = y; w[ixj] = x; e[i] = ey; e[ixj] = ex; }
                }
            }
            __syncthreads();
        }
}

__device__ void sort_int_asc(int* a, int n)
{
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = n + threadIdx.x; i < m; i += blockDim.x) a[i] = 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    bool asc = ((i & k) == 0);
                    int x = a[i], y = a[ixj];
                    if (asc ? (x > y) : (x < y)) { a[i] = y; a[ixj] = x; }
                }
            }
            __syncthreads();
        }
}

// Exact walk over the sorted candidates; the significant pairs, ordered by
// (orientation, translation), go to a per-particle strided list.
// force (tests): 1 sends every particle, 2 every odd-numbered one, to the radix select.
__global__ __launch_bounds__(1024) void k_cand_select(WeightSrc s, int B, SelParams prm,
                                                      SelResult* res, const double* cand_w,
                                                      const int* cand_e, const int* cand_count,
                                                      int2* pairs_strided, long long* counts,
                                                      int* overflow, int force)
{
    const int b = blockIdx.x;
    if (b >= B) return;
    __shared__ double w[CAND_CAP];
    __shared__ int e[CAND_CAP];
    __shared__ int nsig_pairs;
    const int n = cand_count[b];
    if (n > CAND_CAP || n <= 0 || force == 1 || (force == 2 && (b & 1))) {
        if (threadIdx.x == 0) overflow[b] = 1;
        return;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        w[i] = cand_w[(size_t)b * CAND_CAP + i];
        e[i] = cand_e[(size_t)b * CAND_CAP + i];
    }
    __syncthreads();
    sort_pairs_desc(w, e, n);
    if (threadIdx.x == 0) {
        const double target = prm.fraction * res[b].sum_w;
        double frac = 0., sig = 0.;
        long long cnt = 0;
        int i = 0;
        bool crossed = false;
        for (; i < n; i++) {
            if (prm.max_sig > 0) {
                if (cnt < prm.max_sig) {
                    if (prm.count_sig) cnt++;
                    sig = w[i];
                }
            } else {
                if (prm.count_sig) cnt++;
                sig = w[i];
            }
            frac += w[i];
            if (frac > target) { crossed = true; break; }
        }
        if (!crossed) overflow[b] = 2;   // the walk must end inside the candidates
        res[b].sig_w = sig;
        res[b].nsig = cnt;
        int m = 0;
        while (m < n && w[m] >= sig) m++;
        nsig_pairs = m;
    }
    __syncthreads();
    const int m = nsig_pairs;
    // order the significant prefix by (orientation, translation): key = n*T + t
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        int ee = e[i];
        int t = ee / s.O_all, nn = ee - t * s.O_all;
        e[i] = nn * s.T + t;
    }
    __syncthreads();
    sort_int_asc(e, m);
    for (int i = threadIdx.x; i < m; i += blockDim.x)
        pairs_strided[(size_t)b * CAND_CAP + i] = make_int2(e[i] / s.T, e[i] % s.T);
    if (threadIdx.x == 0) {
        int no = 0;
        for (int i = 0; i < m; i++)
            if (i == 0 || e[i] / s.T != e[i - 1] / s.T) no++;
        counts[2 * b] = m;
        counts[2 * b + 1] = no;
    }
}

// Strided per-particle lists -> compact arrays at the batch plan's offsets (particles sent to
// the radix select, flagged in skip, are written by k_sig_write instead).
__global__ void k_cand_scatter(int B, const int2* pairs_strided, const long long* pair_off,
                               const long long* orient_off, int2* pairs, int* orients,
                               long long* orient_pair_start, const int* skip)
{
    const int b = blockIdx.x;
    if (b >= B || (skip && skip[b])) return;
    const long long p0 = pair_off[b], m = pair_off[b + 1] - p0;
    for (long long i = threadIdx.x; i < m; i += blockDim.x)
        pairs[p0 + i] = pairs_strided[(size_t)b * CAND_CAP + i];
    if (threadIdx.x == 0) {
        long long o = orient_off[b];
        for (long long i = 0; i < m; i++) {
            int2 pr = pairs_strided[(size_t)b * CAND_CAP + i];
            if (i == 0 || pr.x != pairs_strided[(size_t)b * CAND_CAP + i - 1].x) {
                orients[o] = pr.x;
                orient_pair_start[o] = p0 + i;
                o++;
            }
        }
    }
}


// The batch's work lists, sized on the device (no host round trip between the coarse and the
// fine pass). Per particle: its status from the coarse selection (no positive weight or a zero /
// NaN sum: T2X_EZEROSUM; no significant sample: T2X_ENOSIG; such a particle gets no work), and
// exclusive prefix sums of its significant pairs, orientations, fine items (orientations x
// oversampled rotations) and fine tiles (ceil(items / rows_per_tile) x column blocks). When the
// pairs or orientations exceed the lists' capacity every offset is written as zero (no kernel
// after this one does any work) and the true totals are kept, so the host can grow the lists and
// plan again. Clears the work-queue counters of the persistent kernels. One workgroup of 1024
// threads; B <= 4 * 1024.
__global__ __launch_bounds__(1024) void k_batch_plan(const long long* __restrict__ counts, const SelResult* __restrict__ selc,
                                                     int B, int over_rot, int rows_per_tile, int ncb,
                                                     long long cap_pairs, long long cap_orients,
                                                     long long* pair_off, long long* orient_off,
                                                     long long* item_off, long long* tile_off, BatchPlan* plan)
{
    __shared__ long long sc[4][1024];
    __shared__ int s_err_b, s_err_code;
    const int tid = threadIdx.x;
    const int per = (B + 1023) / 1024;
    const int i0 = min(B, tid * per), i1 = min(B, i0 + per);
    if (tid == 0) { s_err_b = 0x7fffffff; s_err_code = 0; }
    __syncthreads();
    long long loc[4] = {0, 0, 0, 0};
    for (int i = i0; i < i1; i++) {
        const SelResult& r = selc[i];
        int code = 0;
        if (r.npos == 0 || !(r.sum_w > 0.) || isnan(r.sum_w)) code = T2X_EZEROSUM;
        else if (r.nsig == 0) code = T2X_ENOSIG;
        if (code) {
            atomicMin(&s_err_b, i);
            continue;
        }
        const long long np = counts[2 * i], no = counts[2 * i + 1];
        loc[0] += np;
        loc[1] += no;
        loc[2] += no * over_rot;
        loc[3] += (no * over_rot + rows_per_tile - 1) / rows_per_tile * ncb;
    }
    for (int q = 0; q < 4; q++) sc[q][tid] = loc[q];
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        long long v[4];
        for (int q = 0; q < 4; q++) v[q] = tid >= o ? sc[q][tid - o] : 0;
        __syncthreads();
        for (int q = 0; q < 4; q++) sc[q][tid] += v[q];
        __syncthreads();
    }
    const long long tot_p = sc[0][1023], tot_o = sc[1][1023];
    const bool exceeded = tot_p > cap_pairs || tot_o > cap_orients;
    long long run[4];
    for (int q = 0; q < 4; q++) run[q] = sc[q][tid] - loc[q];
    for (int i = i0; i < i1; i++) {
        pair_off[i] = exceeded ? 0 : run[0];
        orient_off[i] = exceeded ? 0 : run[1];
        item_off[i] = exceeded ? 0 : run[2];
        tile_off[i] = exceeded ? 0 : run[3];
        const SelResult& r = selc[i];
        const bool bad = r.npos == 0 || !(r.sum_w > 0.) || isnan(r.sum_w) || r.nsig == 0;
        if (!bad) {
            const long long np = counts[2 * i], no = counts[2 * i + 1];
            run[0] += np;
            run[1] += no;
            run[2] += no * over_rot;
            run[3] += (no * over_rot + rows_per_tile - 1) / rows_per_tile * ncb;
        }
    }
    if (tid == 1023) {
        pair_off[B] = exceeded ? 0 : sc[0][1023];
        orient_off[B] = exceeded ? 0 : sc[1][1023];
        item_off[B] = exceeded ? 0 : sc[2][1023];
        tile_off[B] = exceeded ? 0 : sc[3][1023];
        plan->npairs = tot_p;
        plan->norients = tot_o;
        plan->nitems = sc[2][1023];
        plan->ntiles = sc[3][1023];
        plan->exceeded = exceeded ? 1 : 0;
        plan->queue[0] = 0;
        plan->queue[1] = 0;
    }
    __syncthreads();
    if (tid == 0) {
        plan->err_particle = s_err_b == 0x7fffffff ? -1 : s_err_b;
        if (s_err_b != 0x7fffffff) {
            const SelResult& r = selc[s_err_b];
            plan->err_code = (r.npos == 0 || !(r.sum_w > 0.) || isnan(r.sum_w)) ? T2X_EZEROSUM : T2X_ENOSIG;
        } else {
            plan->err_code = 0;
        }
    }
}

}  // namespace t2x
