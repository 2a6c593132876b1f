// sx_corr.hip -- time autocorrelation on gfx950: batched Stockham FFTs in LDS.
//
// Reference: smath::auto_correlate_fftw (smath.cpp:141-156) pads each series
// to 2NF, transforms, takes |X|^2, transforms back and scales by
// 1/(2NF (NF - tau)); store() (all_vectors_scatter_device.cpp:231-236) then
// sums C_m(tau) over vectors and accumulates a_m = mean_tau C_m(tau) and
// |a_m|^2.  Because the back transform is linear, this build sums the power
// spectra of all vectors (and, for "self", all atoms) of a shell and
// transforms back once per shell in FP64; a_m is the dot product of |X_m|^2
// with a fixed FP64 weight vector w_f = (1/NF) sum_tau e^{+2 pi i f tau/L} /
// (L (NF - tau)).  Any length L >= 2NF - 1 gives the same linear correlation.
//
// Each series x (NF points, zero padded to L = 2*Lh with Lh >= NF) is
// transformed as two Lh-point FFTs, because the upper half of the padded
// input is zero:  X_{2f} = FFT_Lh(x)_f  and  X_{2f+1} = FFT_Lh(x_n W_L^n)_f.
// The halves run through the same Stockham passes (radix 8/4/2, one odd
// radix 3 or 5 last) in LDS, one pass per barrier pair.  The coherent path
// transforms in FP64 (its FFTs are a negligible share of the work, and FP32
// rounding in the back-transformed tail, where C(tau) averages only NF - tau
// products, would otherwise reach 1e-5 of C(0)); the incoherent path, whose
// cost is dominated by NA x NM FFTs, transforms in FP32, where the rounding
// averages out over the pairs.
#include <cmath>
#include <type_traits>

#include "sx_fft_device.h"

namespace sx {

// ------------------------------------------------------------ planning ----

namespace {
struct Cand {
    int Lh, T, E, odd;
};
// One-block half-lengths: 2^k, 3 2^k, 5 2^k up to 10240 (160 KB of FP64 in LDS).
const Cand kCands[] = {
    {256, 64, 4, 1},    {512, 64, 8, 1},     {1024, 128, 8, 1},  {1536, 64, 24, 3},
    {2048, 256, 8, 1},  {2560, 128, 20, 5},  {3072, 128, 24, 3}, {4096, 256, 16, 1},
    {5120, 256, 20, 5}, {6144, 256, 24, 3},  {8192, 512, 16, 1}, {10240, 512, 20, 5},
};

bool plan_from(const Cand &c, int N1, FftPlan *plan) {
    FftPlan p;
    p.Lh = c.Lh;
    p.N1 = N1;
    p.L = 2 * N1 * c.Lh;
    p.T = c.T;
    p.E = c.E;
    int rem = c.Lh / c.odd;
    int np = 0;
    // powers of two: largest radix dividing both the remainder and E
    while (rem > 1) {
        int r = 8;
        while (r > 1 && (rem % r != 0 || c.E % r != 0)) r >>= 1;
        if (r == 1) return false;
        p.radix[np++] = r;
        rem /= r;
    }
    if (c.odd > 1) p.radix[np++] = c.odd;
    p.npass = np;
    int ns = 1;
    for (int i = 0; i < np; ++i) {
        p.ns[i] = ns;
        ns *= p.radix[i];
    }
    if (ns != c.Lh || np < 2) return false;
    int off = 0;
    for (int i = 1; i < np; ++i) {
        p.toff[i] = off;
        off += p.ns[i] * p.radix[i];
    }
    p.ntw = off;
    *plan = p;
    return true;
}
}  // namespace

bool make_fft_plan(size_t NF, FftPlan *plan) {
    if (NF == 0) return false;
    for (const Cand &c : kCands)
        if ((size_t)c.Lh >= NF) return plan_from(c, 1, plan);
    // Longer series: N1 rows of a one-block length Lh, N1 * Lh >= NF.  Each
    // row evaluates N1 inputs per point (the small DFT that forms it), so the
    // cost per half is about N1 Lh (4 N1 + 1.5 log2 Lh) operations: fewer rows
    // win over a slightly shorter length.
    const Cand *best = nullptr;
    int bestN1 = 0;
    double bestCost = 0.0;
    for (int N1 = 2; N1 <= kMaxRows; ++N1)
        for (const Cand &c : kCands) {
            const size_t len = (size_t)N1 * c.Lh;
            if (len < NF) continue;
            const double cost = (double)len * (4.0 * N1 + 1.5 * std::log2((double)c.Lh));
            if (!best || cost < bestCost) {
                best = &c;
                bestN1 = N1;
                bestCost = cost;
            }
        }
    return best && plan_from(*best, bestN1, plan);
}

template <typename C, int NH, int E>
__global__ __launch_bounds__(512) void corr_all_kernel(CorrArgs a) {
    extern __shared__ double2 lds_d[];
    corr_all_block<C, NH, E>(a, blockIdx.x / a.chunks, blockIdx.x % a.chunks, lds_d);
}

template <typename C, int NH, int E>
__global__ __launch_bounds__(512) void corr_self_kernel(CorrArgs a) {
    using Acc = typename Cx<C>::R;
    static_assert(sizeof(Acc) == 4, "self path runs FP32 FFTs");
    extern __shared__ double2 lds_d[];
    C *buf = reinterpret_cast<C *>(lds_d);
    const FftPlan &P = a.plan;
    double2 *red = reinterpret_cast<double2 *>(buf);  // reused after the last pass, as in corr_all_block
    const int tid = threadIdx.x;
    const int s = blockIdx.x / a.chunks, c = blockIdx.x % a.chunks;
    const long npairs = (long)a.NA * a.NM;  // NA counts the atoms listed in aidx
    const long p0 = npairs * c / a.chunks, p1 = npairs * (c + 1) / a.chunks;

    C v[NH][E];
    Acc acc[2][E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[0][e] = acc[1][e] = Acc(0);
    double sa_r = 0.0, sa_i = 0.0, sa2 = 0.0;
    double *row = a.ppart + ((size_t)s * a.chunks + c) * P.L;
    const size_t plane = a.NAp * a.NFp;
    bool first = true;
    int since = 0;
    for (long pr = p0; pr < p1; ++pr) {
        const int ja = (int)(pr / a.NM), m = (int)(pr % a.NM);
        const int j = a.aidx ? a.aidx[ja] : ja;
        const double *q = a.qrev + ((size_t)s * a.NM + m) * 3;
        SelfInput in;
        in.set(a.coords + (size_t)j * a.NFp, plane, a.NF);
        in.qx = q[0];
        in.qy = q[1];
        in.qz = q[2];
        in.b = a.b[(size_t)(a.brow_per_shell ? s : 0) * a.NAp + j];
        double ar = 0.0, ai = 0.0;
        int t = tid;
        asm volatile("" : "+v"(t));
        fft_power<C, NH, E>(P, v, acc, buf, a.twp, a.twL, a.w, a.NF, in, ar, ai, t);
        const double2 r = block_sum2(ar, ai, red, tid, P.T);
        if (tid == 0) {
            sa_r += r.x;
            sa_i += r.y;
            sa2 += r.x * r.x + r.y * r.y;
        }
        if (++since == kFlushEvery) {
            flush_power<E>(P, acc, row, first, tid);
            first = false;
            since = 0;
        }
    }
    if (since > 0 || first) flush_power<E>(P, acc, row, first, tid);
    if (tid == 0) {
        double *ap = a.apart + ((size_t)s * a.chunks + c) * 4;
        ap[0] = sa_r;
        ap[1] = sa_i;
        ap[2] = sa2;
        ap[3] = 0.0;
    }
}

// One half (H) of every pair of a chunk: the pair loop of corr_self_kernel
// with a single sweep per pair.
template <typename C, int E, int H, typename PT>
__device__ __forceinline__ void self_half_body(const CorrArgs &a, double2 *__restrict__ apair, int s, int c,
                                               C *buf) {
    using Acc = typename Cx<C>::R;
    const FftPlan &P = a.plan;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, nw = P.T >> 6;
    const long npairs = (long)a.NA * a.NM;
    const long p0 = npairs * c / a.chunks, p1 = npairs * (c + 1) / a.chunks;
    C v[1][E];
    Acc acc[2][E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[0][e] = acc[1][e] = Acc(0);
    double *row = a.ppart + ((size_t)s * a.chunks + c) * P.L;
    const size_t plane = a.NAp * a.NFp;
    bool first = true;
    int since = 0;
    using In = typename std::conditional<H == 0, SelfInput, SelfInputShifted>::type;
    using MT = typename std::conditional<std::is_void<PT>::value, int, MidTwiddles<C, PT>>::type;
    MT mt{};
    if constexpr (!std::is_void<PT>::value) mt.load(a.twp, (unsigned)tid);
    for (long pr = p0; pr < p1; ++pr) {
        const int ja = (int)(pr / a.NM), m = (int)(pr % a.NM);
        const int j = a.aidx ? a.aidx[ja] : ja;
        const double *q = a.qrev + ((size_t)s * a.NM + m) * 3;
        In in;
        in.set(a.coords + (size_t)j * a.NFp, plane, a.NF);
        in.qx = q[0];
        in.qy = q[1];
        in.qz = q[2];
        if constexpr (H == 1) in.invL = 1.0 / (double)P.L;
        in.b = a.b[(size_t)(a.brow_per_shell ? s : 0) * a.NAp + j];
        double ar = 0.0, ai = 0.0;
        int t = tid;
        asm volatile("" : "+v"(t));
        if constexpr (std::is_void<PT>::value)
            fft_sweep<C, 1, E>(P, H, v, acc, buf, a.twp, a.twL, a.w, a.NF, in, ar, ai, t);
        else
            fft_sweep_fixed<C, E, PT, H>(v[0], acc, buf, a.twp, mt, a.w, in, ar, ai, (unsigned)t);
        // per-wavefront part of a; the parts are summed per pair in fixed order
        // by self_a_reduce (no block barrier here)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            ar += __shfl_xor(ar, off, 64);
            ai += __shfl_xor(ai, off, 64);
        }
        if ((tid & 63) == 0)
            apair[(((size_t)s * npairs + pr) * 2 + H) * nw + wave] = make_double2(ar, ai);
        if (++since == kFlushEvery) {
            flush_power_half<E, H>(P, acc, row, first, tid);
            first = false;
            since = 0;
        }
    }
    if (since > 0 || first) flush_power_half<E, H>(P, acc, row, first, tid);
}

// Threads and workgroups per CU the split kernel is built for: 256 threads
// (E = 20, halves of 5,120 points: four workgroups per CU, as many as the
// LDS allows, so at most 128 VGPRs), or the 512 of a fixed plan (two
// workgroups per CU).
template <typename PT>
constexpr int split_threads() {
    if constexpr (std::is_void<PT>::value) return 256;
    else return PT::T > 256 ? (int)PT::T : 256;
}
template <typename PT, int E>
constexpr int split_min_blocks() {
    return split_threads<PT>() > 256 ? 2 : (E == 20 ? 4 : 1);
}

// PT: void for the runtime plan, else a StaticPlan the launch's plan matches.
template <typename C, int E, typename PT = void>
__global__ __launch_bounds__(split_threads<PT>(), (split_min_blocks<PT, E>())) void corr_self_split_kernel(
    CorrArgs a, double2 *apair) {
    extern __shared__ double2 lds_d[];
    C *buf = reinterpret_cast<C *>(lds_d);
    const int sc = blockIdx.x >> 1;
    const int s = sc / a.chunks, c = sc % a.chunks;
    if (blockIdx.x & 1)
        self_half_body<C, E, 1, PT>(a, apair, s, c, buf);
    else
        self_half_body<C, E, 0, PT>(a, apair, s, c, buf);
}

// Tail lags of the self path in FP64 (float products are exact in FP64).
// With A_t = x_t and B_u = x_{NF-K+u} (t, u < K, zero beyond K),
// S(NF-1-k) = sum_t conj(A_t) B_{t+d}, d = K-1-k: the sums along the
// diagonals d >= 0 of M = sum over pairs of conj(A) B^T.  Each block sums
// its pairs' M over the 136 tiles (4 x 4, tile row <= tile column) of the
// 64 x 64 upper triangle: threads 0-127 own tiles 0-127 for every staged
// pair; the last wavefront owns tiles 128-135 eight times over, its lane l
// taking tile 128 + l % 8 for staged pair l / 8 only, so no lane idles.
// Each thread does 16 complex products per pair from 8 LDS reads.  Pairs are
// staged eight at a time, one per wavefront at a time (the pair is
// wave-uniform, so its coordinate rows are scalar buffer resources).  The
// block then folds each thread's tile onto its seven diagonals and sums the
// diagonals in a fixed order.
constexpr int kTailThreads = 192;
constexpr int kTailBatch = 8;
constexpr int kTailPairSlots = 128;  // A then B of one staged pair

__device__ __forceinline__ int tail_tile_index(int I, int J) { return I * 16 - (I * (I - 1)) / 2 + (J - I); }

// The partial-sum slot (thread) of tile `tile` for pair class `i` (0 for tiles below 128).
__device__ __forceinline__ int tail_partial_slot(int tile, int i) { return tile < 128 ? tile : 128 + 8 * i + (tile - 128); }

__global__ __launch_bounds__(kTailThreads) void self_tail_kernel(CorrArgs a, double2 *__restrict__ tail,
                                                                 int tblocks) {
    static_assert(kTailLags == 64, "the tiling covers a 64 x 64 product");
    constexpr int kStage = kTailBatch * kTailPairSlots, kParts = kTailThreads * 7;
    __shared__ double2 sh[kStage > kParts ? kStage : kParts];  // staged pairs, then tile partials
    const int s = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const bool last = wave == 2;
    const int tile = last ? 128 + (lane & 7) : tid;
    const int p0 = last ? lane >> 3 : 0, pn = last ? 1 : kTailBatch;
    const int K = min(kTailLags, a.NF);
    const long npairs = (long)a.NA * a.NM;
    const size_t plane = a.NAp * a.NFp;
    int I = 0, rest = tile;
    while (rest >= 16 - I) {
        rest -= 16 - I;
        ++I;
    }
    const int J = I + rest;
    double mr[4][4], mi[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) mr[r][c] = mi[r][c] = 0.0;
    const long nbatch = (npairs + kTailBatch - 1) / kTailBatch;
    for (long bt = b; bt < nbatch; bt += tblocks) {
        __syncthreads();  // the previous batch is consumed
        for (int p = wave; p < kTailBatch; p += 3) {
            const long pr = bt * kTailBatch + p;
            float2 u0 = make_float2(0.f, 0.f), u1 = u0;
            if (pr < npairs) {
                const int ja = (int)(pr / a.NM), m = (int)(pr % a.NM);
                const int j = a.aidx ? a.aidx[ja] : ja;
                const double *q = a.qrev + ((size_t)s * a.NM + m) * 3;
                SelfInput in;
                in.set(a.coords + (size_t)j * a.NFp, plane, a.NF);
                in.qx = q[0];
                in.qy = q[1];
                in.qz = q[2];
                in.b = a.b[(size_t)(a.brow_per_shell ? s : 0) * a.NAp + j];
                if (lane < K) {
                    u0 = in((unsigned)lane);
                    u1 = in((unsigned)(a.NF - K + lane));
                }
            }
            sh[p * kTailPairSlots + lane] = make_double2(u0.x, u0.y);
            sh[p * kTailPairSlots + 64 + lane] = make_double2(u1.x, u1.y);
        }
        __syncthreads();
        for (int i = 0; i < pn; ++i) {
            const double2 *pp = sh + (p0 + i) * kTailPairSlots;
            double2 x[4], y[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                x[r] = pp[4 * I + r];
                y[r] = pp[64 + 4 * J + r];
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    mr[r][c] = fma(x[r].x, y[c].x, mr[r][c]);
                    mr[r][c] = fma(x[r].y, y[c].y, mr[r][c]);
                    mi[r][c] = fma(x[r].x, y[c].y, mi[r][c]);
                    mi[r][c] = fma(-x[r].y, y[c].x, mi[r][c]);
                }
        }
    }
    __syncthreads();
    // entry (r, c) of tile (I, J) lies on diagonal 4 (J - I) + c - r
#pragma unroll
    for (int e = 0; e < 7; ++e) {
        double pr_ = 0.0, pi_ = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = e - 3 + r;
            if (c >= 0 && c < 4) {
                pr_ += mr[r][c];
                pi_ += mi[r][c];
            }
        }
        sh[tid * 7 + e] = make_double2(pr_, pi_);
    }
    __syncthreads();
    if (tid < kTailLags) {
        const int k = tid, d = K - 1 - k;
        double tr = 0.0, ti = 0.0;
        if (d >= 0) {
            for (int D = 0; D < 16; ++D) {
                const int e = d - 4 * D + 3;
                if (e < (D == 0 ? 3 : 0) || e > 6) continue;  // tiles on the diagonal keep c >= r
                for (int I2 = 0; I2 + D < 16; ++I2) {
                    const int t = tail_tile_index(I2, I2 + D);
                    for (int i = 0; i < (t < 128 ? 1 : 8); ++i) {
                        const double2 v = sh[tail_partial_slot(t, i) * 7 + e];
                        tr += v.x;
                        ti += v.y;
                    }
                }
            }
        }
        tail[((size_t)s * tblocks + b) * kTailLags + k] = make_double2(tr, ti);
    }
}

// Per shell: sum over pairs of a (the sum of the pair's wpp parts: two halves
// x wavefronts) and |a|^2, in fixed order.
__global__ __launch_bounds__(256) void self_a_reduce_kernel(const double2 *__restrict__ apair, long npairs,
                                                            int wpp, int rblocks, double *__restrict__ apart) {
    __shared__ double2 red[4];
    __shared__ double red2[4];
    const int s = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
    double sr = 0.0, si = 0.0, s2 = 0.0;
    const double2 *base = apair + (size_t)s * npairs * wpp;
    for (long pr = (long)b * 256 + tid; pr < npairs; pr += (long)rblocks * 256) {
        double xr = 0.0, xi = 0.0;
        for (int k = 0; k < wpp; ++k) {
            const double2 ak = base[(size_t)pr * wpp + k];
            xr += ak.x;
            xi += ak.y;
        }
        sr += xr;
        si += xi;
        s2 += xr * xr + xi * xi;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
        s2 += __shfl_xor(s2, off, 64);
    }
    if ((tid & 63) == 0) {
        red[tid >> 6] = make_double2(sr, si);
        red2[tid >> 6] = s2;
    }
    __syncthreads();
    if (tid == 0) {
        double r0 = 0.0, r1 = 0.0, r2 = 0.0;
        for (int k = 0; k < 4; ++k) {
            r0 += red[k].x;
            r1 += red[k].y;
            r2 += red2[k];
        }
        double *ap = apart + ((size_t)s * rblocks + b) * 4;
        ap[0] = r0;
        ap[1] = r1;
        ap[2] = r2;
        ap[3] = 0.0;
    }
}

// Plans of N1 > 1 rows: block (shell s, chunk c of series, half h, row k1)
// transforms that row of every series of its chunk (one Lh-point FFT each),
// accumulates its bins' power and writes the series' part of a (block sum of
// the weighted power) to apart_rows[s][series][2 N1]; the parts of the 2 N1
// rows of a series are summed by self_a_reduce.  SELF: the series are the
// (atom, vector) pairs of the shell with phases formed here
// (self_vectors_scatter_device.cpp:306-319), else the coherent amplitudes of
// its vectors.
// PT: void for the runtime plan of a row, else a StaticPlan it matches (self
// series only), built for that plan's threads as the split kernel is.
template <typename PT>
constexpr int rows_threads() {
    if constexpr (std::is_void<PT>::value) return 512;
    else return split_threads<PT>();
}
template <typename PT, int E>
constexpr int rows_min_blocks() {
    if constexpr (std::is_void<PT>::value) return 1;
    else return split_min_blocks<PT, E>();
}

template <typename C, int E, bool SELF, typename PT = void>
__global__ __launch_bounds__(rows_threads<PT>(), (rows_min_blocks<PT, E>())) void corr_rows_kernel(
    CorrArgs a, double2 *__restrict__ apart_rows) {
    using Acc = typename Cx<C>::R;
    extern __shared__ double2 lds_d[];
    C *buf = reinterpret_cast<C *>(lds_d);
    double2 *red = reinterpret_cast<double2 *>(buf);  // after the last pass, as in corr_all_block
    const FftPlan &P = a.plan;
    const int rows = 2 * P.N1;
    int bid = blockIdx.x;
    const int hk = bid % rows;
    bid /= rows;
    const int c = bid % a.chunks, s = bid / a.chunks;
    const int h = hk / P.N1, k1 = hk % P.N1;
    const int tid = threadIdx.x;
    const long npairs = SELF ? (long)a.NA * a.NM : (long)a.NM;
    const long p0 = npairs * c / a.chunks, p1 = npairs * (c + 1) / a.chunks;
    const C *tw = pre_twiddles<C>(a);
    const C *twp = pass_twiddles<C>(a);
    const double2 *w = a.w + (size_t)hk * P.Lh;
    C v[1][E];
    Acc acc[2][E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[0][e] = acc[1][e] = Acc(0);
    double *row = a.ppart + ((size_t)s * a.chunks + c) * P.L;
    const size_t plane = a.NAp * a.NFp;
    bool first = true;
    int since = 0;
    using MT = typename std::conditional<std::is_void<PT>::value, int, MidTwiddles<C, PT>>::type;
    MT mt{};
    if constexpr (!std::is_void<PT>::value) mt.load(twp, (unsigned)tid);
    for (long pr = p0; pr < p1; ++pr) {
        double ar = 0.0, ai = 0.0;
        int t = tid;
        asm volatile("" : "+v"(t));
        if constexpr (SELF) {
            const int ja = (int)(pr / a.NM), m = (int)(pr % a.NM);
            const int j = a.aidx ? a.aidx[ja] : ja;
            const double *q = a.qrev + ((size_t)s * a.NM + m) * 3;
            RowInput<C, SelfRowSource> in;
            in.src.X = a.coords + (size_t)j * a.NFp;
            in.src.Y = in.src.X + plane;
            in.src.Z = in.src.X + 2 * plane;
            in.src.qx = q[0];
            in.src.qy = q[1];
            in.src.qz = q[2];
            in.src.hinvL = (double)h / (double)P.L;
            in.src.b = a.b[(size_t)(a.brow_per_shell ? s : 0) * a.NAp + j];
            in.tw = tw;
            in.L = (unsigned)P.L;
            in.NF = (unsigned)a.NF;
            in.Lh = (unsigned)P.Lh;
            in.N1 = P.N1;
            in.k1 = k1;
            if constexpr (std::is_void<PT>::value)
                fft_sweep<C, 1, E>(P, 0, v, acc, buf, twp, tw, w, P.Lh, in, ar, ai, t);
            else
                fft_sweep_fixed<C, E, PT, 0>(v[0], acc, buf, twp, mt, w, in, ar, ai, (unsigned)t);
        } else {
            RowInput<C, AmpRowSource<C>> in;
            in.src.p = a.amp + ((size_t)s * a.NM + pr) * a.NFp;
            in.src.split_stride = a.split_stride;
            in.src.split = a.split;
            in.src.h = h;
            in.src.tw = tw;
            in.tw = tw;
            in.L = (unsigned)P.L;
            in.NF = (unsigned)a.NF;
            in.Lh = (unsigned)P.Lh;
            in.N1 = P.N1;
            in.k1 = k1;
            fft_sweep<C, 1, E>(P, 0, v, acc, buf, twp, tw, w, P.Lh, in, ar, ai, t);
        }
        const double2 r = block_sum2(ar, ai, red, tid, P.T);
        if (tid == 0) apart_rows[((size_t)s * npairs + pr) * rows + hk] = r;
        if (sizeof(Acc) == 4 && ++since == kFlushEvery) {
            flush_power_row<E>(P, acc, row, first, tid, h, k1);
            first = false;
            since = 0;
        }
    }
    if (since > 0 || first || sizeof(Acc) != 4) flush_power_row<E>(P, acc, row, first, tid, h, k1);
}

using RowsFn = void (*)(CorrArgs, double2 *);

template <typename C, bool SELF>
static RowsFn rows_fn(int E) {
    switch (E) {
        case 4: return corr_rows_kernel<C, 4, SELF>;
        case 8: return corr_rows_kernel<C, 8, SELF>;
        case 16: return corr_rows_kernel<C, 16, SELF>;
        case 20: return corr_rows_kernel<C, 20, SELF>;
        case 24: return corr_rows_kernel<C, 24, SELF>;
        default: return nullptr;
    }
}

// The self rows kernel of a row plan fixed at compile time, if there is one.
static RowsFn rows_fixed_self(const FftPlan &P) {
    if (Plan256::matches(P, true)) return corr_rows_kernel<float2, 4, true, Plan256>;
    if (Plan512::matches(P, true)) return corr_rows_kernel<float2, 8, true, Plan512>;
    if (Plan1024::matches(P, true)) return corr_rows_kernel<float2, 8, true, Plan1024>;
    if (Plan1536::matches(P, true)) return corr_rows_kernel<float2, 24, true, Plan1536>;
    if (Plan2048::matches(P, true)) return corr_rows_kernel<float2, 8, true, Plan2048>;
    if (Plan2560::matches(P, true)) return corr_rows_kernel<float2, 20, true, Plan2560>;
    if (Plan3072::matches(P, true)) return corr_rows_kernel<float2, 24, true, Plan3072>;
    if (Plan4096::matches(P, true)) return corr_rows_kernel<float2, 16, true, Plan4096>;
    if (Plan5120::matches(P, true)) return corr_rows_kernel<float2, 20, true, Plan5120>;
    if (Plan6144::matches(P, true)) return corr_rows_kernel<float2, 24, true, Plan6144>;
    if (Plan8192::matches(P, true)) return corr_rows_kernel<float2, 16, true, Plan8192>;
    if (Plan10240::matches(P, true)) return corr_rows_kernel<float2, 20, true, Plan10240>;
    return nullptr;
}

hipError_t launch_corr_rows(const CorrArgs &a, bool self, double2 *apart_rows, hipStream_t stream) {
    if (a.plan.N1 < 2) return hipErrorInvalidValue;
    RowsFn k = self ? rows_fixed_self(a.plan) : nullptr;
    if (!k) k = self ? rows_fn<float2, true>(a.plan.E) : rows_fn<double2, false>(a.plan.E);
    if (!k) return hipErrorInvalidValue;
    // the transform data of one row; the block sum reuses its start
    const size_t lds = (self ? sizeof(float2) : sizeof(double2)) * (size_t)a.plan.Lh;
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3((unsigned)(a.nshells * a.chunks * 2 * a.plan.N1)), dim3(a.plan.T), lds, stream, a,
                       apart_rows);
    return hipGetLastError();
}

// The block reduction reuses the FFT buffer (see corr_all_block), so the
// dynamic LDS is exactly the transform data: 160 KB for 10240 FP64 points.
template <typename C, int NH>
static size_t corr_lds_bytes(const FftPlan &P) {
    return sizeof(C) * NH * P.Lh;
}

using CorrKernel = void (*)(CorrArgs);

template <typename C, int NH>
static hipError_t launch_corr(CorrKernel k4, CorrKernel k8, CorrKernel k16, CorrKernel k20, CorrKernel k24,
                              const CorrArgs &a, hipStream_t stream) {
    CorrKernel k = nullptr;
    switch (a.plan.E) {
        case 4: k = k4; break;
        case 8: k = k8; break;
        case 16: k = k16; break;
        case 20: k = k20; break;
        case 24: k = k24; break;
        default: return hipErrorInvalidValue;
    }
    const size_t lds = corr_lds_bytes<C, NH>(a.plan);
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3((unsigned)(a.nshells * a.chunks)), dim3(a.plan.T), lds, stream, a);
    return hipGetLastError();
}

size_t corr_all_lds_bytes(const FftPlan &P) { return corr_lds_bytes<double2, 1>(P); }

using SelfSplitFn = void (*)(CorrArgs, double2 *);

static SelfSplitFn self_split_fn(const FftPlan &P) {
    if (Plan256::matches(P)) return corr_self_split_kernel<float2, 4, Plan256>;
    if (Plan512::matches(P)) return corr_self_split_kernel<float2, 8, Plan512>;
    if (Plan1024::matches(P)) return corr_self_split_kernel<float2, 8, Plan1024>;
    if (Plan1536::matches(P)) return corr_self_split_kernel<float2, 24, Plan1536>;
    if (Plan2048::matches(P)) return corr_self_split_kernel<float2, 8, Plan2048>;
    if (Plan4096::matches(P)) return corr_self_split_kernel<float2, 16, Plan4096>;
    if (Plan5120::matches(P)) return corr_self_split_kernel<float2, 20, Plan5120>;
    if (Plan2560::matches(P)) return corr_self_split_kernel<float2, 20, Plan2560>;
    if (Plan3072::matches(P)) return corr_self_split_kernel<float2, 24, Plan3072>;
    if (Plan6144::matches(P)) return corr_self_split_kernel<float2, 24, Plan6144>;
    if (Plan8192::matches(P)) return corr_self_split_kernel<float2, 16, Plan8192>;
    if (Plan10240::matches(P)) return corr_self_split_kernel<float2, 20, Plan10240>;
    if (P.T > 256) return nullptr;  // the runtime-plan kernels are built for 256-thread blocks
    switch (P.E) {
        case 4: return corr_self_split_kernel<float2, 4>;
        case 8: return corr_self_split_kernel<float2, 8>;
        case 16: return corr_self_split_kernel<float2, 16>;
        case 20: return corr_self_split_kernel<float2, 20>;
        case 24: return corr_self_split_kernel<float2, 24>;
        default: return nullptr;
    }
}

bool self_split_supported(const FftPlan &P) { return P.N1 == 1 && self_split_fn(P) != nullptr; }

hipError_t self_split_blocks_per_cu(const FftPlan &P, int *blocks) {
    *blocks = 0;
    const SelfSplitFn k = self_split_fn(P);
    if (!k) return hipErrorInvalidValue;
    const size_t lds = corr_lds_bytes<float2, 1>(P);
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, (const void *)k, P.T, lds);
}

hipError_t launch_corr_self_split(const CorrArgs &a, double2 *apair, hipStream_t stream) {
    const SelfSplitFn k = self_split_fn(a.plan);
    if (!k) return hipErrorInvalidValue;
    const size_t lds = corr_lds_bytes<float2, 1>(a.plan);
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3((unsigned)(2 * a.nshells * a.chunks)), dim3(a.plan.T), lds, stream, a, apair);
    return hipGetLastError();
}

hipError_t launch_self_tail(const CorrArgs &a, double2 *tail, int tblocks, hipStream_t stream) {
    hipLaunchKernelGGL(self_tail_kernel, dim3((unsigned)tblocks, (unsigned)a.nshells), dim3(kTailThreads), 0, stream, a,
                       tail, tblocks);
    return hipGetLastError();
}

hipError_t launch_self_a_reduce(const double2 *apair, long npairs, int wpp, int nshells, int rblocks,
                                double *apart, hipStream_t stream) {
    hipLaunchKernelGGL(self_a_reduce_kernel, dim3((unsigned)rblocks, (unsigned)nshells), dim3(256), 0, stream, apair,
                       npairs, wpp, rblocks, apart);
    return hipGetLastError();
}

hipError_t launch_corr_all(const CorrArgs &a, hipStream_t stream) {
    return launch_corr<double2, 1>(corr_all_kernel<double2, 1, 4>, corr_all_kernel<double2, 1, 8>,
                                   corr_all_kernel<double2, 1, 16>, corr_all_kernel<double2, 1, 20>,
                                   corr_all_kernel<double2, 1, 24>, a, stream);
}

hipError_t launch_corr_self(const CorrArgs &a, int halves, hipStream_t stream) {
    if (halves == 2)
        return launch_corr<float2, 2>(corr_self_kernel<float2, 2, 4>, corr_self_kernel<float2, 2, 8>,
                                      corr_self_kernel<float2, 2, 16>, corr_self_kernel<float2, 2, 20>,
                                      corr_self_kernel<float2, 2, 24>, a, stream);
    return launch_corr<float2, 1>(corr_self_kernel<float2, 1, 4>, corr_self_kernel<float2, 1, 8>,
                                  corr_self_kernel<float2, 1, 16>, corr_self_kernel<float2, 1, 20>,
                                  corr_self_kernel<float2, 1, 24>, a, stream);
}

// ------------------------------------------------------ FP64 transforms ----

// w_f = sum_{tau<NF} c_tau e^{+2 pi i f tau / L}, c_tau = 1 / (NF L (NF - tau)),
// the phasor stepped by e^{+2 pi i f / L} and re-anchored on the FP64 table
// tw64 every 64 lags.  Stored in the order the last FFT pass of the plan
// visits the bins: bin f = 2 (k1 + N1 k2) + h, k2 = j + q Lh / R, at
// (h N1 + k1) Lh + q Lh / R + j (N1 = 1: the one-block halves).
__global__ __launch_bounds__(256) void weights_kernel(int L, int NF, int R, int N1, int Lh,
                                                      const double2 *__restrict__ tw64, double2 *__restrict__ w) {
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= L) return;
    double wr = 0.0, wi = 0.0;
    const double2 st = tw64[f];
    long anchor = 0;
    const long jump = (64L * f) % L;
    for (int t0 = 0; t0 < NF; t0 += 64) {
        double2 z = tw64[anchor];
        anchor += jump;
        if (anchor >= L) anchor -= L;
        const int t1 = min(t0 + 64, NF);
        for (int tau = t0; tau < t1; ++tau) {
            const double c = 1.0 / ((double)NF * (double)L * (double)(NF - tau));
            wr = fma(c, z.x, wr);
            wi = fma(c, z.y, wi);
            const double zr = z.x * st.x - z.y * st.y;
            z.y = z.x * st.y + z.y * st.x;
            z.x = zr;
        }
    }
    const int stride = Lh / R;
    const int h = f & 1, fh = f >> 1;
    const int k1 = fh % N1, k2 = fh / N1;
    const int q = k2 / stride, j = k2 - q * stride;
    w[((size_t)(h * N1 + k1)) * Lh + (size_t)q * stride + j] = make_double2(wr, wi);
}

hipError_t launch_weights(const FftPlan &P, int NF, const double2 *tw64, double2 *w, hipStream_t stream) {
    hipLaunchKernelGGL(weights_kernel, dim3((unsigned)((P.L + 255) / 256)), dim3(256), 0, stream, P.L, NF,
                       P.radix[P.npass - 1], P.N1, P.Lh, tw64, w);
    return hipGetLastError();
}

// fqt(tau) = R(tau) / (L (NF - tau) NM), R = sum_f P_f e^{+2 pi i f tau / L},
// P = sum over chunks of the partial power spectra (all FP64); fq and fq2
// from the per-chunk sums of a and |a|^2 (scaled by 1/NM only, as in
// all_vectors_scatter_device.cpp:355-360 and self_vectors_scatter_device.cpp:233-238).
__global__ __launch_bounds__(256) void final_kernel(FinalArgs a) {
    extern __shared__ double Psh[];
    final_block(a, blockIdx.y, blockIdx.x, Psh);
}

hipError_t launch_final(const FinalArgs &a, hipStream_t stream) {
    const size_t lds = a.ppart ? final_lds_bytes(a.L) : 0;
    if (lds > 0) {
        hipError_t e =
            hipFuncSetAttribute((const void *)final_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const unsigned tiles = a.ppart ? (unsigned)((a.NF + 255) / 256) : 1u;
    hipLaunchKernelGGL(final_kernel, dim3(tiles, (unsigned)a.nshells), dim3(256), lds, stream, a);
    return hipGetLastError();
}

// ------------------------------------------------- square / plain series ---

// fqt(t) = (1/NM) sum_m S_m(t), S = |A|^2 (square, smath.cpp:168-177) or A (plain)
__global__ __launch_bounds__(256) void series_fqt_kernel(int mode, const float2 *__restrict__ amp, size_t NFp,
                                                         int split, size_t split_stride, int NF, int NM,
                                                         double *__restrict__ fqt, double norm) {
    const int s = blockIdx.y;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= NF) return;
    double re = 0.0, im = 0.0;
    for (int m = 0; m < NM; ++m) {
        const size_t v = (size_t)s * NM + m;
        float2 x = amp[v * NFp + t];
        for (int k = 1; k < split; ++k) {
            const float2 y = amp[(size_t)k * split_stride + v * NFp + t];
            x.x += y.x;
            x.y += y.y;
        }
        if (mode == 1) {
            re += (double)x.x * x.x + (double)x.y * x.y;
        } else {
            re += x.x;
            im += x.y;
        }
    }
    fqt[((size_t)s * NF + t) * 2] = re * norm;
    fqt[((size_t)s * NF + t) * 2 + 1] = im * norm;
}

// per chunk of vectors: sum_m a_m and sum_m |a_m|^2 with a_m = mean_t S_m(t)
__global__ __launch_bounds__(256) void series_mean_kernel(int mode, const float2 *__restrict__ amp, size_t NFp,
                                                          int split, size_t split_stride, int NF, int NM,
                                                          int chunks, double *__restrict__ apart) {
    __shared__ double2 red[4];
    const int s = blockIdx.x / chunks, c = blockIdx.x % chunks;
    const int m0 = (int)((long)NM * c / chunks), m1 = (int)((long)NM * (c + 1) / chunks);
    const int tid = threadIdx.x;
    double sa_r = 0.0, sa_i = 0.0, sa2 = 0.0;
    for (int m = m0; m < m1; ++m) {
        const size_t v = (size_t)s * NM + m;
        double re = 0.0, im = 0.0;
        for (int t = tid; t < NF; t += 256) {
            float2 x = amp[v * NFp + t];
            for (int k = 1; k < split; ++k) {
                const float2 y = amp[(size_t)k * split_stride + v * NFp + t];
                x.x += y.x;
                x.y += y.y;
            }
            if (mode == 1) {
                re += (double)x.x * x.x + (double)x.y * x.y;
            } else {
                re += x.x;
                im += x.y;
            }
        }
        const double2 r = block_sum2(re, im, red, tid, 256);
        if (tid == 0) {
            const double ar = r.x / NF, ai = r.y / NF;
            sa_r += ar;
            sa_i += ai;
            sa2 += ar * ar + ai * ai;
        }
    }
    if (tid == 0) {
        double *ap = apart + ((size_t)s * chunks + c) * 4;
        ap[0] = sa_r;
        ap[1] = sa_i;
        ap[2] = sa2;
        ap[3] = 0.0;
    }
}

hipError_t launch_series_all(int mode, const float2 *amp, size_t NFp, int split, size_t split_stride, int NF, int NM,
                             int nshells, int chunks, double *fqt, double *apart, double norm, hipStream_t stream) {
    hipLaunchKernelGGL(series_fqt_kernel, dim3((unsigned)((NF + 255) / 256), (unsigned)nshells), dim3(256), 0,
                       stream, mode, amp, NFp, split, split_stride, NF, NM, fqt, norm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(series_mean_kernel, dim3((unsigned)(nshells * chunks)), dim3(256), 0, stream, mode, amp,
                       NFp, split, split_stride, NF, NM, chunks, apart);
    return hipGetLastError();
}


// ------------------------------------------------ self square / plain -----

// "self" with dsp.type square or plain (self_vectors_scatter_device.cpp:115-135):
// per pair the series S(t) = |b e^{i phi}|^2 or b e^{i phi}; one block per
// (shell, chunk of pairs, frame tile of at most kMaxSeriesFrames frames) keeps
// its partial sum over pairs of S(t) in LDS (each frame owned by one thread).
// One tile: the block also forms a = mean_t S per pair and its chunk's
// sums of a and |a|^2; several tiles: the tile's part of a per pair goes to
// parts[shell][pair][tile] for self_a_reduce.
__global__ __launch_bounds__(256) void self_series_kernel(int mode, CorrArgs a, double2 *__restrict__ rows,
                                                          int tiles, double2 *__restrict__ parts) {
    extern __shared__ double2 row[];
    __shared__ double2 red[4];
    const int tid = threadIdx.x;
    const int tile = blockIdx.x % tiles, sc = blockIdx.x / tiles;
    const int s = sc / a.chunks, c = sc % a.chunks;
    const int t0 = tile * kMaxSeriesFrames, t1 = min(a.NF, t0 + kMaxSeriesFrames), nt = t1 - t0;
    const long npairs = (long)a.NA * a.NM;
    const long p0 = npairs * c / a.chunks, p1 = npairs * (c + 1) / a.chunks;
    for (int t = tid; t < nt; t += 256) row[t] = make_double2(0.0, 0.0);
    double sa_r = 0.0, sa_i = 0.0, sa2 = 0.0;
    const size_t plane = a.NAp * a.NFp;
    for (long pr = p0; pr < p1; ++pr) {
        const int j = (int)(pr / a.NM), m = (int)(pr % a.NM);
        const double *q = a.qrev + ((size_t)s * a.NM + m) * 3;
        SelfInput in;
        in.set(a.coords + (size_t)j * a.NFp, plane, a.NF);
        in.qx = q[0];
        in.qy = q[1];
        in.qz = q[2];
        in.b = a.b[(size_t)(a.brow_per_shell ? s : 0) * a.NAp + j];
        double ar = 0.0, ai = 0.0;
        for (int t = tid; t < nt; t += 256) {
            const float2 x = in(t0 + t);
            double sr, si;
            if (mode == 1) {
                sr = (double)x.x * x.x + (double)x.y * x.y;
                si = 0.0;
            } else {
                sr = x.x;
                si = x.y;
            }
            row[t].x += sr;
            row[t].y += si;
            ar += sr;
            ai += si;
        }
        const double2 r = block_sum2(ar, ai, red, tid, 256);
        if (tid == 0) {
            const double xr = r.x / a.NF, xi = r.y / a.NF;
            if (tiles > 1) {
                parts[((size_t)s * npairs + pr) * tiles + tile] = make_double2(xr, xi);
            } else {
                sa_r += xr;
                sa_i += xi;
                sa2 += xr * xr + xi * xi;
            }
        }
    }
    double2 *dst = rows + ((size_t)s * a.chunks + c) * a.NF + t0;
    for (int t = tid; t < nt; t += 256) dst[t] = row[t];
    if (tid == 0 && tiles == 1) {
        double *ap = a.apart + ((size_t)s * a.chunks + c) * 4;
        ap[0] = sa_r;
        ap[1] = sa_i;
        ap[2] = sa2;
        ap[3] = 0.0;
    }
}

// fqt(t) = (1/NM) sum over chunks of the partial series sums
__global__ __launch_bounds__(256) void series_rows_final_kernel(int NF, int NM, int chunks,
                                                                const double2 *__restrict__ rows,
                                                                double *__restrict__ fqt) {
    const int s = blockIdx.y;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= NF) return;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < chunks; ++c) {
        const double2 v = rows[((size_t)s * chunks + c) * NF + t];
        re += v.x;
        im += v.y;
    }
    const double f = 1.0 / (double)NM;
    fqt[((size_t)s * NF + t) * 2] = re * f;
    fqt[((size_t)s * NF + t) * 2 + 1] = im * f;
}

hipError_t launch_series_self(int mode, const CorrArgs &a, double2 *rows, double *fqt, double2 *parts,
                              hipStream_t stream) {
    const int tiles = series_tiles(a.NF);
    const size_t lds = sizeof(double2) * (size_t)std::min(a.NF, kMaxSeriesFrames);
    hipError_t e =
        hipFuncSetAttribute((const void *)self_series_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(self_series_kernel, dim3((unsigned)(a.nshells * a.chunks * tiles)), dim3(256), lds, stream,
                       mode, a, rows, tiles, parts);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(series_rows_final_kernel, dim3((unsigned)((a.NF + 255) / 256), (unsigned)a.nshells), dim3(256),
                       0, stream, a.NF, a.NM, a.chunks, rows, fqt);
    return hipGetLastError();
}

}  // namespace sx
