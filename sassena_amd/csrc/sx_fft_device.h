// sx_fft_device.h -- device-side building blocks of the time autocorrelation:
// complex arithmetic in FP32/FP64, radix 2/3/4/5/8 butterflies, Stockham
// passes in LDS, the per-series power spectrum and the per-block bodies that
// the single-trajectory and the batched kernels share.
#pragma once

#include "sx_internal.h"

namespace sx {

// ------------------------------------------------------ complex arithmetic --
// The same Stockham code runs in FP32 (float2) and FP64 (double2).

template <typename C>
struct Cx;
template <>
struct Cx<float2> {
    using R = float;
    static __device__ __forceinline__ float2 make(float x, float y) { return make_float2(x, y); }
};
template <>
struct Cx<double2> {
    using R = double;
    static __device__ __forceinline__ double2 make(double x, double y) { return make_double2(x, y); }
};

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
    return Cx<C>::make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
    return Cx<C>::make(a.x + b.x, a.y + b.y);
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
    return Cx<C>::make(a.x - b.x, a.y - b.y);
}
// multiply by -i (forward transform rotation)
template <typename C>
__device__ __forceinline__ C cmul_mi(C a) {
    return Cx<C>::make(a.y, -a.x);
}

// ------------------------------------------------------------ butterflies --
// Forward DFTs of radix 2, 3, 4, 5, 8 (sign -1), in place.

template <int R, typename C>
struct Dft;

template <typename C>
struct Dft<2, C> {
    static __device__ __forceinline__ void run(C *x) {
        const C a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    }
};

template <typename C>
struct Dft<4, C> {
    static __device__ __forceinline__ void run(C *x) {
        const C a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
        const C c = cadd(x[1], x[3]), d = cmul_mi(csub(x[1], x[3]));
        x[0] = cadd(a, c);
        x[2] = csub(a, c);
        x[1] = cadd(b, d);
        x[3] = csub(b, d);
    }
};

template <typename C>
struct Dft<8, C> {
    static __device__ __forceinline__ void run(C *x) {
        using R = typename Cx<C>::R;
        C e[4] = {x[0], x[2], x[4], x[6]};
        C o[4] = {x[1], x[3], x[5], x[7]};
        Dft<4, C>::run(e);
        Dft<4, C>::run(o);
        const R s = R(0.70710678118654752440);
        // o_k *= W8^k, W8 = exp(-i pi/4)
        o[1] = Cx<C>::make(s * (o[1].x + o[1].y), s * (o[1].y - o[1].x));
        o[2] = cmul_mi(o[2]);
        o[3] = Cx<C>::make(s * (o[3].y - o[3].x), -s * (o[3].x + o[3].y));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[k] = cadd(e[k], o[k]);
            x[k + 4] = csub(e[k], o[k]);
        }
    }
};

template <typename C>
struct Dft<3, C> {
    static __device__ __forceinline__ void run(C *x) {
        using R = typename Cx<C>::R;
        const R s = R(0.86602540378443864676);  // sin(2 pi / 3)
        const C t = cadd(x[1], x[2]), u = csub(x[1], x[2]);
        const C a = Cx<C>::make(x[0].x - R(0.5) * t.x, x[0].y - R(0.5) * t.y);
        x[0] = cadd(x[0], t);
        // y1 = a - i s u, y2 = a + i s u
        x[1] = Cx<C>::make(a.x + s * u.y, a.y - s * u.x);
        x[2] = Cx<C>::make(a.x - s * u.y, a.y + s * u.x);
    }
};

template <typename C>
struct Dft<5, C> {
    static __device__ __forceinline__ void run(C *x) {
        using R = typename Cx<C>::R;
        const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
        const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
        const C t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
        const C t3 = csub(x[1], x[4]), t4 = csub(x[2], x[3]);
        const C a1 = Cx<C>::make(x[0].x + c1 * t1.x + c2 * t2.x, x[0].y + c1 * t1.y + c2 * t2.y);
        const C a2 = Cx<C>::make(x[0].x + c2 * t1.x + c1 * t2.x, x[0].y + c2 * t1.y + c1 * t2.y);
        const C b1 = Cx<C>::make(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
        const C b2 = Cx<C>::make(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
        x[0] = cadd(x[0], cadd(t1, t2));
        // y1 = a1 - i b1, y4 = a1 + i b1, y2 = a2 - i b2, y3 = a2 + i b2
        x[1] = Cx<C>::make(a1.x + b1.y, a1.y - b1.x);
        x[4] = Cx<C>::make(a1.x - b1.y, a1.y + b1.x);
        x[2] = Cx<C>::make(a2.x + b2.y, a2.y - b2.x);
        x[3] = Cx<C>::make(a2.x - b2.y, a2.y + b2.x);
    }
};

// ------------------------------------------------------------ passes ------
// A sweep transforms NH halves at once (NH = 2: both halves of the padded
// series share every barrier; NH = 1: one half per sweep, half the LDS).
// Half k of the sweep lives at buf[k*Lh, (k+1)*Lh) and is global half h0+k.

// tw[base + k] through a buffer resource, k per lane and base uniform.
template <typename C>
__device__ __forceinline__ C twiddle_load(const C *tw, unsigned k, unsigned base) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)tw, (short)0, 0x7fffffff, kBufferDword3);
    if constexpr (sizeof(C) == 8) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, k * 8u, base * 8u, 0);
        return Cx<C>::make(__uint_as_float(v[0]), __uint_as_float(v[1]));
    } else {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, k * 16u, base * 16u, 0);
        return Cx<C>::make(__hiloint2double((int)v[1], (int)v[0]), __hiloint2double((int)v[3], (int)v[2]));
    }
}

// w[base + j] through a buffer resource, j per lane and base uniform.
__device__ __forceinline__ double2 weight_load(const double2 *w, unsigned j, unsigned base) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)w, (short)0, 0x7fffffff, kBufferDword3);
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, j * 16u, base * 16u, 0);
    return make_double2(__hiloint2double((int)v[1], (int)v[0]), __hiloint2double((int)v[3], (int)v[2]));
}

// Middle pass p: read radix-R groups from LDS, twiddle, butterfly, write back.
template <typename C, int NH, int E, int R>
__device__ __forceinline__ void pass_mid(const FftPlan &P, int p, C (&v)[NH][E], C *buf,
                                         const C *__restrict__ twp, int tid) {
    if constexpr (E % R == 0) {
        const unsigned Lh = P.Lh, T = P.T, Ns = P.ns[p];
        const unsigned stride = Lh / R;
        const unsigned ut = (unsigned)tid;
        // With Ns <= T (both powers of two) every group of this thread has the
        // same twiddle index kk = tid mod Ns: load its R-1 twiddles once, and
        // before the barrier, so the load latency hides behind it.
        const bool shared = Ns <= T;
        C ws[R];
        if (shared) {
            const unsigned kk = (ut & (Ns - 1)) * R;
#pragma unroll
            for (int q = 1; q < R; ++q) ws[q] = twiddle_load(twp, kk, P.toff[p] + q);
        }
#pragma unroll
        for (int i = 0; i < E / R; ++i) {
            const unsigned j = ut + i * T;
#pragma unroll
            for (int q = 0; q < R; ++q)
#pragma unroll
                for (int k = 0; k < NH; ++k) v[k][i * R + q] = buf[k * Lh + j + q * stride];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < E / R; ++i) {
            const unsigned j = ut + i * T;
            const unsigned kk = j & (Ns - 1);
            if (shared) {
#pragma unroll
                for (int q = 1; q < R; ++q)
#pragma unroll
                    for (int k = 0; k < NH; ++k) v[k][i * R + q] = cmul(v[k][i * R + q], ws[q]);
            } else {
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    const C w = twiddle_load(twp, kk * R, P.toff[p] + q);
#pragma unroll
                    for (int k = 0; k < NH; ++k) v[k][i * R + q] = cmul(v[k][i * R + q], w);
                }
            }
            const unsigned o = (j - kk) * R + kk;
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                Dft<R, C>::run(&v[k][i * R]);
#pragma unroll
                for (int q = 0; q < R; ++q) buf[k * Lh + o + q * Ns] = v[k][i * R + q];
            }
        }
        __syncthreads();
    }
}

// Last pass: its outputs land on bins f' = j + q*Lh/R of each half, the same
// bins for every series, so the power spectrum accumulates in registers; the
// full-length bin is f = 2 f' + h.
template <typename C, int NH, int E, int R, typename Acc>
__device__ __forceinline__ void pass_last(const FftPlan &P, int h0, C (&v)[NH][E], Acc (&acc)[2][E],
                                          const C *buf, const C *__restrict__ twp,
                                          const double2 *__restrict__ w, double &ar, double &ai, int tid) {
    if constexpr (E % R == 0) {
        const unsigned Lh = P.Lh, T = P.T;
        const unsigned stride = Lh / R;  // == Ns of this pass
        const unsigned ut = (unsigned)tid;
#pragma unroll
        for (int i = 0; i < E / R; ++i) {
            const unsigned j = ut + i * T;
#pragma unroll
            for (int q = 0; q < R; ++q)
#pragma unroll
                for (int k = 0; k < NH; ++k) v[k][i * R + q] = buf[k * Lh + j + q * stride];
        }
        __syncthreads();  // the buffer may now be overwritten by the next sweep
#pragma unroll
        for (int i = 0; i < E / R; ++i) {
            const unsigned j = ut + i * T;  // k == j in the last pass
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const C t = twiddle_load(twp, j * R, P.toff[P.npass - 1] + q);
#pragma unroll
                for (int k = 0; k < NH; ++k) v[k][i * R + q] = cmul(v[k][i * R + q], t);
            }
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                Dft<R, C>::run(&v[k][i * R]);
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const C x = v[k][i * R + q];
                    const typename Cx<C>::R pw = x.x * x.x + x.y * x.y;
                    acc[h0 + k][i * R + q] += (Acc)pw;
                    // bin f = 2 (j + q stride) + h0 + k, weights in visiting order:
                    // the thread's j as the vector offset, the rest as the
                    // scalar one (no 64-bit address per load)
                    const double2 wf = weight_load(w, j, ((unsigned)(h0 + k) * R + q) * stride);
                    ar = fma(wf.x, (double)pw, ar);
                    ai = fma(wf.y, (double)pw, ai);
                }
            }
        }
    }
}

// First pass: inputs come from a functor (HBM amplitudes or on-the-fly
// phases); half 1 gets the pre-twiddle W_L^n = pre[n].  Stride Ns = 1, so the
// pass itself needs no twiddles.
template <typename C, int NH, int E, int R, typename Input>
__device__ __forceinline__ void pass_first(const FftPlan &P, int h0, C (&v)[NH][E], C *buf,
                                           const C *__restrict__ pre, int NF, const Input &in, int tid) {
    if constexpr (E % R == 0) {
        const unsigned Lh = P.Lh, T = P.T;
        const unsigned stride = Lh / R;
        const unsigned unf = (unsigned)NF;
#pragma unroll
        for (int i = 0; i < E / R; ++i) {
            const unsigned j = (unsigned)tid + i * T;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const unsigned n = j + q * stride;
                C x = Cx<C>::make(0, 0);
                if constexpr (Input::kBounded)
                    x = in(n);  // 0 past NF by itself
                else if (n < unf)
                    x = in(n);
#pragma unroll
                for (int k = 0; k < NH; ++k)
                    v[k][i * R + q] = (Input::kPretwiddled || h0 + k == 0 || n >= unf) ? x : cmul(x, pre[n]);
            }
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                Dft<R, C>::run(&v[k][i * R]);
#pragma unroll
                for (int q = 0; q < R; ++q) buf[k * Lh + j * R + q] = v[k][i * R + q];
            }
        }
        __syncthreads();
    }
}

// One sweep: the Stockham passes of NH halves starting at half h0.
template <typename C, int NH, int E, typename Acc, typename Input>
__device__ __forceinline__ void fft_sweep(const FftPlan &P, int h0, C (&v)[NH][E], Acc (&acc)[2][E], C *buf,
                                          const C *__restrict__ tw, const C *__restrict__ pre,
                                          const double2 *__restrict__ w, int NF, const Input &in, double &ar,
                                          double &ai, int tid) {
    switch (P.radix[0]) {
        case 2: pass_first<C, NH, E, 2>(P, h0, v, buf, pre, NF, in, tid); break;
        case 4: pass_first<C, NH, E, 4>(P, h0, v, buf, pre, NF, in, tid); break;
        case 8: pass_first<C, NH, E, 8>(P, h0, v, buf, pre, NF, in, tid); break;
        default: break;
    }
    for (int p = 1; p < P.npass - 1; ++p) {
        switch (P.radix[p]) {
            case 2: pass_mid<C, NH, E, 2>(P, p, v, buf, tw, tid); break;
            case 4: pass_mid<C, NH, E, 4>(P, p, v, buf, tw, tid); break;
            case 8: pass_mid<C, NH, E, 8>(P, p, v, buf, tw, tid); break;
            default: break;
        }
    }
    switch (P.radix[P.npass - 1]) {
        case 2: pass_last<C, NH, E, 2>(P, h0, v, acc, buf, tw, w, ar, ai, tid); break;
        case 3: pass_last<C, NH, E, 3>(P, h0, v, acc, buf, tw, w, ar, ai, tid); break;
        case 4: pass_last<C, NH, E, 4>(P, h0, v, acc, buf, tw, w, ar, ai, tid); break;
        case 5: pass_last<C, NH, E, 5>(P, h0, v, acc, buf, tw, w, ar, ai, tid); break;
        case 8: pass_last<C, NH, E, 8>(P, h0, v, acc, buf, tw, w, ar, ai, tid); break;
        default: break;
    }
}

// ------------------------------------------- passes of a fixed plan ------
// The same Stockham sweep (one half, NH = 1) for a plan known at compile
// time: strides, Ns and twiddle offsets become constants, so every LDS
// access is the lane's base address plus an immediate offset and the integer
// index arithmetic of the runtime plan disappears.  The twiddle table is the
// runtime plan's (StaticPlan::matches checks that the layouts agree).
template <unsigned LH, unsigned TT, int... RS>
struct StaticPlan {
    static constexpr unsigned Lh = LH, T = TT;
    static constexpr int npass = sizeof...(RS);
    static constexpr int kRadix[npass] = {RS...};
    static constexpr int radix(int p) { return kRadix[p]; }
    static constexpr unsigned ns(int p) {
        unsigned n = 1;
        for (int i = 0; i < p; ++i) n *= (unsigned)kRadix[i];
        return n;
    }
    static constexpr unsigned toff(int p) {
        unsigned o = 0;
        for (int i = 1; i < p; ++i) o += ns(i) * (unsigned)kRadix[i];
        return o;
    }
    // The LDS layout of the buffer pass p writes and pass p + 1 reads: float2
    // index o lives at swz(p, o) = o ^ (((o >> S) mod 2^n) << B).  Pass p
    // writes o = (j - kk) R + kk + q Ns (kk = j mod Ns): the bits below
    // log2 Ns vary over the lanes, those from log2 Ns to log2(Ns R) hold q,
    // the same for every lane of one store, so with Ns < 16 a store covers
    // only Ns (or, 16-byte stores of the first pass, 2 R) of the 16 float2
    // positions of a 32-bank row.  XOR-ing the bits from B = max(1, log2 Ns)
    // to 3 with lane bits from S = max(4, log2(Ns R)) up spreads every store
    // over all 16 (bit 0 is kept, so pairs stay adjacent for 16-byte stores).
    // Readers take ut + C with C a multiple of 2^(S+n) (required below), so
    // their address is the lane's swz(p, ut) plus the same immediate C.
    static constexpr int lg2(unsigned x) {
        int l = 0;
        while ((1u << l) < x) ++l;
        return l;
    }
    static constexpr bool pow2(unsigned x) { return (x & (x - 1)) == 0; }
    static constexpr int kSwzTop = 4;
    static constexpr int swz_b(int p) { return lg2(ns(p)) < 1 ? 1 : lg2(ns(p)); }
    static constexpr int swz_s(int p) {
        return lg2(ns(p) * (unsigned)radix(p)) > kSwzTop ? lg2(ns(p) * (unsigned)radix(p)) : kSwzTop;
    }
    static constexpr int swz_n(int p) {
        if (p + 1 >= npass || swz_b(p) >= kSwzTop || ns(p) > T || !pow2(ns(p)) || !pow2((unsigned)radix(p)))
            return 0;
        const unsigned m = 1u << (swz_s(p) + kSwzTop - swz_b(p));
        if (T % m != 0 || (Lh / (unsigned)radix(p + 1)) % m != 0) return 0;
        return kSwzTop - swz_b(p);
    }
    static constexpr unsigned swz(int p, unsigned o) {
        return swz_n(p) == 0 ? o : o ^ (((o >> swz_s(p)) & ((1u << swz_n(p)) - 1u)) << swz_b(p));
    }
    // rows: the plan of one row of a long series (any N1) rather than a one-block plan
    static bool matches(const FftPlan &P, bool rows = false) {
        if (P.Lh != (int)Lh || P.T != (int)T || P.npass != npass || (rows ? P.N1 < 2 : P.N1 != 1)) return false;
        for (int i = 0; i < npass; ++i)
            if (P.radix[i] != kRadix[i] || P.ns[i] != (int)ns(i) || (i > 0 && P.toff[i] != (int)toff(i)))
                return false;
        return true;
    }
};

template <typename C, int E, typename PT, typename Input>
__device__ __forceinline__ void spass_first(C (&v)[E], C *buf, const Input &in, unsigned ut) {
    constexpr int R = PT::radix(0);
    constexpr unsigned T = PT::T, stride = PT::Lh / R;
    static_assert(E % R == 0, "first pass of a fixed plan");
    if constexpr (!Input::kSplitLoad) {
        // a row of a long series: every point of the row exists, formed in one call
#pragma unroll
        for (int i = 0; i < E / R; ++i) {
#pragma unroll
            for (int q = 0; q < R; ++q) v[i * R + q] = in(ut + i * T + q * stride);
            Dft<R, C>::run(&v[i * R]);
#pragma unroll
            for (int q = 0; q < R; ++q) buf[PT::swz(0, ut * R + (q & ~1)) + (q & 1) + i * T * R] = v[i * R + q];
        }
    } else {
        static_assert(Input::kBounded, "first pass of a fixed plan");
        // the coordinate loads of half the pass's points in flight before the
        // first is used, those of the other half issued once the first half is
        // being consumed (all at once: 128 VGPRs and a spill)
        float xs[E], ys[E], zs[E];
#pragma unroll
        for (int e = 0; e < E / 2; ++e) in.load(ut + (e / R) * T + (e % R) * stride, xs[e], ys[e], zs[e]);
#pragma unroll
        for (int i = 0; i < E / R; ++i) {
            if (i == (E / R) / 2) {
#pragma unroll
                for (int e = E / 2; e < E; ++e) in.load(ut + (e / R) * T + (e % R) * stride, xs[e], ys[e], zs[e]);
            }
#pragma unroll
            for (int q = 0; q < R; ++q)
                v[i * R + q] = in.at(ut + i * T + q * stride, xs[i * R + q], ys[i * R + q], zs[i * R + q]);
            Dft<R, C>::run(&v[i * R]);
#pragma unroll
            for (int q = 0; q < R; ++q) buf[PT::swz(0, ut * R + (q & ~1)) + (q & 1) + i * T * R] = v[i * R + q];
        }
    }
    __syncthreads();
}

// The twiddles of the middle passes of a fixed plan for one thread: every
// group of the thread has the index kk = ut mod Ns in pass p, so each pass
// needs R - 1 of them, the same for every series; loaded once per kernel and
// kept in registers.
template <typename C, typename PT>
struct MidTwiddles {
    C w[PT::npass][8];
    const C *twp;  // the pass twiddle table (passes with Ns > T load theirs per series)
    template <int p>
    __device__ __forceinline__ void load_from(const C *__restrict__ twp, unsigned ut) {
        if constexpr (p < PT::npass - 1) {
            constexpr int R = PT::radix(p);
            if constexpr (PT::ns(p) <= PT::T) {
                const unsigned kk = ut & (PT::ns(p) - 1);
#pragma unroll
                for (int q = 1; q < R; ++q) w[p][q] = twiddle_load(twp, kk * R, PT::toff(p) + q);
            }
            load_from<p + 1>(twp, ut);
        }
    }
    __device__ __forceinline__ void load(const C *__restrict__ tw, unsigned ut) {
        twp = tw;
        load_from<1>(tw, ut);
    }
};

template <typename C, typename PT>
__device__ __forceinline__ const C *twp_of(const MidTwiddles<C, PT> &mt) {
    return mt.twp;
}

template <typename C, int E, typename PT, int p>
__device__ __forceinline__ void spass_mid(C (&v)[E], C *buf, const MidTwiddles<C, PT> &mt, unsigned ut) {
    constexpr int R = PT::radix(p);
    constexpr unsigned T = PT::T, Ns = PT::ns(p), stride = PT::Lh / R;
    static_assert(E % R == 0, "middle pass of a fixed plan");
    if constexpr (Ns > T) {
        // groups j = ut + i T have kk = ut + (i T mod Ns) (Ns a multiple of
        // T): twiddle q is the lane's exp(-2 pi i q ut / (Ns R)) (R - 1 loads)
        // times a constant of the plan, and the outputs go to
        // (i T - (i T mod Ns)) R + kk + q Ns, the lane's base plus a constant
        static_assert(Ns % T == 0, "middle pass of a fixed plan");
        using Rl = typename Cx<C>::R;
        C base[R];
#pragma unroll
        for (int q = 1; q < R; ++q) base[q] = twiddle_load(twp_of(mt), ut * R, PT::toff(p) + q);
        const unsigned rb = PT::swz(p - 1, ut);
#pragma unroll
        for (int i = 0; i < E / R; ++i)
#pragma unroll
            for (int q = 0; q < R; ++q) v[i * R + q] = buf[rb + i * T + q * stride];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < E / R; ++i) {
            constexpr double kTurn = -6.283185307179586476925286766559 / (double)(Ns * R);
            const unsigned im = (i * T) % Ns;
#pragma unroll
            for (int q = 1; q < R; ++q) {
                C t = base[q];
                if (im != 0)
                    t = cmul(t, Cx<C>::make((Rl)__builtin_cos(kTurn * (double)(q * im)),
                                            (Rl)__builtin_sin(kTurn * (double)(q * im))));
                v[i * R + q] = cmul(v[i * R + q], t);
            }
            Dft<R, C>::run(&v[i * R]);
#pragma unroll
            for (int q = 0; q < R; ++q) buf[ut + (i * T - im) * R + im + q * Ns] = v[i * R + q];
        }
        __syncthreads();
    } else {
        // every group of the thread has the twiddle index kk = ut mod Ns
        const unsigned kk = ut & (Ns - 1);
        const C(&ws)[8] = mt.w[p];
        const unsigned rb = PT::swz(p - 1, ut);
#pragma unroll
        for (int i = 0; i < E / R; ++i)
#pragma unroll
            for (int q = 0; q < R; ++q) v[i * R + q] = buf[rb + i * T + q * stride];
        __syncthreads();
        const unsigned ob = (ut - kk) * R + kk;
#pragma unroll
        for (int i = 0; i < E / R; ++i) {
#pragma unroll
            for (int q = 1; q < R; ++q) v[i * R + q] = cmul(v[i * R + q], ws[q]);
            Dft<R, C>::run(&v[i * R]);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if constexpr (PT::swz_n(p) > 0)
                    buf[PT::swz(p, ob + q * Ns) + i * T * R] = v[i * R + q];
                else
                    buf[ob + i * T * R + q * Ns] = v[i * R + q];
            }
        }
        __syncthreads();
    }
}

// The last pass's twiddles exp(-2 pi i q (ut + i T) / Lh) are the thread's
// exp(-2 pi i q ut / Lh) (R - 1 loads, issued before the pass's barrier)
// times constants of the plan.
template <typename C, int E, typename PT, int H, typename Acc>
__device__ __forceinline__ void spass_last(C (&v)[E], Acc (&acc)[2][E], const C *buf, const C *__restrict__ twp,
                                           const double2 *__restrict__ w, double &ar, double &ai, unsigned ut) {
    constexpr int p = PT::npass - 1, R = PT::radix(p);
    constexpr unsigned T = PT::T, stride = PT::Lh / R;
    static_assert(E % R == 0 && PT::ns(p) == stride, "last pass of a fixed plan");
    using Rl = typename Cx<C>::R;
    C base[R];
#pragma unroll
    for (int q = 1; q < R; ++q) base[q] = twiddle_load(twp, ut * R, PT::toff(p) + q);
    const unsigned rb = PT::swz(p - 1, ut);
#pragma unroll
    for (int i = 0; i < E / R; ++i)
#pragma unroll
        for (int q = 0; q < R; ++q) v[i * R + q] = buf[rb + i * T + q * stride];
    __syncthreads();  // the buffer may now be overwritten by the next series
#pragma unroll
    for (int i = 0; i < E / R; ++i) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
            // exp(-2 pi i q i T / Lh), a constant of the plan
            constexpr double kTurn = -6.283185307179586476925286766559 / (double)PT::Lh;
            C t = base[q];
            if (i > 0) {
                const Rl cr = (Rl)__builtin_cos(kTurn * (double)(q * i * T));
                const Rl ci = (Rl)__builtin_sin(kTurn * (double)(q * i * T));
                t = cmul(t, Cx<C>::make(cr, ci));
            }
            v[i * R + q] = cmul(v[i * R + q], t);
        }
        Dft<R, C>::run(&v[i * R]);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const C x = v[i * R + q];
            const typename Cx<C>::R pw = x.x * x.x + x.y * x.y;
            acc[H][i * R + q] += (Acc)pw;
            const double2 wf = weight_load(w, ut, ((unsigned)H * R + q) * stride + i * T);
            ar = fma(wf.x, (double)pw, ar);
            ai = fma(wf.y, (double)pw, ai);
        }
    }
}

template <typename C, int E, typename PT, int p, int END>
__device__ __forceinline__ void spass_mids(C (&v)[E], C *buf, const MidTwiddles<C, PT> &mt, unsigned ut) {
    if constexpr (p < END) {
        spass_mid<C, E, PT, p>(v, buf, mt, ut);
        spass_mids<C, E, PT, p + 1, END>(v, buf, mt, ut);
    }
}

constexpr int inverse_mod(int a, int m) {
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 0;
}
constexpr int gcd_int(int a, int b) { return b == 0 ? a : gcd_int(b, a % b); }

// Whether the last two passes of PT (radices RA and RB coprime, the first of
// them on Ns = T, so that each thread holds one group of RA RB = E points)
// fuse into one radix-RA RB pass in registers.
template <typename PT, int E>
constexpr bool fused_last() {
    if (PT::npass < 3) return false;
    constexpr int pa = PT::npass - 2, pb = PT::npass - 1;
    return PT::ns(pa) == PT::T && PT::radix(pa) * PT::radix(pb) == E &&
           gcd_int(PT::radix(pa), PT::radix(pb)) == 1;
}

// The last two passes fused: the thread's group of R = RA RB points
// x[ut + q T] times exp(-2 pi i q ut / Lh), then the R-point DFT by the
// Good-Thomas mapping (input n = (RB n1 + RA n2) mod R, output
// k = (RB (RB^-1 mod RA) k1 + RA (RA^-1 mod RB) k2) mod R: RB DFTs of RA
// points, then RA DFTs of RB points, no twiddles between them).  Output k is
// bin ut + k T of the half; its power is accumulated in the slot the runtime
// plan's last pass would give it (radix RB, groups i = k mod RA, q = k / RA).
// Twiddle q = RB a + b is exp(-2 pi i RB a ut / Lh) exp(-2 pi i b ut / Lh):
// the first factor is pass RA's own twiddle a (kept in registers), the
// second pass RB's twiddle b at kk = ut (R - 1 ... RB - 1 loads).
template <typename C, int E, typename PT, int H, typename Acc>
__device__ __forceinline__ void spass_last_fused(C (&v)[E], Acc (&acc)[2][E], const C *buf,
                                                 const C *__restrict__ twp, const MidTwiddles<C, PT> &mt,
                                                 const double2 *__restrict__ w, double &ar, double &ai,
                                                 unsigned ut) {
    constexpr int pa = PT::npass - 2, pb = PT::npass - 1;
    constexpr int RA = PT::radix(pa), RB = PT::radix(pb), R = RA * RB;
    constexpr unsigned T = PT::T;
    static_assert(R == E && PT::Lh == T * R, "fused last passes of a fixed plan");
    C base[RB];
#pragma unroll
    for (int b = 1; b < RB; ++b) base[b] = twiddle_load(twp, ut * RB, PT::toff(pb) + b);
    const unsigned rb = PT::swz(pa - 1, ut);
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = buf[rb + q * T];
    __syncthreads();  // the buffer may now be overwritten by the next series
#pragma unroll
    for (int a = 0; a < RA; ++a)
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if (a == 0 && b == 0) continue;
            const C t = a == 0 ? base[b] : b == 0 ? mt.w[pa][a] : cmul(mt.w[pa][a], base[b]);
            v[RB * a + b] = cmul(v[RB * a + b], t);
        }
    C y[RA][RB];
#pragma unroll
    for (int n2 = 0; n2 < RB; ++n2) {
        C t[RA];
#pragma unroll
        for (int n1 = 0; n1 < RA; ++n1) t[n1] = v[(RB * n1 + RA * n2) % R];
        Dft<RA, C>::run(t);
#pragma unroll
        for (int k1 = 0; k1 < RA; ++k1) y[k1][n2] = t[k1];
    }
    constexpr int MA = RB * inverse_mod(RB % RA, RA), MB = RA * inverse_mod(RA % RB, RB);
#pragma unroll
    for (int k1 = 0; k1 < RA; ++k1) {
        C z[RB];
#pragma unroll
        for (int n2 = 0; n2 < RB; ++n2) z[n2] = y[k1][n2];
        Dft<RB, C>::run(z);
#pragma unroll
        for (int k2 = 0; k2 < RB; ++k2) {
            const int k = (MA * k1 + MB * k2) % R;
            const C x = z[k2];
            const typename Cx<C>::R pw = x.x * x.x + x.y * x.y;
            acc[H][(k % RA) * RB + k / RA] += (Acc)pw;
            const double2 wf = weight_load(w, ut, (unsigned)H * PT::Lh + (unsigned)k * T);
            ar = fma(wf.x, (double)pw, ar);
            ai = fma(wf.y, (double)pw, ai);
        }
    }
}

// One half H of a padded series through the fixed plan PT.
template <typename C, int E, typename PT, int H, typename Acc, typename Input>
__device__ __forceinline__ void fft_sweep_fixed(C (&v)[E], Acc (&acc)[2][E], C *buf, const C *__restrict__ tw,
                                                const MidTwiddles<C, PT> &mt, const double2 *__restrict__ w,
                                                const Input &in, double &ar, double &ai, unsigned ut) {
    static_assert(H == 0 || Input::kPretwiddled, "the odd-bin half needs its input pre-twiddled");
    spass_first<C, E, PT>(v, buf, in, ut);
    if constexpr (fused_last<PT, E>()) {
        spass_mids<C, E, PT, 1, PT::npass - 2>(v, buf, mt, ut);
        spass_last_fused<C, E, PT, H>(v, acc, buf, tw, mt, w, ar, ai, ut);
    } else {
        spass_mids<C, E, PT, 1, PT::npass - 1>(v, buf, mt, ut);
        spass_last<C, E, PT, H>(v, acc, buf, tw, w, ar, ai, ut);
    }
}

// The one-block half-lengths of the self path (the runtime plans of
// make_fft_plan, radix by radix): 256 = 4^4 on 64 threads, 512 = 8^3 on 64,
// 1,024 = 8^3 x 2 on 128, 1,536 = 8^3 x 3 on 64 (last two passes fused:
// radix 24), 2,048 = 8^3 x 4 on 256, 2,560 = 4^4 x 2 x 5 on 128, 3,072 =
// 8^3 x 2 x 3 on 128, 4,096 = 8^4 on 256, 5,120 = 4^5 x 5 on 256 (last two
// fused: radix 20), 6,144 = 8^3 x 4 x 3 on 256,
using Plan256 = StaticPlan<256, 64, 4, 4, 4, 4>;
using Plan512 = StaticPlan<512, 64, 8, 8, 8>;
using Plan1024 = StaticPlan<1024, 128, 8, 8, 8, 2>;
using Plan1536 = StaticPlan<1536, 64, 8, 8, 8, 3>;
using Plan2048 = StaticPlan<2048, 256, 8, 8, 8, 4>;
using Plan4096 = StaticPlan<4096, 256, 8, 8, 8, 8>;
using Plan5120 = StaticPlan<5120, 256, 4, 4, 4, 4, 4, 5>;
using Plan2560 = StaticPlan<2560, 128, 4, 4, 4, 4, 2, 5>;
using Plan3072 = StaticPlan<3072, 128, 8, 8, 8, 2, 3>;
using Plan6144 = StaticPlan<6144, 256, 8, 8, 8, 4, 3>;
// and on 512 threads (two workgroups per CU, 64 and 80 KB of LDS each):
// 8,192 = 8^4 x 2, 10,240 = 4^5 x 2 x 5
using Plan8192 = StaticPlan<8192, 512, 8, 8, 8, 8, 2>;
using Plan10240 = StaticPlan<10240, 512, 4, 4, 4, 4, 4, 2, 5>;

// |X_f|^2 of one padded series accumulated into acc; a += sum_f w_f |X_f|^2.
// tw: packed pass twiddles; pre: exp(-2 pi i n / L) for the odd bins.
template <typename C, int NH, int E, typename Acc, typename Input>
__device__ __forceinline__ void fft_power(const FftPlan &P, C (&v)[NH][E], Acc (&acc)[2][E], C *buf,
                                          const C *__restrict__ tw, const C *__restrict__ pre,
                                          const double2 *__restrict__ w, int NF, const Input &in, double &ar,
                                          double &ai, int tid) {
    if constexpr (NH == 2) {
        fft_sweep<C, 2, E>(P, 0, v, acc, buf, tw, pre, w, NF, in, ar, ai, tid);
    } else {
        fft_sweep<C, 1, E>(P, 0, v, acc, buf, tw, pre, w, NF, in, ar, ai, tid);
        fft_sweep<C, 1, E>(P, 1, v, acc, buf, tw, pre, w, NF, in, ar, ai, tid);
    }
}

// Sum of (x, y) over the block; the result is valid in thread 0.
__device__ __forceinline__ double2 block_sum2(double x, double y, double2 *red, int tid, int T) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        x += __shfl_xor(x, off, 64);
        y += __shfl_xor(y, off, 64);
    }
    if ((tid & 63) == 0) red[tid >> 6] = make_double2(x, y);
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (tid == 0)
        for (int k = 0; k < T / 64; ++k) {
            r.x += red[k].x;
            r.y += red[k].y;
        }
    __syncthreads();
    return r;
}

// Adds the register power spectrum to this block's FP64 row and clears it.
template <int E, typename Acc>
__device__ __forceinline__ void flush_power(const FftPlan &P, Acc (&acc)[2][E], double *row, bool first, int tid) {
    const int R = P.radix[P.npass - 1];
    const int stride = P.Lh / R;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = e / R, q = e % R;
        const int f = tid + i * P.T + q * stride;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const double add = (double)acc[h][e];
            double *dst = row + 2 * f + h;
            *dst = first ? add : *dst + add;
            acc[h][e] = Acc(0);
        }
    }
}

// As flush_power for the bins of half H only (the other half's row entries
// belong to another block).
template <int E, int H, typename Acc>
__device__ __forceinline__ void flush_power_half(const FftPlan &P, Acc (&acc)[2][E], double *row, bool first,
                                                 int tid) {
    const int R = P.radix[P.npass - 1];
    const int stride = P.Lh / R;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = e / R, q = e % R;
        const int f = tid + i * P.T + q * stride;
        const double add = (double)acc[H][e];
        double *dst = row + 2 * f + H;
        *dst = first ? add : *dst + add;
        acc[H][e] = Acc(0);
    }
}

// ------------------------------------------------------------ kernels -----

// FP32 power spectra are folded into the FP64 row every kFlushEvery series
// (bounding the FP32 accumulation to a few terms); FP64 ones only at the end.
constexpr int kFlushEvery = 16;

template <typename C>
struct AmpInput {
    static constexpr bool kPretwiddled = false;
    static constexpr bool kBounded = false;
    const float2 *p;
    size_t split_stride;
    int split;
    __device__ __forceinline__ C operator()(unsigned n) const {
        float2 s = p[n];
        for (int k = 1; k < split; ++k) {
            const float2 t = p[(size_t)k * split_stride + n];
            s.x += t.x;
            s.y += t.y;
        }
        return Cx<C>::make(s.x, s.y);
    }
};

template <typename C>
__device__ __forceinline__ const C *pass_twiddles(const CorrArgs &a);
template <>
__device__ __forceinline__ const float2 *pass_twiddles<float2>(const CorrArgs &a) {
    return a.twp;
}
template <>
__device__ __forceinline__ const double2 *pass_twiddles<double2>(const CorrArgs &a) {
    return a.twp64;
}
template <typename C>
__device__ __forceinline__ const C *pre_twiddles(const CorrArgs &a);
template <>
__device__ __forceinline__ const float2 *pre_twiddles<float2>(const CorrArgs &a) {
    return a.twL;
}
template <>
__device__ __forceinline__ const double2 *pre_twiddles<double2>(const CorrArgs &a) {
    return a.twL64;
}

// "all": one block per (shell, chunk of vectors); the series of vector v is
// its coherent amplitude A_v(t) from HBM (sum of the atom-range partials).
// One block of the coherent correlation: shell s, vector chunk c.
template <typename C, int NH, int E>
__device__ __forceinline__ void corr_all_block(const CorrArgs &a, int s, int c, double2 *lds_d) {
    using Acc = typename Cx<C>::R;
    C *buf = reinterpret_cast<C *>(lds_d);
    const FftPlan &P = a.plan;
    // The block reduction of a_m runs after the last pass has read the
    // buffer behind a barrier, and ends with a barrier before the next series
    // writes it, so it can use the start of the FFT buffer.
    double2 *red = reinterpret_cast<double2 *>(buf);
    const C *tw = pass_twiddles<C>(a);
    const C *pre = pre_twiddles<C>(a);
    const int tid = threadIdx.x;
    const int m0 = (int)((long)a.NM * c / a.chunks), m1 = (int)((long)a.NM * (c + 1) / a.chunks);

    C v[NH][E];
    Acc acc[2][E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[0][e] = acc[1][e] = Acc(0);
    double sa_r = 0.0, sa_i = 0.0, sa2 = 0.0;
    double *row = a.ppart + ((size_t)s * a.chunks + c) * P.L;
    const bool fp32 = sizeof(Acc) == 4;
    bool first = true;
    int since = 0;
    for (int m = m0; m < m1; ++m) {
        const size_t vidx = (size_t)s * a.NM + m;
        AmpInput<C> in{a.amp + vidx * a.NFp, a.split_stride, a.split};
        double ar = 0.0, ai = 0.0;
        // Opaque copy of the thread index: keeps the compiler from hoisting
        // every pass's address arithmetic out of the series loop (which
        // would not fit in registers).
        int t = tid;
        asm volatile("" : "+v"(t));
        fft_power<C, NH, E>(P, v, acc, buf, tw, pre, a.w, a.NF, in, ar, ai, t);
        const double2 r = block_sum2(ar, ai, red, tid, P.T);
        if (tid == 0) {
            sa_r += r.x;
            sa_i += r.y;
            sa2 += r.x * r.x + r.y * r.y;
        }
        if (fp32 && ++since == kFlushEvery) {
            flush_power<E>(P, acc, row, first, tid);
            first = false;
            since = 0;
        }
    }
    if (since > 0 || first || !fp32) flush_power<E>(P, acc, row, first, tid);
    if (tid == 0) {
        double *ap = a.apart + ((size_t)s * a.chunks + c) * 4;
        ap[0] = sa_r;
        ap[1] = sa_i;
        ap[2] = sa2;
        ap[3] = 0.0;
    }
}

// "self": the series of pair (atom j, vector m) is b_j exp(i q_m . r_j(t)),
// evaluated on the fly in the first FFT pass from the SoA coordinates
// (self_vectors_scatter_device.cpp:306-319).  The rows are read through
// buffer resources built once per pair and bounded at NF frames, so a read
// past the series returns 0 and the padding needs no branch (b is zeroed by
// a select): the first pass can issue the loads of all its points together,
// where a branch per point waited for each point's loads in turn.
struct SelfRows {
    __amdgpu_buffer_rsrc_t rx, ry, rz;
    unsigned nf;
    __device__ __forceinline__ void set(const float *X, size_t plane, int NF) {
        nf = (unsigned)NF;
        rx = __builtin_amdgcn_make_buffer_rsrc((void *)X, (short)0, nf * 4u, kBufferDword3);
        ry = __builtin_amdgcn_make_buffer_rsrc((void *)(X + plane), (short)0, nf * 4u, kBufferDword3);
        rz = __builtin_amdgcn_make_buffer_rsrc((void *)(X + 2 * plane), (short)0, nf * 4u, kBufferDword3);
    }
    __device__ __forceinline__ void load(unsigned n, float &x, float &y, float &z) const {
        x = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rx, n * 4u, 0, 0));
        y = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ry, n * 4u, 0, 0));
        z = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rz, n * 4u, 0, 0));
    }
};

struct SelfInput : SelfRows {
    static constexpr bool kPretwiddled = false;
    static constexpr bool kBounded = true;  // 0 past NF without a branch
    static constexpr bool kSplitLoad = true;  // load() then at()
    double qx, qy, qz;
    float b;
    // the sample at frame n from its loaded coordinates
    __device__ __forceinline__ float2 at(unsigned n, float xs, float ys, float zs) const {
        double p = qx * (double)xs;
        p = fma(qy, (double)ys, p);
        p = fma(qz, (double)zs, p);
        const float f = (float)__builtin_amdgcn_fract(p);
        const float bn = n < nf ? b : 0.f;
        return make_float2(bn * __builtin_amdgcn_cosf(f), bn * __builtin_amdgcn_sinf(f));
    }
    __device__ __forceinline__ float2 operator()(unsigned n) const {
        float xs, ys, zs;
        load(n, xs, ys, zs);
        return at(n, xs, ys, zs);
    }
};

// The odd-bin input x_n W_L^n of a self series with the pre-twiddle folded
// into the phase: b exp(2 pi i (q.r_n - n/L)), reduced in FP64.
struct SelfInputShifted : SelfRows {
    static constexpr bool kPretwiddled = true;
    static constexpr bool kBounded = true;
    static constexpr bool kSplitLoad = true;
    double qx, qy, qz, invL;
    float b;
    __device__ __forceinline__ float2 at(unsigned n, float xs, float ys, float zs) const {
        double p = fma(-(double)n, invL, qx * (double)xs);
        p = fma(qy, (double)ys, p);
        p = fma(qz, (double)zs, p);
        const float f = (float)__builtin_amdgcn_fract(p);
        const float bn = n < nf ? b : 0.f;
        return make_float2(bn * __builtin_amdgcn_cosf(f), bn * __builtin_amdgcn_sinf(f));
    }
    __device__ __forceinline__ float2 operator()(unsigned n) const {
        float xs, ys, zs;
        load(n, xs, ys, zs);
        return at(n, xs, ys, zs);
    }
};

// One block of the shell back-transform: shell s, frame tile `tile` (256
// frames); Psh holds final_lds_bytes(L) of LDS.  fqt(tau) = R(tau) / (L (NF -
// tau) NM) with R = sum_f P_f e^{+2 pi i f tau / L} in FP64, P the sum of the
// chunk partials, staged through LDS kFinalChunk bins at a time; tile 0 also
// writes fq = sum a / NM and fq2 = sum |a|^2 / NM.
__device__ __forceinline__ void final_block(const FinalArgs &a, int s, int tile, double *Psh) {
    const int tid = threadIdx.x;
    if (a.ppart != nullptr) {
        const double *base = a.ppart + (size_t)s * a.chunks * a.L;
        const int tau = tile * 256 + tid;
        const bool live = tau < a.NF;
        // e^{+2 pi i f tau / L} by repeated products with e^{+2 pi i tau / L},
        // re-anchored on the FP64 table every 64 bins
        const long L = a.L;
        const double2 st = a.tw64[live ? tau : 0];
        long anchor = 0;  // (f * tau) mod L at the anchors
        const long jump = (64L * (live ? tau : 0)) % L;
        double rr0 = 0.0, ri0 = 0.0, rr1 = 0.0, ri1 = 0.0;
        for (int fc = 0; fc < a.L; fc += kFinalChunk) {
            const int nc = min(kFinalChunk, a.L - fc);
            __syncthreads();  // every thread is done with the previous chunk
            for (int f = tid; f < nc; f += 256) {
                double acc = 0.0;
                for (int c = 0; c < a.chunks; ++c) acc += base[(size_t)c * a.L + fc + f];
                Psh[f] = acc;
            }
            __syncthreads();
            if (live) {
                for (int f0 = 0; f0 < nc; f0 += 64) {
                    double2 z = a.tw64[anchor];
                    anchor += jump;
                    if (anchor >= L) anchor -= L;
                    const int f1 = min(f0 + 64, nc);
                    int f = f0;
                    for (; f + 1 < f1; f += 2) {
                        const double p0 = Psh[f], p1 = Psh[f + 1];
                        rr0 = fma(p0, z.x, rr0);
                        ri0 = fma(p0, z.y, ri0);
                        double zr = z.x * st.x - z.y * st.y, zi = z.x * st.y + z.y * st.x;
                        rr1 = fma(p1, zr, rr1);
                        ri1 = fma(p1, zi, ri1);
                        z.x = zr * st.x - zi * st.y;
                        z.y = zr * st.y + zi * st.x;
                    }
                    if (f < f1) {
                        rr0 = fma(Psh[f], z.x, rr0);
                        ri0 = fma(Psh[f], z.y, ri0);
                    }
                }
            }
        }
        if (live) {
            const double norm = a.norm != 0.0 ? a.norm : 1.0 / (double)a.NM;
            double scale = norm / ((double)a.L * (double)(a.NF - tau));
            double re = (rr0 + rr1) * scale, im = (ri0 + ri1) * scale;
            if (a.tail && tau >= a.NF - a.tail_k) {
                // the last lags from the direct FP64 sums
                const double2 *tp = a.tail + (size_t)s * a.tail_blocks * kTailLags + (a.NF - 1 - tau);
                double tr = 0.0, ti = 0.0;
                for (int bb = 0; bb < a.tail_blocks; ++bb) {
                    tr += tp[(size_t)bb * kTailLags].x;
                    ti += tp[(size_t)bb * kTailLags].y;
                }
                scale = norm / (double)(a.NF - tau);
                re = tr * scale;
                im = ti * scale;
            }
            a.fqt[((size_t)s * a.NF + tau) * 2] = re;
            a.fqt[((size_t)s * a.NF + tau) * 2 + 1] = a.conj ? -im : im;
        }
    }
    if (tile == 0 && tid == 0) {
        double sr = 0.0, si = 0.0, s2 = 0.0;
        const int na = a.achunks > 0 ? a.achunks : a.chunks;
        const double *ap = a.apart + (size_t)s * na * 4;
        for (int c = 0; c < na; ++c) {
            sr += ap[4 * c];
            si += ap[4 * c + 1];
            s2 += ap[4 * c + 2];
        }
        const double factor = a.norm != 0.0 ? a.norm : 1.0 / (double)a.NM;
        a.fq[2 * s] = sr * factor;
        a.fq[2 * s + 1] = (a.conj ? -si : si) * factor;
        a.fq2[2 * s] = s2 * factor;
        a.fq2[2 * s + 1] = 0.0;
    }
}

// ------------------------------------------------- series of N1 > 1 rows ---

// Input of row k1 of half h of a series x (plan with N1 > 1 rows of Lh
// points, M = N1 Lh, L = 2 M):
//   in(n2) = W_M^{n2 k1} sum_{n1 < N1} y(n2 + Lh n1) W_{N1}^{n1 k1},
//   y(n) = x(n) W_L^{h n} for n < NF, else 0,
// the column DFT and twiddle of the four-step factorisation
// Y[k1 + N1 k2] = sum_{n2} W_Lh^{n2 k2} in(n2).  tw: [L] exp(-2 pi i e / L),
// so W_{N1} = W_L^{2 Lh} and W_M = W_L^2.  Src(n) returns y(n) (its
// pre-twiddle applied).
template <typename C, typename Src>
struct RowInput {
    static constexpr bool kPretwiddled = true;
    static constexpr bool kBounded = false;
    static constexpr bool kSplitLoad = false;
    Src src;
    const C *tw;
    unsigned L, NF, Lh;
    int N1, k1;
    __device__ __forceinline__ C operator()(unsigned n2) const {
        C acc = Cx<C>::make(0, 0);
        for (int n1 = 0; n1 < N1; ++n1) {
            const unsigned n = n2 + Lh * (unsigned)n1;
            if (n >= NF) break;
            const C y = src(n);
            const int e = (n1 * k1) % N1;
            acc = cadd(acc, e == 0 ? y : cmul(y, tw[2u * Lh * (unsigned)e]));
        }
        return k1 == 0 ? acc : cmul(acc, tw[(2u * n2 * (unsigned)k1) % L]);
    }
};

// Row source of a coherent series: the sum of the atom-range partial
// amplitudes at frame n, times W_L^{h n}.
template <typename C>
struct AmpRowSource {
    const float2 *p;
    size_t split_stride;
    int split, h;
    const C *tw;
    __device__ __forceinline__ C operator()(unsigned n) const {
        float2 s = p[n];
        for (int k = 1; k < split; ++k) {
            const float2 t = p[(size_t)k * split_stride + n];
            s.x += t.x;
            s.y += t.y;
        }
        const C x = Cx<C>::make(s.x, s.y);
        return h ? cmul(x, tw[n]) : x;
    }
};

// Row source of a self series: b exp(2 pi i (q.r_n - h n / L)), the phase
// and its pre-twiddle formed and reduced in FP64.
struct SelfRowSource {
    const float *X, *Y, *Z;
    double qx, qy, qz, hinvL;
    float b;
    __device__ __forceinline__ float2 operator()(unsigned n) const {
        double p = fma(-(double)n, hinvL, qx * (double)X[n]);
        p = fma(qy, (double)Y[n], p);
        p = fma(qz, (double)Z[n], p);
        const float f = (float)__builtin_amdgcn_fract(p);
        return make_float2(b * __builtin_amdgcn_cosf(f), b * __builtin_amdgcn_sinf(f));
    }
};

// As flush_power_half for row k1 of half h of an N1-row plan: bin k2 of the
// row is bin 2 (k1 + N1 k2) + h of the power-spectrum row.
template <int E, typename Acc>
__device__ __forceinline__ void flush_power_row(const FftPlan &P, Acc (&acc)[2][E], double *row, bool first, int tid,
                                                int h, int k1) {
    const int R = P.radix[P.npass - 1];
    const int stride = P.Lh / R;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = e / R, q = e % R;
        const int k2 = tid + i * P.T + q * stride;
        const double add = (double)acc[0][e];
        double *dst = row + 2 * ((size_t)k1 + (size_t)P.N1 * k2) + h;
        *dst = first ? add : *dst + add;
        acc[0][e] = Acc(0);
    }
}

}  // namespace sx
