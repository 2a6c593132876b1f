// sx_batch.hip -- many uneven coherent trajectories in one captured HIP graph.
//
// Per trajectory the arithmetic is that of sx_compute_shells_oriented.  The
// amplitude stage is one persistent kernel: the (trajectory, frame tile,
// orientation group, shell group) tiles of every trajectory are sorted
// largest first and cut into 8 contiguous, cost-balanced runs, one per XCD
// (workgroup b starts on XCD b mod 8); each workgroup pulls tiles from its
// XCD's run through a global counter until the run is drained, so every
// wavefront exits.  The correlation stage launches once per FFT length class
// over the (trajectory, shell, chunk) blocks of that class, and one
// finalisation launch covers every (trajectory, shell, frame tile).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <map>
#include <cstdio>
#include <memory>
#include <numeric>
#include <vector>

#include "sx_amp_device.h"
#include "sx_ctx.h"
#include "sx_fft_device.h"

namespace sx {

// KIND 0: tiles of trajectories whose shells take one phase each or an
// exact progression (modes 0 and 1); KIND 1: tiles of trajectories stepped
// with residual corrections on their tile lines (mode 2: both walks
// corrected, also where one of them needs none), KIND 2: such tiles whose
// shells all lie on their line (mode 3); KINDs 1 and 2 keep their FP64
// totals in the dynamic LDS (kResidualLds<VM, VS>), each wavefront touching
// only its own slice, so tiles need no barrier between them.  The kinds run
// as persistent launches with a work queue each, so that no kernel carries
// another's body (one kernel with the bodies of modes 2-5 ran the
// 64-trajectory batch 19 % slower).
// The LDS word through which thread 0 hands the next tile index to its
// workgroup.  The residual kernel's dynamic LDS fills the CU at two
// workgroups (80 KiB for 1 x 20), and a static word on top would leave room
// for one, so there the index passes through the first word of wave 0's
// slice of the totals: thread 0 writes it only after its own wave has
// finished the previous tile, and the other waves never touch that slice.
template <int KIND>
__device__ inline int *tile_slot(double *dyn) {
    if constexpr (KIND > 0) {
        return reinterpret_cast<int *>(dyn);
    } else {
        __shared__ int cur;
        return &cur;
    }
}

template <int VM, int VS, int KIND>
__global__ __launch_bounds__(256, 2) void amp_persistent_kernel(const OrientKernelArgs *__restrict__ targs,
                                                             const int4 *__restrict__ items,
                                                             const int *__restrict__ qstart,
                                                             unsigned *__restrict__ counters) {
    extern __shared__ double tot_lds[];
    int *cur = tile_slot<KIND>(tot_lds);
    const int x = blockIdx.x & 7;
    const int begin = qstart[x], n = qstart[x + 1] - begin;
    for (;;) {
        if (threadIdx.x == 0) *cur = (int)atomicAdd(&counters[x], 1u);
        __syncthreads();
        const int it = *cur;
        __syncthreads();  // every thread has read cur before it is overwritten
        if (it >= n) break;
        const int4 w = items[begin + it];
        if constexpr (KIND == 1)
            oriented_tile_body<VM, VS, false, 2, true>(targs[w.x], w.y, 0, w.z & 0xffff, w.z >> 16, tot_lds);
        else if constexpr (KIND == 2)
            oriented_tile_body<VM, VS, false, 3, true>(targs[w.x], w.y, 0, w.z & 0xffff, w.z >> 16, tot_lds);
        else
            oriented_tile<VM, VS, false, false>(targs[w.x], w.y, 0, w.z & 0xffff, w.z >> 16);
    }
}

__global__ void reset_counters_kernel(unsigned *c, int n) {
    const int i = threadIdx.x;
    if (i < n) c[i] = 0u;
}

template <typename C, int NH, int E>
__global__ __launch_bounds__(512) void corr_batch_kernel(const CorrArgs *__restrict__ cargs,
                                                         const int4 *__restrict__ items) {
    extern __shared__ double2 lds_d[];
    const int4 w = items[blockIdx.x];
    corr_all_block<C, NH, E>(cargs[w.x], w.y, w.z, lds_d);
}

__global__ __launch_bounds__(256) void final_batch_kernel(const FinalArgs *__restrict__ fargs,
                                                          const int4 *__restrict__ items) {
    extern __shared__ double Psh[];
    const int4 w = items[blockIdx.x];
    final_block(fargs[w.x], w.y, w.z, Psh);
}

using PersistentFn = void (*)(const OrientKernelArgs *, const int4 *, const int *, unsigned *);
using CorrBatchFn = void (*)(const CorrArgs *, const int4 *);

// Plans with mixed tiles: each trajectory's shells are covered by the
// cheapest mix of 1 x 8, 1 x 16 and 1 x 20 tiles (fixed shapes 2, 3 and 4,
// by mixed code 0, 1, 2), whose queues run one after the other.
constexpr int kMixedShape = kOrientShapes;
constexpr int kMixedVS[3] = {8, 16, 20};
constexpr int kMixedFixed[3] = {2, 3, 4};

static size_t persistent_res_lds(int shape) {
    switch (shape) {
        case 0: return kResidualLds<2, 8>;
        case 1: return kResidualLds<4, 4>;
        case 2: return kResidualLds<1, 8>;
        case 3: return kResidualLds<1, 16>;
        default: return kResidualLds<1, 20>;
    }
}

template <int KIND>
static PersistentFn persistent_fn_of(int shape) {
    switch (shape) {
        case 0: return amp_persistent_kernel<2, 8, KIND>;
        case 1: return amp_persistent_kernel<4, 4, KIND>;
        case 2: return amp_persistent_kernel<1, 8, KIND>;
        case 3: return amp_persistent_kernel<1, 16, KIND>;
        default: return amp_persistent_kernel<1, 20, KIND>;
    }
}

constexpr int kQueueKinds = 3;
static PersistentFn persistent_fn(int shape, int kind) {
    return kind == 2 ? persistent_fn_of<2>(shape) : kind == 1 ? persistent_fn_of<1>(shape) : persistent_fn_of<0>(shape);
}

static CorrBatchFn corr_batch_fn(int E) {
    switch (E) {
        case 4: return corr_batch_kernel<double2, 1, 4>;
        case 8: return corr_batch_kernel<double2, 1, 8>;
        case 16: return corr_batch_kernel<double2, 1, 16>;
        case 20: return corr_batch_kernel<double2, 1, 20>;
        case 24: return corr_batch_kernel<double2, 1, 24>;
        default: return nullptr;
    }
}

}  // namespace sx

using namespace sxh;

namespace {

constexpr int kXcds = 8;

struct TrajWork {
    sx_ctx *ctx = nullptr;  // staged coordinates, factors and FFT tables
    DevBuf qlen, u, qres, amp, ppart, apart, fqt, fq, fq2;
    int mode = 0;          // shell stepping mode (sx::shell_progression)
    sx::ShellGrid grid;              // the shells' grid and exact runs (mode 2)
    std::vector<double> qturns;      // and their |q| / 2 pi
    double qstep = 0.0;    // step and origin of the shells in |q| / 2 pi
    double q0 = 0.0;
    int chunks = 1;
};

// One persistent amplitude launch: its tiles in 8 per-XCD runs.
struct Queue {
    DevBuf items, qstart, counters;
    int nitems = 0, persistent_blocks = 0;
    int kind = 0;  // 0 no residuals, 1 residual-corrected tiles, 2 tiles on their line
    int shape = 0;     // fixed tile shape of the kernel
};

// Device-side launch plan of a set of trajectories.
struct Plan {
    std::vector<int> ks;
    DevBuf oka, cargs, citems, fargs, fitems;
    // persistent launches in order: tiles of modes 0 / 1, then of mode 2,
    // each by tile shape (largest first)
    std::vector<Queue> queues;
    int nitems = 0, nfitems = 0;
    size_t terms_exact_runs = 0;  // terms of mode-2 tiles stepped without corrections (kind 2)
    size_t final_lds = 0;
    struct CorrGroup {
        int E, T, first, count;
        size_t lds;
    };
    std::vector<CorrGroup> groups;
    hipGraphExec_t exec = nullptr;
    ~Plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (DevBuf *d : {&oka, &cargs, &citems, &fargs, &fitems}) d->release();
        for (Queue &q : queues)
            for (DevBuf *d : {&q.items, &q.qstart, &q.counters}) d->release();
    }
};

}  // namespace

struct sx_batch {
    sx_config cfg{};
    std::string err;
    hipStream_t stream = nullptr;
    std::vector<sx_trajectory> trajs;
    std::vector<TrajWork> work;
    int shape = 0;
    // persistent workgroups that fit the GPU at once, per kernel (plain,
    // residual) and tile shape
    int max_resident[sx::kQueueKinds][sx::kOrientShapes] = {};
    std::unique_ptr<Plan> all;
    std::vector<std::unique_ptr<Plan>> one;
    sx_batch_timing timing{};
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

int bfail(sx_batch *b, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

#define SXB_HIP(b, expr)                                                                                  \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return bfail((b), e_ == hipErrorOutOfMemory ? SX_ERR_NOMEM : SX_ERR_HIP, "%s failed: %s (%s:%d)", \
                         #expr, hipGetErrorString(e_), __FILE__, __LINE__);                               \
    } while (0)

sx::OrientKernelArgs orient_args(const TrajWork &w, const sx_trajectory &t, int shape) {
    sx::OrientArgs oa;
    oa.coords = w.ctx->coords.as<float>();
    oa.NAp = w.ctx->NAp;
    oa.NFp = w.ctx->NFp;
    oa.qlen = w.qlen.as<double>();
    oa.u = w.u.as<double>();
    oa.b = w.ctx->b.as<float>();
    oa.NM = (int)t.NM;
    oa.ns = (int)t.nshells;
    oa.split = 1;
    oa.per_shell_b = 0;
    oa.shape = shape;
    oa.amp = w.amp.as<float2>();
    oa.ap = w.mode;
    oa.qstep = w.qstep;
    oa.q0 = w.q0;
    // mode 2: [ns] tile lines, then [ns] residuals (build_plan)
    oa.lines = w.mode == 2 ? w.qres.as<sx::TileLine>() : nullptr;
    oa.qres = w.mode == 2 ? reinterpret_cast<const float *>(oa.lines + t.nshells) : nullptr;
    return sx::make_orient_args(oa);
}

// Cost of one shell slot (one shell of a tile, for one orientation, atom and
// frame) of the mixed plans' 1 x 8, 1 x 16 and 1 x 20 tiles relative to
// 1 x 20, for exactly stepped tiles (modes 0 and 1) and residual-corrected
// ones (mode 2); measured on an MI355X through the persistent kernels, four
// trajectories of 20,000 atoms x 2,000 frames and 200 orientations with 8,
// 16 and 20 shells (residual: 2.137, 1.482 and 1.378 ms per shell; exact:
// 1.731, 1.193 and 1.073 ms).
constexpr double kSlotCost[2][3] = {{1.614, 1.112, 1.0}, {1.551, 1.076, 1.0}};

struct ShellGroup {
    int first, code;  // first shell; tile code (kMixedVS[code] shells)
};

// The cheapest cover of ns shells by mixed tiles, by dynamic programming
// over the shells left: best[n] = min over codes g of kSlotCost[g] * VS_g +
// best[max(0, n - VS_g)].  Groups run from shell 0; only the last one can be
// partial.
std::vector<ShellGroup> shell_groups(int ns, bool res) {
    const double *c = kSlotCost[res ? 1 : 0];
    std::vector<double> best(ns + 1, 0.0);
    std::vector<int> pick(ns + 1, 0);
    for (int n = 1; n <= ns; ++n) {
        best[n] = 1e300;
        for (int g = 0; g < 3; ++g) {
            const double v = c[g] * sx::kMixedVS[g] + best[std::max(0, n - sx::kMixedVS[g])];
            if (v < best[n]) {
                best[n] = v;
                pick[n] = g;
            }
        }
    }
    std::vector<ShellGroup> out;
    int first = 0;
    for (int n = ns; n > 0; n -= sx::kMixedVS[pick[n]]) {
        out.push_back({first, pick[n]});
        first += sx::kMixedVS[pick[n]];
    }
    return out;
}

// Builds the device work lists of trajectories ks for tile shape `shape`
// (a fixed shape, or sx::kMixedShape).
int build_plan(sx_batch *b, const std::vector<int> &ks, int shape, Plan &P) {
    P.ks = ks;
    const bool mixed = shape == sx::kMixedShape;
    // mixed tiles have one orientation per wavefront, as 1 x 20
    const int args_shape = mixed ? 4 : shape;
    const int VS = sx::orient_shape_vs(args_shape);
    std::vector<sx::OrientKernelArgs> oka(b->trajs.size());
    P.terms_exact_runs = 0;
    struct Item {
        int4 w;
        long cost;
    };
    // [queue kind][kernel shape]
    std::vector<Item> all_items[sx::kQueueKinds][sx::kOrientShapes];
    for (int k : ks) {
        const sx_trajectory &t = b->trajs[k];
        const bool res = b->work[k].mode == 2;
        oka[k] = orient_args(b->work[k], t, args_shape);
        const int ftiles = oka[k].ftiles, og = oka[k].ogroups;
        std::vector<ShellGroup> groups;
        if (mixed) {
            groups = shell_groups((int)t.nshells, res);
        } else {
            for (int s = 0; s < oka[k].sgroups; ++s) groups.push_back({s * VS, 0});
        }
        std::vector<sx::TileLine> lines;
        if (res) {
            // the lines of this trajectory's tiles and the residuals against them
            TrajWork &w = b->work[k];
            const int ns = (int)t.nshells;
            std::vector<int> firsts, vs;
            for (const ShellGroup &sgp : groups) {
                firsts.push_back(sgp.first);
                vs.push_back(mixed ? sx::kMixedVS[sgp.code] : VS);
            }
            lines.resize(ns);
            std::vector<float> rr(ns, 0.f);
            sx::tile_lines(w.qturns.data(), ns, w.ctx->rmax, w.grid, firsts.data(), vs.data(), (int)firsts.size(),
                           lines.data(), rr.data());
            const size_t lb = sizeof(sx::TileLine) * ns;
            SXB_HIP(b, hipMemcpy(w.qres.p, lines.data(), lb, hipMemcpyHostToDevice));
            SXB_HIP(b, hipMemcpy(static_cast<char *>(w.qres.p) + lb, rr.data(), sizeof(float) * ns,
                                 hipMemcpyHostToDevice));
        }
        // tile cost in atom shell slots (x 1000 in kSlotCost units)
        for (size_t g = 0; g < groups.size(); ++g) {
            const int code = groups[g].code;
            const long cost = mixed ? (long)(oka[k].atoms_per_split * sx::kMixedVS[code] *
                                             kSlotCost[res ? 1 : 0][code] * 1000.0)
                                    : (long)oka[k].atoms_per_split;
            const int first = groups[g].first;
            const bool plain = res && lines[first].flags == 0;
            if (plain)
                P.terms_exact_runs += (size_t)std::min(mixed ? sx::kMixedVS[code] : VS, (int)t.nshells - first) *
                                      t.NM * t.NA * t.NF;
            std::vector<Item> &items = all_items[plain ? 2 : res ? 1 : 0][mixed ? sx::kMixedFixed[code] : shape];
            for (int ft = 0; ft < ftiles; ++ft)
                for (int o = 0; o < og; ++o) items.push_back({make_int4(k, ft, o | (groups[g].first << 16), 0), cost});
        }
    }
    P.nitems = 0;
    P.queues.clear();
    static const int kLaunchOrder[sx::kOrientShapes] = {4, 3, 0, 1, 2};  // 1x20, 1x16, 2x8, 4x4, 1x8
    for (int qi = 0; qi < sx::kQueueKinds; ++qi)
        for (int sh : kLaunchOrder) {
            if (all_items[qi][sh].empty()) continue;
            P.queues.emplace_back();
            P.queues.back().kind = qi;
            P.queues.back().shape = sh;
        }
    for (Queue &Q : P.queues) {
        const int qi = Q.kind;
        std::vector<Item> &items = all_items[qi][Q.shape];
        // largest tiles first; within a trajectory, tiles sharing a frame tile stay adjacent
        std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &c) { return a.cost > c.cost; });
        // 8 contiguous runs of equal cost, one per XCD
        long total = 0;
        for (const Item &it : items) total += it.cost;
        std::vector<int> qstart(kXcds + 1, (int)items.size());
        qstart[0] = 0;
        long acc = 0;
        int q = 1;
        for (size_t i = 0; i < items.size() && q < kXcds; ++i) {
            acc += items[i].cost;
            while (q < kXcds && acc * kXcds >= total * q) qstart[q++] = (int)i + 1;
        }
        std::vector<int4> w(items.size());
        for (size_t i = 0; i < items.size(); ++i) w[i] = items[i].w;
        Q.nitems = (int)items.size();
        P.nitems += Q.nitems;
        SXB_HIP(b, Q.items.ensure(sizeof(int4) * std::max<size_t>(1, w.size())));
        if (!w.empty()) SXB_HIP(b, hipMemcpy(Q.items.p, w.data(), sizeof(int4) * w.size(), hipMemcpyHostToDevice));
        SXB_HIP(b, Q.qstart.ensure(sizeof(int) * qstart.size()));
        SXB_HIP(b, hipMemcpy(Q.qstart.p, qstart.data(), sizeof(int) * qstart.size(), hipMemcpyHostToDevice));
        SXB_HIP(b, Q.counters.ensure(sizeof(unsigned) * kXcds));
        Q.persistent_blocks =
            std::max(kXcds, std::min(b->max_resident[qi][Q.shape], (Q.nitems + kXcds - 1) / kXcds * kXcds));
        Q.persistent_blocks = (Q.persistent_blocks + kXcds - 1) / kXcds * kXcds;
    }
    SXB_HIP(b, P.oka.ensure(sizeof(sx::OrientKernelArgs) * oka.size()));
    SXB_HIP(b, hipMemcpy(P.oka.p, oka.data(), sizeof(sx::OrientKernelArgs) * oka.size(), hipMemcpyHostToDevice));

    // correlation blocks grouped by FFT length; finalisation blocks
    std::vector<sx::CorrArgs> ca(b->trajs.size());
    std::vector<sx::FinalArgs> fa(b->trajs.size());
    std::map<int, std::vector<int4>> by_len;
    std::vector<int4> fitems;
    size_t final_lds = 8;
    for (int k : ks) {
        const sx_trajectory &t = b->trajs[k];
        TrajWork &tw = b->work[k];
        sx_ctx *c = tw.ctx;
        sx::CorrArgs &a = ca[k];
        a.plan = c->plan;
        a.NF = (int)t.NF;
        a.NM = (int)t.NM;
        a.nshells = (int)t.nshells;
        a.chunks = tw.chunks;
        a.twL = c->twL.as<float2>();
        a.twL64 = c->twL64.as<double2>();
        a.twp = c->twp.as<float2>();
        a.twp64 = c->twp64.as<double2>();
        a.w = c->w.as<double2>();
        a.ppart = tw.ppart.as<double>();
        a.apart = tw.apart.as<double>();
        a.amp = tw.amp.as<float2>();
        a.NFp = c->NFp;
        a.split = 1;
        a.split_stride = t.nshells * t.NM * c->NFp;
        for (int s = 0; s < (int)t.nshells; ++s)
            for (int ch = 0; ch < tw.chunks; ++ch) by_len[c->plan.Lh].push_back(make_int4(k, s, ch, 0));
        sx::FinalArgs &f = fa[k];
        f.L = c->plan.L;
        f.NF = (int)t.NF;
        f.NM = (int)t.NM;
        f.nshells = (int)t.nshells;
        f.chunks = tw.chunks;
        f.ppart = tw.ppart.as<double>();
        f.apart = tw.apart.as<double>();
        f.tw64 = c->tw64.as<double2>();
        f.conj = b->cfg.dsp_method == SX_METHOD_DIRECT ? 1 : 0;
        f.fqt = tw.fqt.as<double>();
        f.fq = tw.fq.as<double>();
        f.fq2 = tw.fq2.as<double>();
        const int tiles = (int)((t.NF + 255) / 256);
        for (int s = 0; s < (int)t.nshells; ++s)
            for (int ti = 0; ti < tiles; ++ti) fitems.push_back(make_int4(k, s, ti, 0));
        final_lds = std::max(final_lds, sx::final_lds_bytes(c->plan.L));
    }
    std::vector<int4> citems;
    P.groups.clear();
    for (auto &kv : by_len) {
        const sx::FftPlan &pl = b->work[kv.second[0].x].ctx->plan;
        // corr_all_block reduces a_m in the start of the FFT buffer, so the
        // dynamic LDS is exactly the transform data (160 KB at Lh = 10240)
        Plan::CorrGroup g{pl.E, pl.T, (int)citems.size(), (int)kv.second.size(), sx::corr_all_lds_bytes(pl)};
        P.groups.push_back(g);
        citems.insert(citems.end(), kv.second.begin(), kv.second.end());
    }
    SXB_HIP(b, P.cargs.ensure(sizeof(sx::CorrArgs) * ca.size()));
    SXB_HIP(b, hipMemcpy(P.cargs.p, ca.data(), sizeof(sx::CorrArgs) * ca.size(), hipMemcpyHostToDevice));
    SXB_HIP(b, P.citems.ensure(sizeof(int4) * std::max<size_t>(1, citems.size())));
    if (!citems.empty())
        SXB_HIP(b, hipMemcpy(P.citems.p, citems.data(), sizeof(int4) * citems.size(), hipMemcpyHostToDevice));
    SXB_HIP(b, P.fargs.ensure(sizeof(sx::FinalArgs) * fa.size()));
    SXB_HIP(b, hipMemcpy(P.fargs.p, fa.data(), sizeof(sx::FinalArgs) * fa.size(), hipMemcpyHostToDevice));
    SXB_HIP(b, P.fitems.ensure(sizeof(int4) * std::max<size_t>(1, fitems.size())));
    if (!fitems.empty())
        SXB_HIP(b, hipMemcpy(P.fitems.p, fitems.data(), sizeof(int4) * fitems.size(), hipMemcpyHostToDevice));
    P.nfitems = (int)fitems.size();
    P.final_lds = final_lds;
    return SX_OK;
}

int launch_amplitude(sx_batch *b, Plan &P) {
    for (Queue &Q : P.queues) {
        if (Q.nitems == 0) continue;
        const int shape = Q.shape;
        hipLaunchKernelGGL(sx::reset_counters_kernel, dim3(1), dim3(kXcds), 0, b->stream, Q.counters.as<unsigned>(),
                           kXcds);
        SXB_HIP(b, hipGetLastError());
        // (the dynamic-LDS attribute of the residual kernels is set once, in sx_batch_stage)
        const size_t lds = Q.kind ? sx::persistent_res_lds(shape) : 0;
        hipLaunchKernelGGL(sx::persistent_fn(shape, Q.kind), dim3(Q.persistent_blocks), dim3(256), lds, b->stream,
                           P.oka.as<sx::OrientKernelArgs>(), Q.items.as<int4>(), Q.qstart.as<int>(),
                           Q.counters.as<unsigned>());
        SXB_HIP(b, hipGetLastError());
    }
    return SX_OK;
}

int launch_rest(sx_batch *b, Plan &P) {
    for (const Plan::CorrGroup &g : P.groups) {
        sx::CorrBatchFn fn = sx::corr_batch_fn(g.E);
        if (!fn) return bfail(b, SX_ERR_UNSUPPORTED, "no correlation kernel for E=%d", g.E);
        hipLaunchKernelGGL(fn, dim3(g.count), dim3(g.T), g.lds, b->stream, P.cargs.as<sx::CorrArgs>(),
                           P.citems.as<int4>() + g.first);
        SXB_HIP(b, hipGetLastError());
    }
    hipLaunchKernelGGL(sx::final_batch_kernel, dim3(P.nfitems), dim3(256), P.final_lds, b->stream,
                       P.fargs.as<sx::FinalArgs>(), P.fitems.as<int4>());
    SXB_HIP(b, hipGetLastError());
    return SX_OK;
}

int capture(sx_batch *b, Plan &P) {
    // kernel attributes are set before capture (dynamic LDS above 64 KB)
    for (const Plan::CorrGroup &g : P.groups)
        SXB_HIP(b, hipFuncSetAttribute((const void *)sx::corr_batch_fn(g.E),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)g.lds));
    SXB_HIP(b, hipFuncSetAttribute((const void *)sx::final_batch_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)P.final_lds));
    hipGraph_t g = nullptr;
    SXB_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_amplitude(b, P);
    if (rc == SX_OK) rc = launch_rest(b, P);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != SX_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, SX_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&P.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(b, SX_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    return SX_OK;
}

int copy_out(sx_batch *b, const Plan &P) {
    for (int k : P.ks) {
        const sx_trajectory &t = b->trajs[k];
        const TrajWork &w = b->work[k];
        SXB_HIP(b, hipMemcpyAsync(t.fqt, w.fqt.p, sizeof(double) * 2 * t.nshells * t.NF, hipMemcpyDeviceToHost,
                                  b->stream));
        SXB_HIP(b, hipMemcpyAsync(t.fq, w.fq.p, sizeof(double) * 2 * t.nshells, hipMemcpyDeviceToHost, b->stream));
        SXB_HIP(b, hipMemcpyAsync(t.fq2, w.fq2.p, sizeof(double) * 2 * t.nshells, hipMemcpyDeviceToHost, b->stream));
    }
    SXB_HIP(b, hipStreamSynchronize(b->stream));
    return SX_OK;
}

}  // namespace

extern "C" {

int sx_batch_create(const sx_config *cfg, sx_batch **out) {
    if (!cfg || !out) return SX_ERR_ARG;
    *out = nullptr;
    if (cfg->scattering_type != SX_SCATTERING_ALL || cfg->dsp_type != SX_DSP_AUTOCORRELATE) return SX_ERR_UNSUPPORTED;
    if (cfg->dsp_method != SX_METHOD_FFT && cfg->dsp_method != SX_METHOD_DIRECT) return SX_ERR_ARG;
    if (cfg->tile_shape < 0 || cfg->tile_shape > sx::kOrientShapes) return SX_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) return SX_ERR_HIP;
    if (hipSetDevice(cfg->device) != hipSuccess) return SX_ERR_HIP;
    sx_batch *b = new (std::nothrow) sx_batch();
    if (!b) return SX_ERR_NOMEM;
    b->cfg = *cfg;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess) {
        sx_batch_destroy(b);
        return SX_ERR_HIP;
    }
    *out = b;
    return SX_OK;
}

void sx_batch_destroy(sx_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->cfg.device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->all.reset();
    b->one.clear();
    for (TrajWork &w : b->work) {
        for (DevBuf *d : {&w.qlen, &w.u, &w.qres, &w.amp, &w.ppart, &w.apart, &w.fqt, &w.fq, &w.fq2}) d->release();
        sx_destroy(w.ctx);
    }
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *sx_batch_last_error(const sx_batch *b) { return b ? b->err.c_str() : "null batch"; }

int sx_batch_stage(sx_batch *b, const sx_trajectory *trajs, size_t ntraj) {
    if (!b) return SX_ERR_ARG;
    if (!trajs || ntraj == 0 || ntraj > (1u << 16)) return bfail(b, SX_ERR_ARG, "sx_batch_stage: bad trajectory list");
    if (!b->trajs.empty()) return bfail(b, SX_ERR_STATE, "sx_batch_stage: batch already staged");
    SXB_HIP(b, hipSetDevice(b->cfg.device));
    for (size_t k = 0; k < ntraj; ++k) {
        const sx_trajectory &t = trajs[k];
        if (!t.xyz || !t.b || !t.qlen || !t.u || !t.fqt || !t.fq || !t.fq2 || t.NA == 0 || t.NF == 0 ||
            t.nshells == 0 || t.NM == 0 || t.nshells > 65535 || t.NM > 200000)
            return bfail(b, SX_ERR_ARG, "sx_batch_stage: trajectory %zu is incomplete", k);
        for (size_t s = 0; s < t.nshells; ++s)
            if (!std::isfinite(t.qlen[s])) return bfail(b, SX_ERR_ARG, "trajectory %zu: non-finite |q|", k);
        for (size_t m = 0; m < 3 * t.NM; ++m)
            if (!std::isfinite(t.u[m])) return bfail(b, SX_ERR_ARG, "trajectory %zu: non-finite orientation", k);
    }
    b->trajs.assign(trajs, trajs + ntraj);
    b->work.resize(ntraj);
    for (size_t k = 0; k < ntraj; ++k) {
        const sx_trajectory &t = b->trajs[k];
        TrajWork &w = b->work[k];
        sx_config c = b->cfg;
        int rc = sx_create(&c, &w.ctx);
        if (rc != SX_OK) return bfail(b, rc, "trajectory %zu: sx_create failed", k);
        rc = sx_stage(w.ctx, t.xyz, t.NA, t.NF, t.layout);
        if (rc == SX_OK) rc = sx_set_factors(w.ctx, t.b, t.NA);
        if (rc != SX_OK) return bfail(b, rc, "trajectory %zu: %s", k, sx_last_error(w.ctx));
        if (!w.ctx->plan_ok || w.ctx->plan.N1 > 1)
            return bfail(b, SX_ERR_UNSUPPORTED,
                         "trajectory %zu: NF=%zu above the batch's one-block FFT lengths (at most 10240 frames); "
                         "run it through sx_compute_shells_oriented", k, t.NF);
        std::vector<double> ql(t.nshells);
        for (size_t s = 0; s < t.nshells; ++s) ql[s] = t.qlen[s] / kTwoPi;
        sx::ShellGrid g;
        w.mode = sx::shell_progression(ql.data(), (int)t.nshells, w.ctx->rmax, &g);
        w.qstep = g.step;
        w.q0 = g.q0;
        if (w.mode == 2) {
            // tile lines and residuals, per tile plan (build_plan)
            w.grid = g;
            w.qturns = ql;
            SXB_HIP(b, w.qres.ensure(sx::shell_table_bytes((int)t.nshells)));
        }
        w.chunks = (int)std::min<size_t>(t.NM, 8);
        SXB_HIP(b, w.qlen.ensure(sizeof(double) * t.nshells));
        SXB_HIP(b, hipMemcpy(w.qlen.p, ql.data(), sizeof(double) * t.nshells, hipMemcpyHostToDevice));
        SXB_HIP(b, w.u.ensure(sizeof(double) * 3 * t.NM));
        SXB_HIP(b, hipMemcpy(w.u.p, t.u, sizeof(double) * 3 * t.NM, hipMemcpyHostToDevice));
        SXB_HIP(b, w.amp.ensure(sizeof(float2) * t.nshells * t.NM * w.ctx->NFp));
        SXB_HIP(b, w.ppart.ensure(sizeof(double) * t.nshells * w.chunks * w.ctx->plan.L));
        SXB_HIP(b, w.apart.ensure(sizeof(double) * 4 * t.nshells * w.chunks));
        SXB_HIP(b, w.fqt.ensure(sizeof(double) * 2 * t.nshells * t.NF));
        SXB_HIP(b, w.fq.ensure(sizeof(double) * 2 * t.nshells));
        SXB_HIP(b, w.fq2.ensure(sizeof(double) * 2 * t.nshells));
    }
    // persistent grid: the workgroups that are resident at once
    int dev = b->cfg.device, cus = 0, per_cu = 0;
    SXB_HIP(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    for (int qi = 0; qi < sx::kQueueKinds; ++qi)
        for (int sh = 0; sh < sx::kOrientShapes; ++sh) {
            const size_t lds = qi ? sx::persistent_res_lds(sh) : 0;
            const void *fn = (const void *)sx::persistent_fn(sh, qi);
            if (lds > 0) SXB_HIP(b, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            SXB_HIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds));
            b->max_resident[qi][sh] = cus * std::max(per_cu, 1);
        }

    std::vector<int> all(ntraj);
    std::iota(all.begin(), all.end(), 0);
    // a fixed tile shape when asked for, else the mixed kernel
    b->shape = b->cfg.tile_shape > 0 ? b->cfg.tile_shape - 1 : sx::kMixedShape;
    b->all = std::make_unique<Plan>();
    int rc = build_plan(b, all, b->shape, *b->all);
    if (rc != SX_OK) return rc;
    b->timing.tiles = (size_t)b->all->nitems;
    b->timing.tile_shape = b->shape;
    size_t terms = 0;
    b->timing.terms_stepped = b->timing.terms_residual = 0;
    for (size_t k = 0; k < ntraj; ++k) {
        const sx_trajectory &t = b->trajs[k];
        const size_t n = t.nshells * t.NM * t.NA * t.NF;
        terms += n;
        if (b->work[k].mode == 1) b->timing.terms_stepped += n;
        if (b->work[k].mode == 2) b->timing.terms_residual += n;
    }
    b->timing.terms_exact_runs = b->all->terms_exact_runs;
    b->timing.terms = terms;
    rc = capture(b, *b->all);
    if (rc != SX_OK) return rc;
    b->one.resize(ntraj);
    return SX_OK;
}

int sx_batch_run(sx_batch *b) {
    if (!b) return SX_ERR_ARG;
    if (!b->all) return bfail(b, SX_ERR_STATE, "sx_batch_run: stage the batch first");
    SXB_HIP(b, hipSetDevice(b->cfg.device));
    SXB_HIP(b, hipEventRecord(b->e0, b->stream));
    SXB_HIP(b, hipGraphLaunch(b->all->exec, b->stream));
    SXB_HIP(b, hipEventRecord(b->e1, b->stream));
    int rc = copy_out(b, *b->all);
    if (rc != SX_OK) return rc;
    b->timing.total_ms = elapsed(b->e0, b->e1);
    return SX_OK;
}

int sx_batch_time_amplitude(sx_batch *b, double *ms) {
    if (!b || !ms) return SX_ERR_ARG;
    if (!b->all) return bfail(b, SX_ERR_STATE, "sx_batch_time_amplitude: stage the batch first");
    SXB_HIP(b, hipSetDevice(b->cfg.device));
    SXB_HIP(b, hipEventRecord(b->e0, b->stream));
    int rc = launch_amplitude(b, *b->all);
    if (rc != SX_OK) return rc;
    SXB_HIP(b, hipEventRecord(b->e1, b->stream));
    SXB_HIP(b, hipEventSynchronize(b->e1));
    *ms = elapsed(b->e0, b->e1);
    b->timing.amplitude_ms = *ms;
    return SX_OK;
}

int sx_batch_run_one(sx_batch *b, size_t k, double *ms) {
    if (!b) return SX_ERR_ARG;
    if (!b->all) return bfail(b, SX_ERR_STATE, "sx_batch_run_one: stage the batch first");
    if (k >= b->trajs.size()) return bfail(b, SX_ERR_ARG, "sx_batch_run_one: no trajectory %zu", k);
    SXB_HIP(b, hipSetDevice(b->cfg.device));
    if (!b->one[k]) {
        auto P = std::make_unique<Plan>();
        int rc = build_plan(b, {(int)k}, b->shape, *P);
        if (rc == SX_OK) rc = capture(b, *P);
        if (rc != SX_OK) return rc;
        b->one[k] = std::move(P);
    }
    SXB_HIP(b, hipEventRecord(b->e0, b->stream));
    SXB_HIP(b, hipGraphLaunch(b->one[k]->exec, b->stream));
    SXB_HIP(b, hipEventRecord(b->e1, b->stream));
    int rc = copy_out(b, *b->one[k]);
    if (rc != SX_OK) return rc;
    if (ms) *ms = elapsed(b->e0, b->e1);
    return SX_OK;
}

int sx_batch_get_timing(const sx_batch *b, sx_batch_timing *out) {
    if (!b || !out) return SX_ERR_ARG;
    *out = b->timing;
    return SX_OK;
}

}  // extern "C"
