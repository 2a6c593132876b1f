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

// Trajectories whose shells step with residual corrections (mode 2) keep
// their FP64 totals in the dynamic LDS, which the host sizes to
// kResidualLds<VM, VS> when the batch holds any such trajectory; each
// wavefront touches only its own slice, so tiles need no barrier between them.
template <int VM, int VS>
__global__ __launch_bounds__(256, 2) void amp_persistent_kernel(const OrientKernelArgs *__restrict__ targs,
                                                             const int4 *__restrict__ items,
                                                             const int *__restrict__ qstart,
                                                             unsigned *__restrict__ counters) {
    extern __shared__ double tot_lds[];
    __shared__ int cur;
    const int x = blockIdx.x & 7;
    const int begin = qstart[x], n = qstart[x + 1] - begin;
    for (;;) {
        if (threadIdx.x == 0) cur = (int)atomicAdd(&counters[x], 1u);
        __syncthreads();
        const int it = cur;
        __syncthreads();  // every thread has read cur before it is overwritten
        if (it >= n) break;
        const int4 w = items[begin + it];
        const OrientKernelArgs &ta = targs[w.x];
        if (ta.ap == 2)
            oriented_tile_body<VM, VS, false, 2, false>(ta, w.y, w.w, w.z & 0xffff, w.z >> 16, tot_lds);
        else
            oriented_tile<VM, VS, false, false>(ta, w.y, w.w, w.z & 0xffff, w.z >> 16);
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

static size_t persistent_res_lds(int shape) {
    switch (shape) {
        case 0: return kResidualLds<2, 8>;
        case 1: return kResidualLds<4, 4>;
        case 2: return kResidualLds<1, 8>;
        case 3: return kResidualLds<1, 16>;
        default: return kResidualLds<1, 20>;
    }
}

static PersistentFn persistent_fn(int shape) {
    switch (shape) {
        case 0: return amp_persistent_kernel<2, 8>;
        case 1: return amp_persistent_kernel<4, 4>;
        case 2: return amp_persistent_kernel<1, 8>;
        case 3: return amp_persistent_kernel<1, 16>;
        default: return amp_persistent_kernel<1, 20>;
    }
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
    double qstep = 0.0;    // step and origin of the shells in |q| / 2 pi
    double q0 = 0.0;
    int chunks = 1;
};

// Device-side launch plan of a set of trajectories.
struct Plan {
    std::vector<int> ks;
    DevBuf oka, items, qstart, counters, cargs, citems, fargs, fitems;
    int nitems = 0, nfitems = 0, persistent_blocks = 0;
    size_t final_lds = 0;
    struct CorrGroup {
        int E, T, first, count;
        size_t lds;
    };
    std::vector<CorrGroup> groups;
    hipGraphExec_t exec = nullptr;
    ~Plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (DevBuf *d : {&oka, &items, &qstart, &counters, &cargs, &citems, &fargs, &fitems}) d->release();
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
    int max_resident = 0;  // persistent blocks that fit the GPU at once
    bool residuals = false;  // some trajectory steps its shells with residual corrections
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
    oa.qres = w.qres.as<float>();
    return sx::make_orient_args(oa);
}

// Builds the device work lists of trajectories ks for tile shape `shape`;
// stride > 1 keeps every stride-th amplitude tile (autotuning samples), but
// never fewer than about kMinSampleTiles tiles.
constexpr size_t kMinSampleTiles = 2048;

int build_plan(sx_batch *b, const std::vector<int> &ks, int shape, int stride, Plan &P) {
    P.ks = ks;
    const int VM = sx::orient_shape_vm(shape), VS = sx::orient_shape_vs(shape);
    std::vector<sx::OrientKernelArgs> oka(b->trajs.size());
    struct Item {
        int4 w;
        long cost;
    };
    std::vector<Item> items;
    for (int k : ks) {
        const sx_trajectory &t = b->trajs[k];
        oka[k] = orient_args(b->work[k], t, shape);
        const int ftiles = oka[k].ftiles, og = oka[k].ogroups, sg = oka[k].sgroups;
        const long cost = (long)oka[k].atoms_per_split;
        for (int ft = 0; ft < ftiles; ++ft)
            for (int o = 0; o < og; ++o)
                for (int s = 0; s < sg; ++s) items.push_back({make_int4(k, ft, o | (s << 16), 0), cost});
    }
    (void)VM;
    (void)VS;
    // largest tiles first; within a trajectory, tiles sharing a frame tile stay adjacent
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &c) { return a.cost > c.cost; });
    if (stride > 1) stride = std::max(1, std::min(stride, (int)(items.size() / kMinSampleTiles)));
    if (stride > 1) {
        std::vector<Item> keep;
        for (size_t i = 0; i < items.size(); i += stride) keep.push_back(items[i]);
        items.swap(keep);
    }
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
    P.nitems = (int)items.size();
    SXB_HIP(b, P.oka.ensure(sizeof(sx::OrientKernelArgs) * oka.size()));
    SXB_HIP(b, hipMemcpy(P.oka.p, oka.data(), sizeof(sx::OrientKernelArgs) * oka.size(), hipMemcpyHostToDevice));
    SXB_HIP(b, P.items.ensure(sizeof(int4) * std::max<size_t>(1, w.size())));
    if (!w.empty()) SXB_HIP(b, hipMemcpy(P.items.p, w.data(), sizeof(int4) * w.size(), hipMemcpyHostToDevice));
    SXB_HIP(b, P.qstart.ensure(sizeof(int) * qstart.size()));
    SXB_HIP(b, hipMemcpy(P.qstart.p, qstart.data(), sizeof(int) * qstart.size(), hipMemcpyHostToDevice));
    SXB_HIP(b, P.counters.ensure(sizeof(unsigned) * kXcds));
    P.persistent_blocks = std::max(kXcds, std::min(b->max_resident, (P.nitems + kXcds - 1) / kXcds * kXcds));
    P.persistent_blocks = (P.persistent_blocks + kXcds - 1) / kXcds * kXcds;

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

int launch_amplitude(sx_batch *b, Plan &P, int shape) {
    hipLaunchKernelGGL(sx::reset_counters_kernel, dim3(1), dim3(kXcds), 0, b->stream, P.counters.as<unsigned>(),
                       kXcds);
    SXB_HIP(b, hipGetLastError());
    // (the dynamic-LDS attribute of every shape is set once, in sx_batch_stage)
    const size_t lds = b->residuals ? sx::persistent_res_lds(shape) : 0;
    hipLaunchKernelGGL(sx::persistent_fn(shape), dim3(P.persistent_blocks), dim3(256), lds, b->stream,
                       P.oka.as<sx::OrientKernelArgs>(), P.items.as<int4>(), P.qstart.as<int>(),
                       P.counters.as<unsigned>());
    SXB_HIP(b, hipGetLastError());
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
    int rc = launch_amplitude(b, P, b->shape);
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
            std::vector<float> rr(t.nshells);
            for (size_t s = 0; s < t.nshells; ++s) rr[s] = (float)(kTwoPi * g.res[s]);  // radians per Angstrom
            SXB_HIP(b, w.qres.ensure(sizeof(float) * t.nshells));
            SXB_HIP(b, hipMemcpy(w.qres.p, rr.data(), sizeof(float) * t.nshells, hipMemcpyHostToDevice));
            b->residuals = true;
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
    int min_per_cu = 1 << 30;
    for (int sh = 0; sh < sx::kOrientShapes; ++sh) {
        const size_t lds = b->residuals ? sx::persistent_res_lds(sh) : 0;
        if (lds > 0)
            SXB_HIP(b, hipFuncSetAttribute((const void *)sx::persistent_fn(sh),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        SXB_HIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)sx::persistent_fn(sh), 256,
                                                               lds));
        min_per_cu = std::min(min_per_cu, std::max(per_cu, 1));
    }
    b->max_resident = cus * min_per_cu;

    std::vector<int> all(ntraj);
    std::iota(all.begin(), all.end(), 0);
    // autotune: each shape on every 64th amplitude tile of the batch (at least
    // kMinSampleTiles tiles)
    if (b->cfg.tile_shape > 0) {
        b->shape = b->cfg.tile_shape - 1;
    } else {
        float best = 1e30f;
        for (int sh = 0; sh < sx::kOrientShapes; ++sh) {
            Plan P;
            int rc = build_plan(b, all, sh, 64, P);
            if (rc != SX_OK) return rc;
            SXB_HIP(b, hipEventRecord(b->e0, b->stream));
            rc = launch_amplitude(b, P, sh);
            if (rc != SX_OK) return rc;
            SXB_HIP(b, hipEventRecord(b->e1, b->stream));
            SXB_HIP(b, hipEventSynchronize(b->e1));
            const float ms = elapsed(b->e0, b->e1);
            if (ms < best) {
                best = ms;
                b->shape = sh;
            }
        }
    }
    b->all = std::make_unique<Plan>();
    int rc = build_plan(b, all, b->shape, 1, *b->all);
    if (rc != SX_OK) return rc;
    b->timing.tiles = (size_t)b->all->nitems;
    b->timing.tile_shape = b->shape;
    size_t terms = 0;
    for (const sx_trajectory &t : b->trajs) terms += t.nshells * t.NM * t.NA * t.NF;
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
    int rc = launch_amplitude(b, *b->all, b->shape);
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
        int rc = build_plan(b, {(int)k}, b->shape, 1, *P);
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
