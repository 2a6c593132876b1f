// sx_capi.hip -- the C ABI (include/sassena_amd.h) over the gfx950 kernels.
//
// A context owns one HIP stream (default priority), the staged SoA block and
// the work buffers, which grow on demand and are kept between calls.  Host
// code here only validates, lays out small tables (q/(2 pi), twiddles) and
// sequences launches; every arithmetic step of the path runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <cmath>
#include <numeric>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sx_ctx.h"

using namespace sxh;

// Work memory of the allocation-free entry point (sx_compute_shells_oriented_noalloc):
// one file-scope device array per GPU, held until the process ends.  Host code
// never passes the array's name to a kernel or a copy (on the host it is only
// a stand-in symbol): it resolves the array's address on each device once,
// with that device current, and uses that address.
constexpr size_t kScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on the current device `dev`, resolved on first use.
hipError_t scratch_address(int dev, void **out) {
    *out = nullptr;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

float sxh::elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

namespace {

int fail(sx_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define SX_HIP(ctx, expr)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail((ctx), e_ == hipErrorOutOfMemory ? SX_ERR_NOMEM : SX_ERR_HIP, "%s failed: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                             \
    } while (0)

int pick_vector_block(const sx_ctx *c) {
    if (c->cfg.vector_block == 4 || c->cfg.vector_block == 8 || c->cfg.vector_block == 16)
        return c->cfg.vector_block;
    return 8;
}

// Atom ranges per amplitude tile: enough workgroups (>= 2048, eight
// wavefronts per SIMD) to fill the 256 CUs, at most 16 partial sums.
int pick_split_blocks(const sx_ctx *c, long base) {
    base = std::max<long>(base, 1);
    if (c->cfg.atom_split > 0) return std::max(1, std::min(c->cfg.atom_split, (int)(c->NAp / 16)));
    long split = (2048 + base - 1) / base;
    split = std::min<long>(split, 16);
    split = std::min<long>(split, (long)(c->NAp / 64));
    return (int)std::max<long>(split, 1);
}

// Atom ranges of the oriented kernel.  Its base*split workgroups do equal
// work, so a count just above a multiple of the workgroups resident at once
// (occupancy x CUs) leaves a last round with most of the GPU idle (3,200 tiles
// on 512 slots: 7 rounds for 6.25 of work).  Among the splits from the rule
// above to 16 (at least 64 atoms a range), take the one with the fewest
// rounds per unit of work; ties go to the smaller split.
int round_filling_split(long base, int first, long resident, size_t NAp, size_t split_bytes);

int pick_oriented_split(sx_ctx *c, long base, int shape, bool per_row, int mode, size_t V, int *split) {
    base = std::max<long>(base, 1);
    *split = pick_split_blocks(c, base);
    if (c->cfg.atom_split > 0 || c->noalloc) return SX_OK;
    int per_cu = 0, cus = 0;
    SX_HIP(c, sx::oriented_blocks_per_cu(shape, per_row, mode, &per_cu));
    SX_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->cfg.device));
    const size_t split_bytes = sizeof(float2) * c->NFp * std::max<size_t>(V, 1);
    *split = round_filling_split(base, *split, (long)std::max(per_cu, 1) * std::max(cus, 1), c->NAp, split_bytes);
    return SX_OK;
}

int pick_split(const sx_ctx *c, int V, int vb) {
    if (c->cfg.atom_split > 0) return std::max(1, std::min(c->cfg.atom_split, (int)(c->NAp / sx::kAtomChunk)));
    const long ftiles = (long)(c->NFp / sx::kWave);
    const long vgroups = (V + 4L * vb - 1) / (4L * vb);
    const long base = ftiles * vgroups;
    long split = (2048 + base - 1) / base;
    split = std::min<long>(split, 16);
    split = std::min<long>(split, (long)(c->NAp / 64));
    return (int)std::max<long>(split, 1);
}

// Amplitude bytes the extra atom ranges of round_filling_split may add.
constexpr size_t kSplitAmpBudget = (size_t)4 << 30;
// Largest buffer of per-(atom, frame) spherical terms of the sphere multipole path.
constexpr size_t kSphereTermBudget = (size_t)16 << 30;
// The sphere's table of j_l per (shell, degree, atom, frame): at most this
// much, and only while this much device memory stays free beside it.
constexpr size_t kBesselTableBudget = (size_t)96 << 30;
constexpr size_t kBesselTableMargin = (size_t)4 << 30;

// Splits from `first` to 16 (at least 64 atoms a range, partial amplitude
// sets of split_bytes each within kSplitAmpBudget) for `base` equal
// workgroups per split: the one with the fewest rounds of `resident`
// workgroups per unit of work, ties to the smaller split.
int round_filling_split(long base, int first, long resident, size_t NAp, size_t split_bytes) {
    const long smem = (long)(kSplitAmpBudget / std::max<size_t>(split_bytes, 1));
    const long smax = std::max<long>(first, std::min<long>({16L, (long)(NAp / 64), smem}));
    int split = first;
    double best = 0.0;
    for (long s = first; s <= smax; ++s) {
        const long blocks = (base * s + 7) / 8 * 8;
        const double rounds = (double)((blocks + resident - 1) / resident) / (double)s;
        if (s == first || rounds < best * 0.99) {
            best = rounds;
            split = (int)s;
        }
    }
    return split;
}

// Whether the sphere's Bessel table of `bytes` may be held: within its
// budget, and with kBesselTableMargin of device memory left beside it.
bool bessel_table_fits(sx_ctx *c, size_t bytes) {
    if (bytes > kBesselTableBudget) return false;
    if (c->mpjt.p && bytes <= c->mpjt.cap) return true;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
    return fr + c->mpjt.cap >= bytes + kBesselTableMargin;
}

// Atom ranges of the multipole kernel.  Its workgroups each sum every atom of
// their range for one (frame tile, shell, group of 4 moment blocks): the rule
// above, then the split that fills whole rounds of resident workgroups.
int pick_multipole_split(sx_ctx *c, int kind, int ns, int nmo, int *split) {
    const long base = sx::multipole_base_blocks(c->NFp, ns, nmo, kind);
    *split = pick_split_blocks(c, base);
    if (c->cfg.atom_split > 0) return SX_OK;
    int per_cu = 0, cus = 0;
    SX_HIP(c, sx::multipole_blocks_per_cu(kind, &per_cu));
    SX_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->cfg.device));
    *split = round_filling_split(base, *split, (long)std::max(per_cu, 1) * std::max(cus, 1), c->NAp,
                                 sizeof(float2) * c->NFp * (size_t)ns * nmo);
    return SX_OK;
}

// Atom ranges of the general-vector kernel: the rule above, then the split
// that fills whole rounds of resident workgroups (see pick_oriented_split).
int pick_all_split(sx_ctx *c, int V, int vb, bool per_row, int *split) {
    *split = pick_split(c, V, vb);
    if (c->cfg.atom_split > 0) return SX_OK;
    int per_cu = 0, cus = 0;
    SX_HIP(c, sx::all_blocks_per_cu(vb, per_row, &per_cu));
    SX_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->cfg.device));
    const long base = (long)(c->NFp / sx::kWave) * ((V + 4L * vb - 1) / (4L * vb));
    *split = round_filling_split(base, *split, (long)std::max(per_cu, 1) * std::max(cus, 1), c->NAp,
                                 sizeof(float2) * c->NFp * (size_t)V);
    return SX_OK;
}

// Largest buffer of per-pair parts of a the split "self" path may use.
constexpr size_t kSplitBudgetBytes = (size_t)16 << 30;

int prepare_tables(sx_ctx *c) {
    c->plan_ok = sx::make_fft_plan(c->NF, &c->plan);
    if (!c->plan_ok) return SX_OK;  // only the autocorrelation needs it
    const int L = c->plan.L;
    std::vector<float2> tw(L);
    std::vector<double2> twf(L), tw64(L);
    for (int e = 0; e < L; ++e) {
        const double ang = kTwoPi * (double)e / (double)L;
        const double cs = std::cos(ang), sn = std::sin(ang);
        tw[e] = make_float2((float)cs, (float)-sn);  // exp(-2 pi i e / L)
        twf[e] = make_double2(cs, -sn);              // exp(-2 pi i e / L)
        tw64[e] = make_double2(cs, sn);              // exp(+2 pi i e / L)
    }
    SX_HIP(c, c->twL.ensure(sizeof(float2) * L));
    SX_HIP(c, c->tw64.ensure(sizeof(double2) * L));
    SX_HIP(c, c->twL64.ensure(sizeof(double2) * L));
    SX_HIP(c, hipMemcpyAsync(c->twL64.p, twf.data(), sizeof(double2) * L, hipMemcpyHostToDevice, c->stream));
    SX_HIP(c, c->w.ensure(sizeof(double2) * L));
    SX_HIP(c, hipMemcpyAsync(c->twL.p, tw.data(), sizeof(float2) * L, hipMemcpyHostToDevice, c->stream));
    SX_HIP(c, hipMemcpyAsync(c->tw64.p, tw64.data(), sizeof(double2) * L, hipMemcpyHostToDevice, c->stream));
    SX_HIP(c, sx::launch_weights(c->plan, (int)c->NF, c->tw64.as<double2>(), c->w.as<double2>(), c->stream));
    // packed pass twiddles: exp(-2 pi i q kk / (Ns R)) at toff[p] + kk R + q
    const sx::FftPlan &P = c->plan;
    std::vector<float2> tp(P.ntw);
    std::vector<double2> tp64(P.ntw);
    for (int p = 1; p < P.npass; ++p) {
        const int Ns = P.ns[p], R = P.radix[p];
        for (int kk = 0; kk < Ns; ++kk)
            for (int q = 0; q < R; ++q) {
                const double ang = kTwoPi * (double)(q * kk) / (double)(Ns * R);
                const double cs = std::cos(ang), sn = std::sin(ang);
                tp[P.toff[p] + kk * R + q] = make_float2((float)cs, (float)-sn);
                tp64[P.toff[p] + kk * R + q] = make_double2(cs, -sn);
            }
    }
    SX_HIP(c, c->twp.ensure(sizeof(float2) * P.ntw));
    SX_HIP(c, c->twp64.ensure(sizeof(double2) * P.ntw));
    SX_HIP(c, hipMemcpyAsync(c->twp.p, tp.data(), sizeof(float2) * P.ntw, hipMemcpyHostToDevice, c->stream));
    SX_HIP(c, hipMemcpyAsync(c->twp64.p, tp64.data(), sizeof(double2) * P.ntw, hipMemcpyHostToDevice, c->stream));
    SX_HIP(c, hipStreamSynchronize(c->stream));
    return SX_OK;
}


// Tile shape of the oriented amplitude kernel.  A fixed cfg.tile_shape wins;
// otherwise, for calls of at least kTuneTerms terms, every shape is timed once
// on the real data (HIP events on the context stream) and the fastest is kept
// for this problem size.  The trials sum over a prefix of the atoms of about
// kTuneSampleTerms terms (the launch grid is that of the full call; its time
// is linear in the atoms summed), so tuning a large call costs a few
// milliseconds per shape.  Smaller calls take a default: shells stepped by the
// recurrence amortise the per-atom phase work over the tile's shells, so more
// than 16 such shells take 1 x 20 and 9..16 take 1 x 16 (with one factor per
// atom: per-shell factors need more registers than those shapes have without
// spilling); everything else 2 x 8.
constexpr double kTuneTerms = 2e10;
constexpr double kTuneSampleTerms = 5e10;

int default_shape(const sx_ctx *c, int ns, bool per_shell) {
    if (!c->q_mode || per_shell) return 0;
    return ns > 16 ? 4 : ns > 8 ? 3 : 0;
}

// The tile lines and residuals of a residual-corrected launch of this tile
// shape, onto the stream ahead of it (the tiles start at multiples of VS);
// *plain: the shells it steps without corrections.
int upload_tile_lines(sx_ctx *c, int ns, int shape, int *plain) {
    *plain = 0;
    if (c->q_mode != 2) return SX_OK;
    const int VS = sx::orient_shape_vs(shape);
    std::vector<int> firsts, vs;
    for (int s0 = 0; s0 < ns; s0 += VS) {
        firsts.push_back(s0);
        vs.push_back(VS);
    }
    // the copies queued from these host tables by an earlier launch are done
    // before they are overwritten
    SX_HIP(c, hipStreamSynchronize(c->stream));
    c->tl_lines.assign(ns, sx::TileLine{});
    c->tl_res.assign(ns, 0.f);
    *plain = sx::tile_lines(c->q_turns.data(), ns, c->rmax, c->q_grid, firsts.data(), vs.data(), (int)firsts.size(),
                            c->tl_lines.data(), c->tl_res.data());
    const size_t lb = sizeof(sx::TileLine) * ns;
    SX_HIP(c, hipMemcpyAsync(c->qres.p, c->tl_lines.data(), lb, hipMemcpyHostToDevice, c->stream));
    SX_HIP(c, hipMemcpyAsync(static_cast<char *>(c->qres.p) + lb, c->tl_res.data(), sizeof(float) * ns,
                             hipMemcpyHostToDevice, c->stream));
    return SX_OK;
}

// OrientArgs of the current call (stepping mode and tables from run_batch).
sx::OrientArgs oriented_args(const sx_ctx *c, int ns, int NM, const float *bptr, bool per_shell, int shape,
                             int split) {
    sx::OrientArgs oa;
    oa.coords = c->coords.as<float>();
    oa.NAp = c->NAp;
    oa.NFp = c->NFp;
    oa.qlen = c->qlen.as<double>();
    oa.u = c->uvec.as<double>();
    oa.b = bptr;
    oa.NM = NM;
    oa.ns = ns;
    oa.split = split;
    oa.per_shell_b = per_shell ? 1 : 0;
    oa.shape = shape;
    oa.amp = c->amp.as<float2>();
    oa.ap = c->q_mode;
    oa.qstep = c->q_step;
    oa.q0 = c->q_origin;
    // mode 2: [ns] tile lines, then [ns] residuals (upload_tile_lines)
    oa.lines = c->q_mode == 2 ? c->qres.as<sx::TileLine>() : nullptr;
    oa.qres = c->q_mode == 2 ? reinterpret_cast<const float *>(oa.lines + ns) : nullptr;
    return oa;
}

int pick_shape(sx_ctx *c, int ns, int NM, const float *bptr, bool per_shell, int *shape) {
    if (c->cfg.tile_shape > 0) {
        if (c->cfg.tile_shape > sx::kOrientShapes)
            return fail(c, SX_ERR_ARG, "tile_shape %d out of range", c->cfg.tile_shape);
        *shape = c->cfg.tile_shape - 1;
        return SX_OK;
    }
    *shape = default_shape(c, ns, per_shell);
    const double terms = (double)ns * NM * (double)c->NA * (double)c->NF;
    if (terms < kTuneTerms || c->noalloc) return SX_OK;
    const std::array<size_t, 6> key = {c->NFp, c->NAp, (size_t)NM, (size_t)ns, per_shell ? 1u : 0u,
                                       (size_t)c->q_mode};
    auto it = c->tuned.find(key);
    if (it != c->tuned.end()) {
        *shape = it->second;
        return SX_OK;
    }
    float best = 1e30f;
    hipEvent_t e0, e1;
    SX_HIP(c, hipEventCreate(&e0));
    SX_HIP(c, hipEventCreate(&e1));
    for (int sh = 0; sh < sx::kOrientShapes; ++sh) {
        int split = 1;
        const int src = pick_oriented_split(c, sx::oriented_base_blocks(c->NFp, NM, ns, sh), sh, per_shell,
                                            c->q_mode, (size_t)ns * NM, &split);
        if (src != SX_OK) return src;
        SX_HIP(c, c->amp.ensure(sizeof(float2) * (size_t)ns * NM * c->NFp * split));
        sx::OrientArgs oa = oriented_args(c, ns, NM, bptr, per_shell, sh, split);
        // atom prefix: whole 64-atom rows, at least one per atom range
        const double frac = std::min(1.0, kTuneSampleTerms / terms);
        size_t na = (size_t)std::ceil(frac * (double)c->NAp / 64.0) * 64;
        na = std::max(na, (size_t)64 * split);
        oa.natoms = std::min(na, c->NAp);
        int plain = 0;
        const int urc = upload_tile_lines(c, ns, sh, &plain);
        if (urc != SX_OK) return urc;
        SX_HIP(c, hipEventRecord(e0, c->stream));
        SX_HIP(c, sx::launch_amp_oriented(oa, c->stream));
        SX_HIP(c, hipEventRecord(e1, c->stream));
        SX_HIP(c, hipEventSynchronize(e1));
        const float ms = elapsed(e0, e1);
        if (ms < best) {
            best = ms;
            *shape = sh;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->tuned[key] = *shape;
    return SX_OK;
}

// Multipole averaging of one call: moments (l, m), per-shell q data.
struct MPSpec {
    int kind;                    // 0 sphere, 1 cylinder
    std::vector<int> moments;    // [nmo][2]
    // sphere, per moment: c_|m| = sqrt((2|m|+1)/(4 pi) prod_{i<=|m|} (2i-1)/(2i)), the start of the
    // Legendre recurrence, and 4 pi f_l|m|, f = sqrt((2l+1)/(2|m|+1) (l-|m|)! (2|m|)! / (l+|m|)!)
    // from log-gamma: no factorial is formed (finite for l <= 200)
    std::vector<double> ynorm;
    // followed by [l1][l1][2] the coefficients of the Legendre recurrence in the degree,
    // ((2 ll - 1) / (ll - m), (ll + m - 1) / (ll - m)) at [ll][m] for m + 2 <= ll < l1
    int plm_l1 = 0;
    std::vector<double> qinfo;   // per shell: |q| (sphere) or q_r, phi_q, q_z (cylinder)
    double base[9];              // cylinder frame of the axis
};

// One batch of shells [s0, s0+ns) through the configured path; outputs land in
// the device buffers at shell offset 0 and are copied to host by the caller.
int run_batch(sx_ctx *c, int ns, size_t NM, const double *qv, const double *qlen, const double *u,
              const double *b_shell, double *fqt_h, double *fq_h, double *fq2_h, const MPSpec *mp = nullptr) {
    const bool self = c->cfg.scattering_type == SX_SCATTERING_SELF;
    const int dsp = c->cfg.dsp_type;
    const int V = (int)(ns * NM);
    const int NF = (int)c->NF;
    const bool prof = c->cfg.profile != 0;

    const bool oriented = qv == nullptr && mp == nullptr;
    // q / (2 pi) per series (the kernels evaluate phases in turns): read by the
    // general-vector amplitude kernel and the self kernels; the coherent
    // oriented kernel forms q from |q| and the orientations
    const bool need_qrev = mp == nullptr && (!oriented || self);
    std::vector<double> qh(need_qrev ? 3 * (size_t)V : 0);
    if (need_qrev && qv) {
        for (size_t i = 0; i < qh.size(); ++i) qh[i] = qv[i] / kTwoPi;
    } else if (need_qrev) {
        for (int s = 0; s < ns; ++s)
            for (size_t m = 0; m < NM; ++m)
                for (int k = 0; k < 3; ++k) qh[(s * NM + m) * 3 + k] = qlen[s] * u[3 * m + k] / kTwoPi;
    }
    if (oriented) {
        std::vector<double> ql(ns);
        for (int s = 0; s < ns; ++s) ql[s] = qlen[s] / kTwoPi;
        sx::ShellGrid g;
        c->q_mode = sx::shell_progression(ql.data(), ns, c->rmax, &g);
        // residual-corrected steps take one factor per atom
        if (c->q_mode == 2 && b_shell) c->q_mode = 0;
        c->q_step = g.step;
        c->q_origin = g.q0;
        if (c->q_mode == 2) {
            // tile lines and residuals, per tile shape at each launch
            c->q_grid = g;
            c->q_turns = ql;
            SX_HIP(c, c->qres.ensure(sx::shell_table_bytes(ns)));
        }
        SX_HIP(c, c->qlen.ensure(sizeof(double) * ns));
        SX_HIP(c, hipMemcpyAsync(c->qlen.p, ql.data(), sizeof(double) * ns, hipMemcpyHostToDevice, c->stream));
        SX_HIP(c, c->uvec.ensure(sizeof(double) * 3 * NM));
        SX_HIP(c, hipMemcpyAsync(c->uvec.p, u, sizeof(double) * 3 * NM, hipMemcpyHostToDevice, c->stream));
    }
    if (need_qrev) {
        SX_HIP(c, c->qrev.ensure(sizeof(double) * qh.size()));
        SX_HIP(c, hipMemcpyAsync(c->qrev.p, qh.data(), sizeof(double) * qh.size(), hipMemcpyHostToDevice, c->stream));
    }

    const float *bptr = c->b.as<float>();
    const int *browp = nullptr;
    if (b_shell) {
        std::vector<float> bs((size_t)ns * c->NAp, 0.f);
        for (int s = 0; s < ns; ++s)
            for (size_t j = 0; j < c->NA; ++j) bs[(size_t)s * c->NAp + j] = (float)b_shell[(size_t)s * c->NA + j];
        SX_HIP(c, c->bshell.ensure(sizeof(float) * bs.size()));
        SX_HIP(c, hipMemcpyAsync(c->bshell.p, bs.data(), sizeof(float) * bs.size(), hipMemcpyHostToDevice,
                                 c->stream));
        bptr = c->bshell.as<float>();
        if (!self) {
            std::vector<int> br(V);
            for (int v = 0; v < V; ++v) br[v] = v / (int)NM;
            SX_HIP(c, c->brow.ensure(sizeof(int) * V));
            SX_HIP(c, hipMemcpyAsync(c->brow.p, br.data(), sizeof(int) * V, hipMemcpyHostToDevice, c->stream));
            browp = c->brow.as<int>();
        }
    }

    SX_HIP(c, c->fqt.ensure(sizeof(double) * 2 * (size_t)ns * NF));
    SX_HIP(c, c->fq.ensure(sizeof(double) * 2 * ns));
    SX_HIP(c, c->fq2.ensure(sizeof(double) * 2 * ns));

    sx::FinalArgs fa{};
    fa.NF = NF;
    fa.NM = (int)NM;
    fa.nshells = ns;
    fa.conj = (dsp == SX_DSP_AUTOCORRELATE && c->cfg.dsp_method == SX_METHOD_DIRECT) ? 1 : 0;
    // final factor: 1/NM for vector averages (all_vectors_scatter_device.cpp:355), 1/(4 pi) for
    // multipole spheres and 1/(2 pi) for cylinders (multipole_scatter_device.cpp:394, 866)
    const double norm = mp ? (mp->kind == 0 ? 1.0 / (2.0 * kTwoPi) : 1.0 / kTwoPi) : 1.0 / (double)NM;
    fa.norm = norm;
    fa.fqt = c->fqt.as<double>();
    fa.fq = c->fq.as<double>();
    fa.fq2 = c->fq2.as<double>();

    if (prof) SX_HIP(c, hipEventRecord(c->ev[0], c->stream));
    if (!self) {
        const int vb = pick_vector_block(c);
        int shape = 0;
        if (oriented && !mp) {
            int rc = pick_shape(c, ns, (int)NM, bptr, b_shell != nullptr, &shape);
            if (rc != SX_OK) return rc;
        }
        int split = 1;
        if (!oriented && !mp) {
            int rc = pick_all_split(c, V, vb, browp != nullptr, &split);
            if (rc != SX_OK) return rc;
        }
        if (oriented && !mp) {
            int rc = pick_oriented_split(c, sx::oriented_base_blocks(c->NFp, (int)NM, ns, shape), shape,
                                         b_shell != nullptr, c->q_mode, (size_t)ns * NM, &split);
            if (rc != SX_OK) return rc;
        }
        if (mp) {
            int rc = pick_multipole_split(c, mp->kind, ns, (int)NM, &split);
            if (rc != SX_OK) return rc;
        }
        const size_t split_stride = (size_t)V * c->NFp;
        SX_HIP(c, c->amp.ensure(sizeof(float2) * split_stride * split));
        if (mp) {
            SX_HIP(c, c->mpmom.ensure(sizeof(int) * mp->moments.size()));
            SX_HIP(c, hipMemcpyAsync(c->mpmom.p, mp->moments.data(), sizeof(int) * mp->moments.size(),
                                     hipMemcpyHostToDevice, c->stream));
            // q data, the cylinder frame and the sphere's Y_lm factors in one buffer
            SX_HIP(c, c->mpq.ensure(sizeof(double) * (mp->qinfo.size() + 9 + mp->ynorm.size())));
            SX_HIP(c, hipMemcpyAsync(c->mpq.p, mp->qinfo.data(), sizeof(double) * mp->qinfo.size(),
                                     hipMemcpyHostToDevice, c->stream));
            SX_HIP(c, hipMemcpyAsync(c->mpq.as<double>() + mp->qinfo.size(), mp->base, sizeof(double) * 9,
                                     hipMemcpyHostToDevice, c->stream));
            if (!mp->ynorm.empty())
                SX_HIP(c, hipMemcpyAsync(c->mpq.as<double>() + mp->qinfo.size() + 9, mp->ynorm.data(),
                                         sizeof(double) * mp->ynorm.size(), hipMemcpyHostToDevice, c->stream));
            sx::MultipoleArgs ma;
            ma.coords = c->coords.as<float>();
            ma.NAp = c->NAp;
            ma.NFp = c->NFp;
            ma.kind = mp->kind;
            ma.qinfo = c->mpq.as<double>();
            ma.base = c->mpq.as<double>() + mp->qinfo.size();
            ma.moments = c->mpmom.as<int>();
            ma.ynorm = mp->ynorm.empty() ? nullptr : c->mpq.as<double>() + mp->qinfo.size() + 9;
            ma.plm_l1 = mp->plm_l1;
            ma.nmo = (int)NM;
            ma.ns = ns;
            ma.b = bptr;
            ma.per_shell_b = b_shell ? 1 : 0;
            ma.amp = c->amp.as<float2>();
            ma.split = split;
            // The spherical (36 B) or cylinder (12 B) terms of every (atom,
            // frame) once per call when they fit the budget; else the kernel
            // forms them.
            const size_t sph_bytes =
                (mp->kind == 0 ? sx::kSphereTermBytes : sx::kCylinderTermBytes) * c->NAp * c->NFp;
            const int terms = c->cfg.multipole_terms;
            if (terms == 1 || terms == 3 || (terms == 0 && sph_bytes <= kSphereTermBudget)) {
                SX_HIP(c, c->mpsph.ensure(sph_bytes));
                if (mp->kind == 0)
                    SX_HIP(c, sx::launch_multipole_sphere_terms(ma, c->mpsph.p, c->stream));
                else
                    SX_HIP(c, sx::launch_multipole_cylinder_terms(ma, c->mpsph.p, c->stream));
                ma.sph = c->mpsph.p;
                // sphere up to degree 16, cylinder up to order 32: every
                // Bessel function of every (shell, atom, frame) once per call
                // when the table fits; else the two lowest (16 B each) within
                // the terms' budget
                int L1 = mp->plm_l1;
                if (mp->kind == 1) {
                    L1 = 1;
                    for (size_t k = 0; k + 1 < mp->moments.size(); k += 2) {
                        const int l = mp->moments[k], m = mp->moments[k + 1];
                        const int n = (l == 0 || m > 3 || m < 0) ? 0 : (m <= 1 ? 2 * l : 2 * l - 1);
                        L1 = std::max(L1, n + 1);
                    }
                }
                const int L1max = mp->kind == 0 ? sx::kBesselTableL1 : sx::kCylTableN1;
                // (the cylinder's table also holds two phases per (shell, atom, frame))
                const size_t tab_bytes = (sizeof(double) * (size_t)std::max(L1, 1) +
                                          (mp->kind == 1 ? 2 * sizeof(double2) : 0)) *
                                         (size_t)ns * c->NAp * c->NFp;
                const size_t j01_bytes = sizeof(double2) * (size_t)ns * c->NAp * c->NFp;
                if (terms != 3 && L1 >= 1 && L1 <= L1max && bessel_table_fits(c, tab_bytes)) {
                    SX_HIP(c, c->mpjt.ensure(tab_bytes));
                    SX_HIP(c, sx::launch_multipole_bessel_table(ma, L1, c->mpjt.p, c->stream));
                    ma.jtab = c->mpjt.p;
                    ma.jl1 = L1;
                } else if (j01_bytes <= kSphereTermBudget) {
                    SX_HIP(c, c->mpj.ensure(j01_bytes));
                    SX_HIP(c, sx::launch_multipole_bessel01(ma, c->mpj.p, c->stream));
                    ma.cj01 = c->mpj.p;
                }
            }
            SX_HIP(c, sx::launch_multipole(ma, c->stream));
        } else if (oriented) {
            const sx::OrientArgs oa = oriented_args(c, ns, (int)NM, bptr, b_shell != nullptr, shape, split);
            int plain = 0;
            int urc = upload_tile_lines(c, ns, shape, &plain);
            if (urc != SX_OK) return urc;
            SX_HIP(c, sx::launch_amp_oriented(oa, c->stream));
            c->timing.tile_shape = shape + 1;
            c->timing.shell_progression = oa.ap ? 1 : 0;
            c->timing.shell_residuals = oa.ap == 2 ? 1 : 0;
            c->timing.shell_exact_runs = plain;
        } else {
            sx::AmpArgs aa;
            aa.coords = c->coords.as<float>();
            aa.NAp = c->NAp;
            aa.NFp = c->NFp;
            aa.NF = NF;
            aa.qrev = c->qrev.as<double>();
            aa.b = bptr;
            aa.brow = browp;
            aa.V = V;
            aa.vb = vb;
            aa.split = split;
            aa.amp = c->amp.as<float2>();
            SX_HIP(c, sx::launch_amp_all(aa, c->stream));
        }
        c->timing.amplitude_launches++;
        if (prof) SX_HIP(c, hipEventRecord(c->ev[1], c->stream));
        int chunks = (int)std::min<size_t>(NM, std::max<size_t>(1, (1024 + ns - 1) / ns));
        SX_HIP(c, c->apart.ensure(sizeof(double) * 4 * (size_t)ns * chunks));
        fa.chunks = chunks;
        fa.apart = c->apart.as<double>();
        if (dsp == SX_DSP_AUTOCORRELATE) {
            const sx::FftPlan &P = c->plan;
            SX_HIP(c, c->ppart.ensure(sizeof(double) * (size_t)ns * chunks * P.L));
            sx::CorrArgs ca{};
            ca.plan = P;
            ca.NF = NF;
            ca.NM = (int)NM;
            ca.nshells = ns;
            ca.chunks = chunks;
            ca.twL = c->twL.as<float2>();
            ca.twL64 = c->twL64.as<double2>();
            ca.twp = c->twp.as<float2>();
            ca.twp64 = c->twp64.as<double2>();
            ca.w = c->w.as<double2>();
            ca.ppart = c->ppart.as<double>();
            ca.apart = c->apart.as<double>();
            ca.amp = c->amp.as<float2>();
            ca.NFp = c->NFp;
            ca.split = split;
            ca.split_stride = split_stride;
            if (P.N1 > 1) {
                // rows of a long plan: per-row parts of a, summed per vector
                const int wpp = 2 * P.N1;
                const int rblocks = (int)std::max<size_t>(1, std::min<size_t>(256, NM / 256));
                SX_HIP(c, c->apair.ensure(sizeof(double2) * wpp * (size_t)ns * NM));
                SX_HIP(c, c->apart.ensure(sizeof(double) * 4 * (size_t)ns * std::max(chunks, rblocks)));
                ca.apart = c->apart.as<double>();
                SX_HIP(c, sx::launch_corr_rows(ca, false, c->apair.as<double2>(), c->stream));
                SX_HIP(c, sx::launch_self_a_reduce(c->apair.as<double2>(), (long)NM, wpp, ns, rblocks,
                                                   c->apart.as<double>(), c->stream));
                fa.apart = c->apart.as<double>();
                fa.achunks = rblocks;
            } else {
                SX_HIP(c, sx::launch_corr_all(ca, c->stream));
            }
            fa.ppart = c->ppart.as<double>();
        } else {
            SX_HIP(c, sx::launch_series_all(dsp == SX_DSP_SQUARE ? 1 : 2, c->amp.as<float2>(), c->NFp, split,
                                            split_stride, NF, (int)NM, ns, chunks, c->fqt.as<double>(),
                                            c->apart.as<double>(), norm, c->stream));
            fa.ppart = nullptr;
        }
        c->timing.correlate_launches++;
    } else if (dsp != SX_DSP_AUTOCORRELATE) {
        if (prof) SX_HIP(c, hipEventRecord(c->ev[1], c->stream));
        const long pairs = (long)c->NA * (long)NM;
        int chunks = (int)std::min<long>(pairs, std::max<long>(1, (1024 + ns - 1) / ns));
        SX_HIP(c, c->apart.ensure(sizeof(double) * 4 * (size_t)ns * chunks));
        SX_HIP(c, c->ppart.ensure(sizeof(double2) * (size_t)ns * chunks * NF));
        sx::CorrArgs ca{};
        ca.NF = NF;
        ca.NM = (int)NM;
        ca.nshells = ns;
        ca.chunks = chunks;
        ca.apart = c->apart.as<double>();
        ca.coords = c->coords.as<float>();
        ca.NAp = c->NAp;
        ca.NFp = c->NFp;
        ca.NA = (int)c->NA;
        ca.qrev = c->qrev.as<double>();
        ca.b = bptr;
        ca.brow_per_shell = b_shell ? 1 : 0;
        const int tiles = sx::series_tiles(NF);
        if (tiles > 1) SX_HIP(c, c->apair.ensure(sizeof(double2) * (size_t)tiles * ns * (size_t)pairs));
        SX_HIP(c, sx::launch_series_self(dsp == SX_DSP_SQUARE ? 1 : 2, ca, c->ppart.as<double2>(),
                                         c->fqt.as<double>(), tiles > 1 ? c->apair.as<double2>() : nullptr,
                                         c->stream));
        fa.chunks = chunks;
        if (tiles > 1) {
            const int rblocks = (int)std::max<long>(1, std::min<long>(256, pairs / 256));
            SX_HIP(c, c->apart.ensure(sizeof(double) * 4 * (size_t)ns * std::max(chunks, rblocks)));
            SX_HIP(c, sx::launch_self_a_reduce(c->apair.as<double2>(), pairs, tiles, ns, rblocks,
                                               c->apart.as<double>(), c->stream));
            fa.achunks = rblocks;
        }
        c->timing.correlate_launches++;
        fa.apart = c->apart.as<double>();
        fa.ppart = nullptr;
    } else {
        // atoms whose length is zero in every shell of the call add nothing
        // to any sum; the transforms run over the others only
        c->aidx_h.clear();
        for (size_t j = 0; j < c->NA; ++j) {
            bool nz = false;
            if (b_shell) {
                for (int s = 0; s < ns && !nz; ++s) nz = (float)b_shell[(size_t)s * c->NA + j] != 0.f;
            } else {
                nz = c->b_h[j] != 0.f;  // the kernels read the float factors
            }
            if (nz) c->aidx_h.push_back((int)j);
        }
        const bool all_atoms = c->aidx_h.size() == c->NA;
        if (!all_atoms && !c->aidx_h.empty()) {
            SX_HIP(c, c->aidx.ensure(sizeof(int) * c->aidx_h.size()));
            SX_HIP(c, hipMemcpyAsync(c->aidx.p, c->aidx_h.data(), sizeof(int) * c->aidx_h.size(),
                                     hipMemcpyHostToDevice, c->stream));
        }
        if (prof) SX_HIP(c, hipEventRecord(c->ev[1], c->stream));
        const long pairs = (long)c->aidx_h.size() * (long)NM;
        int chunks = (int)std::max<long>(1, std::min<long>(pairs, std::max<long>(1, (2048 + ns - 1) / ns)));
        const sx::FftPlan &P = c->plan;
        const int rblocks = (int)std::max<long>(1, std::min<long>(256, pairs / 256));
        SX_HIP(c, c->apart.ensure(sizeof(double) * 4 * (size_t)ns * std::max(chunks, rblocks)));
        SX_HIP(c, c->ppart.ensure(sizeof(double) * (size_t)ns * chunks * P.L));
        sx::CorrArgs ca{};
        ca.plan = P;
        ca.NF = NF;
        ca.NM = (int)NM;
        ca.nshells = ns;
        ca.chunks = chunks;
        ca.twL = c->twL.as<float2>();
        ca.twp = c->twp.as<float2>();
        ca.w = c->w.as<double2>();
        ca.ppart = c->ppart.as<double>();
        ca.apart = c->apart.as<double>();
        ca.coords = c->coords.as<float>();
        ca.NAp = c->NAp;
        ca.NFp = c->NFp;
        ca.NA = (int)c->aidx_h.size();
        ca.aidx = all_atoms ? nullptr : c->aidx.as<int>();
        ca.qrev = c->qrev.as<double>();
        ca.b = bptr;
        ca.brow_per_shell = b_shell ? 1 : 0;
        // halves of a padded series: in different blocks (3, or 0 when the plan
        // allows and the per-pair parts of a fit kSplitBudgetBytes), both in one
        // sweep (2, or 0 otherwise), one after the other in one block (1)
        const int wpp = 2 * (P.T / 64);  // parts of a per pair: halves x wavefronts
        const size_t apair_bytes = sizeof(double2) * wpp * (size_t)ns * (size_t)pairs;
        int halves = c->cfg.fft_halves;
        if (halves == 0) halves = (sx::self_split_supported(P) && apair_bytes <= kSplitBudgetBytes) ? 3 : 2;
        if (halves == 3 && P.N1 == 1 && !sx::self_split_supported(P))
            return fail(c, SX_ERR_UNSUPPORTED, "fft_halves=3: no split kernel for this FFT plan (NF=%zu)", NF);
        if (P.N1 > 1) {
            // rows of a long plan (the halves always in different blocks):
            // one part of a per (pair, half, row), summed per pair
            const int rpp = 2 * P.N1;
            SX_HIP(c, c->apair.ensure(std::max<size_t>(sizeof(double2) * rpp * (size_t)ns * (size_t)pairs,
                                                       sizeof(double2))));
            SX_HIP(c, sx::launch_corr_rows(ca, true, c->apair.as<double2>(), c->stream));
            SX_HIP(c, sx::launch_self_a_reduce(c->apair.as<double2>(), pairs, rpp, ns, rblocks,
                                               c->apart.as<double>(), c->stream));
            fa.achunks = rblocks;
        } else if (halves == 3) {
            // Equal blocks, 2 ns chunks of them: a count that is not a multiple of
            // the blocks resident at once leaves a last round with the GPU mostly
            // idle (4,120 blocks on 1,024 slots ran 5 rounds for 4.02 of work).
            // Among half to twice the chunk count (at least one pair a chunk),
            // take the one with the fewest rounds per unit of work, ties (within
            // 0.5 %) to the count nearest the original.
            int per_cu = 0, cus = 0;
            SX_HIP(c, sx::self_split_blocks_per_cu(P, &per_cu));
            SX_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->cfg.device));
            const long resident = (long)std::max(per_cu, 1) * std::max(cus, 1);
            const long c0 = chunks;
            long pick = c0;
            double best = 0.0;
            for (long ch = std::max(1L, c0 / 2); ch <= std::min(pairs, 2 * c0); ++ch) {
                const double per = (double)((2L * ns * ch + resident - 1) / resident) / (double)ch;
                const bool closer = std::labs(ch - c0) < std::labs(pick - c0);
                if (best == 0.0 || per < best * 0.995 || (per <= best * 1.005 && closer)) {
                    if (per < best || best == 0.0) best = per;
                    pick = ch;
                }
            }
            if (pick != c0) {
                chunks = (int)pick;
                ca.chunks = chunks;
                SX_HIP(c, c->apart.ensure(sizeof(double) * 4 * (size_t)ns * std::max(chunks, rblocks)));
                SX_HIP(c, c->ppart.ensure(sizeof(double) * (size_t)ns * chunks * P.L));
                ca.ppart = c->ppart.as<double>();
                ca.apart = c->apart.as<double>();
            }
            SX_HIP(c, c->apair.ensure(std::max<size_t>(apair_bytes, sizeof(double2))));
            SX_HIP(c, sx::launch_corr_self_split(ca, c->apair.as<double2>(), c->stream));
            SX_HIP(c, sx::launch_self_a_reduce(c->apair.as<double2>(), pairs, wpp, ns, rblocks,
                                               c->apart.as<double>(), c->stream));
            fa.achunks = rblocks;
        } else {
            SX_HIP(c, sx::launch_corr_self(ca, halves, c->stream));
        }
        // the last lags by direct FP64 sums
        const int tblocks = (int)std::max<long>(1, std::min<long>(1024, (pairs + 63) / 64));
        SX_HIP(c, c->tail.ensure(sizeof(double2) * (size_t)ns * tblocks * sx::kTailLags));
        SX_HIP(c, sx::launch_self_tail(ca, c->tail.as<double2>(), tblocks, c->stream));
        fa.tail = c->tail.as<double2>();
        fa.tail_k = (int)std::min<size_t>(sx::kTailLags, NF);
        fa.tail_blocks = tblocks;
        c->timing.correlate_launches++;
        fa.chunks = chunks;
        fa.apart = c->apart.as<double>();
        fa.ppart = c->ppart.as<double>();
    }
    if (prof) SX_HIP(c, hipEventRecord(c->ev[2], c->stream));
    fa.L = c->plan_ok ? c->plan.L : 0;
    fa.tw64 = c->tw64.as<double2>();
    SX_HIP(c, sx::launch_final(fa, c->stream));
    if (prof) SX_HIP(c, hipEventRecord(c->ev[3], c->stream));

    SX_HIP(c, hipMemcpyAsync(fqt_h, c->fqt.p, sizeof(double) * 2 * (size_t)ns * NF, hipMemcpyDeviceToHost,
                             c->stream));
    SX_HIP(c, hipMemcpyAsync(fq_h, c->fq.p, sizeof(double) * 2 * ns, hipMemcpyDeviceToHost, c->stream));
    SX_HIP(c, hipMemcpyAsync(fq2_h, c->fq2.p, sizeof(double) * 2 * ns, hipMemcpyDeviceToHost, c->stream));
    SX_HIP(c, hipStreamSynchronize(c->stream));
    if (prof) {
        if (self) {
            c->timing.amplitude_ms += elapsed(c->ev[1], c->ev[2]);
        } else {
            c->timing.amplitude_ms += elapsed(c->ev[0], c->ev[1]);
            c->timing.correlate_ms += elapsed(c->ev[1], c->ev[2]);
        }
        c->timing.finalize_ms += elapsed(c->ev[2], c->ev[3]);
        c->timing.total_ms += elapsed(c->ev[0], c->ev[3]);
    }
    c->timing.terms += (size_t)V * c->NA * c->NF;
    return SX_OK;
}

}  // namespace

extern "C" {

const char *sx_version(void) { return "sassena_amd 0.1.0 (gfx950)"; }

const char *sx_last_error(const sx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

size_t sx_pad_length(size_t NF) {
    sx::FftPlan p;
    return sx::make_fft_plan(NF, &p) ? (size_t)p.L : 0;
}

int sx_create(const sx_config *cfg, sx_ctx **out) {
    if (!cfg || !out) return SX_ERR_ARG;
    *out = nullptr;
    if (cfg->scattering_type != SX_SCATTERING_ALL && cfg->scattering_type != SX_SCATTERING_SELF) return SX_ERR_ARG;
    if (cfg->dsp_type < SX_DSP_AUTOCORRELATE || cfg->dsp_type > SX_DSP_PLAIN) return SX_ERR_ARG;
    if (cfg->dsp_method != SX_METHOD_FFT && cfg->dsp_method != SX_METHOD_DIRECT) return SX_ERR_ARG;
    if (cfg->fft_halves < 0 || cfg->fft_halves > 3) return SX_ERR_ARG;
    if (cfg->multipole_terms < 0 || cfg->multipole_terms > 3) return SX_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) return SX_ERR_HIP;
    if (hipSetDevice(cfg->device) != hipSuccess) return SX_ERR_HIP;
    sx_ctx *c = new (std::nothrow) sx_ctx();
    if (!c) return SX_ERR_NOMEM;
    c->cfg = *cfg;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SX_ERR_HIP;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            sx_destroy(c);
            return SX_ERR_HIP;
        }
    // the allocation-free entry point's scratch on this device (device current)
    if (scratch_address(cfg->device, &c->scratch) != hipSuccess || !c->scratch) {
        sx_destroy(c);
        return SX_ERR_HIP;
    }
    *out = c;
    return SX_OK;
}

void sx_destroy(sx_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf *d : {&c->coords, &c->b, &c->twL, &c->twL64, &c->twp, &c->twp64, &c->tw64, &c->w, &c->aidx, &c->apair, &c->tail, &c->qrev, &c->qlen, &c->uvec, &c->brow, &c->bshell, &c->mpmom, &c->mpq, &c->mpsph, &c->mpj, &c->mpjt, &c->qres, &c->redbuf, &c->amp,
                      &c->ppart, &c->apart, &c->fqt, &c->fq, &c->fq2, &c->stage, &c->stage2})
        d->release();
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"

namespace {

// Host-to-device staging through two pinned bounce buffers shared by the
// process (allocated once, kStageChunkFloats floats each, or one piece when
// larger): while the DMA drains one, host threads copy the next piece of the
// caller's block into the other.  Pieces are atom rows of one plane (SoA,
// copied with a 2D copy straight into the padded rows), frames (frames-AoS)
// or atoms (atoms-AoS), the AoS pieces transposed on the device from a
// staging buffer per bounce buffer.
struct PinPool {
    std::mutex mu;
    void *buf[2] = {nullptr, nullptr};
    size_t bytes = 0;
};
PinPool g_pins;

// copies n bytes on up to 8 host threads
void parallel_copy(void *dst, const void *src, size_t n) {
    const size_t kMin = (size_t)4 << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t parts = std::min<size_t>(std::min(8u, hw), std::max<size_t>(1, n / kMin));
    if (parts <= 1) {
        std::memcpy(dst, src, n);
        return;
    }
    std::vector<std::thread> th;
    const size_t per = (n + parts - 1) / parts;
    for (size_t k = 0; k < parts; ++k) {
        const size_t o = k * per;
        if (o >= n) break;
        const size_t len = std::min(per, n - o);
        th.emplace_back([=] { std::memcpy((char *)dst + o, (const char *)src + o, len); });
    }
    for (auto &t : th) t.join();
}

int stage_upload(sx_ctx *c, const float *xyz, size_t NA, size_t NF, int layout, float *dst, size_t plane) {
    struct Piece {
        const float *src;
        size_t floats;
        size_t k, first, count;  // plane (SoA), first row / frame / atom, how many
    };
    std::vector<Piece> pieces;
    size_t maxf = 0;
    if (layout == SX_LAYOUT_SOA) {
        const size_t per = std::max<size_t>(1, kStageChunkFloats / NF);
        for (size_t k = 0; k < 3; ++k)
            for (size_t a0 = 0; a0 < NA; a0 += per) {
                const size_t na = std::min(per, NA - a0);
                pieces.push_back({xyz + (k * NA + a0) * NF, na * NF, k, a0, na});
            }
    } else if (layout == SX_LAYOUT_FRAMES_AOS) {
        // the transpose grid covers 64 frames per y block (<= 65535)
        const size_t per = std::min<size_t>(std::max<size_t>(1, kStageChunkFloats / (3 * NA)), size_t(65535) * 64);
        for (size_t f0 = 0; f0 < NF; f0 += per) {
            const size_t nf = std::min(per, NF - f0);
            pieces.push_back({xyz + f0 * NA * 3, 3 * NA * nf, 0, f0, nf});
        }
    } else {
        // one y block per atom (<= 65535)
        const size_t per = std::min<size_t>(std::max<size_t>(1, kStageChunkFloats / (3 * NF)), size_t(65535));
        for (size_t a0 = 0; a0 < NA; a0 += per) {
            const size_t na = std::min(per, NA - a0);
            pieces.push_back({xyz + a0 * NF * 3, 3 * NF * na, 0, a0, na});
        }
    }
    for (const Piece &p : pieces) maxf = std::max(maxf, p.floats);
    const size_t bytes = sizeof(float) * std::max(maxf, kStageChunkFloats);
    std::lock_guard<std::mutex> lock(g_pins.mu);
    if (g_pins.bytes < bytes) {
        for (void *&b : g_pins.buf) {
            if (b) (void)hipHostFree(b);
            b = nullptr;
        }
        g_pins.bytes = 0;
        for (void *&b : g_pins.buf) SX_HIP(c, hipHostMalloc(&b, bytes, hipHostMallocDefault));
        g_pins.bytes = bytes;
    }
    const bool aos = layout != SX_LAYOUT_SOA;
    if (aos) {
        SX_HIP(c, c->stage.ensure(sizeof(float) * maxf));
        SX_HIP(c, c->stage2.ensure(sizeof(float) * maxf));
    }
    hipEvent_t done[2] = {nullptr, nullptr};
    SX_HIP(c, hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
    SX_HIP(c, hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
    int rc = SX_OK;
    for (size_t i = 0; i < pieces.size() && rc == SX_OK; ++i) {
        const Piece &p = pieces[i];
        const int bb = (int)(i & 1);
        // the DMA (and transpose) of piece i - 2 has drained this bounce buffer
        if (i >= 2) {
            hipError_t e = hipEventSynchronize(done[bb]);
            if (e != hipSuccess) {
                rc = fail(c, SX_ERR_HIP, "sx_stage: %s", hipGetErrorString(e));
                break;
            }
        }
        parallel_copy(g_pins.buf[bb], p.src, sizeof(float) * p.floats);
        hipError_t e = hipSuccess;
        if (!aos) {
            e = hipMemcpy2DAsync(dst + p.k * plane + p.first * c->NFp, sizeof(float) * c->NFp, g_pins.buf[bb],
                                 sizeof(float) * NF, sizeof(float) * NF, p.count, hipMemcpyHostToDevice, c->stream);
        } else {
            float *stg = bb ? c->stage2.as<float>() : c->stage.as<float>();
            e = hipMemcpyAsync(stg, g_pins.buf[bb], sizeof(float) * p.floats, hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess)
                e = layout == SX_LAYOUT_FRAMES_AOS
                        ? sx::launch_stage_frames(stg, NA, p.first, p.count, dst, c->NAp, c->NFp, c->stream)
                        : sx::launch_stage_atoms(stg, p.first, p.count, NF, dst, c->NAp, c->NFp, c->stream);
        }
        if (e == hipSuccess) e = hipEventRecord(done[bb], c->stream);
        if (e != hipSuccess) rc = fail(c, SX_ERR_HIP, "sx_stage: %s", hipGetErrorString(e));
    }
    // the bounce buffers are free again once the stream has drained
    hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipEventDestroy(done[0]);
    (void)hipEventDestroy(done[1]);
    if (rc == SX_OK && e != hipSuccess) rc = fail(c, SX_ERR_HIP, "sx_stage: %s", hipGetErrorString(e));
    return rc;
}

}  // namespace

extern "C" {

int sx_stage(sx_ctx *c, const float *xyz, size_t NA, size_t NF, int layout) {
    if (!c) return SX_ERR_ARG;
    if (!xyz || NA == 0 || NF == 0) return fail(c, SX_ERR_ARG, "sx_stage: empty coordinate block");
    if (NA > (size_t)1 << 28 || NF > (size_t)1 << 24)
        return fail(c, SX_ERR_ARG, "sx_stage: block too large (NA=%zu NF=%zu)", NA, NF);
    if (layout != SX_LAYOUT_FRAMES_AOS && layout != SX_LAYOUT_ATOMS_AOS && layout != SX_LAYOUT_SOA)
        return fail(c, SX_ERR_ARG, "sx_stage: unknown layout %d", layout);
    SX_HIP(c, hipSetDevice(c->cfg.device));
    c->staged = false;
    c->have_b = false;
    c->NA = NA;
    c->NF = NF;
    c->NAp = round_up(NA, 64);
    c->NFp = round_up(NF, sx::kWave);
    const size_t plane = c->NAp * c->NFp;
    const size_t block = sizeof(float) * 3 * plane;
    // the block must fit the free device memory (beyond the allocation it
    // reuses), with room for the FFT tables
    size_t free_b = 0, total_b = 0;
    SX_HIP(c, hipMemGetInfo(&free_b, &total_b));
    const size_t need = (block > c->coords.cap ? block - c->coords.cap : 0) + ((size_t)64 << 20);
    if (need > free_b)
        return fail(c, SX_ERR_NOMEM,
                    "sx_stage: %zu atoms x %zu frames need %zu bytes of device memory, %zu of %zu are free", NA, NF,
                    block, free_b, total_b);
    SX_HIP(c, c->coords.ensure(block));  // kept when NA and NF need no more room than before
    SX_HIP(c, hipMemsetAsync(c->coords.p, 0, block, c->stream));
    float *dst = c->coords.as<float>();
    int rc = stage_upload(c, xyz, NA, NF, layout, dst, plane);
    if (rc != SX_OK) return rc;
    // largest |coordinate|: |d| = |u . r| <= sqrt(3) of it bounds the residual phases
    SX_HIP(c, c->redbuf.ensure(sizeof(unsigned)));
    SX_HIP(c, sx::launch_absmax(dst, 3 * plane, c->redbuf.as<unsigned>(), c->stream));
    unsigned bits = 0;
    SX_HIP(c, hipMemcpyAsync(&bits, c->redbuf.p, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    SX_HIP(c, hipStreamSynchronize(c->stream));
    float amax = 0.f;
    std::memcpy(&amax, &bits, sizeof amax);
    c->rmax = std::sqrt(3.0) * (double)amax;
    if (!std::isfinite(c->rmax)) return fail(c, SX_ERR_ARG, "sx_stage: non-finite coordinates");
    c->stage.release();
    c->stage2.release();
    rc = prepare_tables(c);
    if (rc != SX_OK) return rc;
    c->staged = true;
    return SX_OK;
}

int sx_set_factors(sx_ctx *c, const double *b, size_t NA) {
    if (!c) return SX_ERR_ARG;
    if (!c->staged) return fail(c, SX_ERR_STATE, "sx_set_factors: stage coordinates first");
    if (!b || NA != c->NA) return fail(c, SX_ERR_ARG, "sx_set_factors: expected %zu factors, got %zu", c->NA, NA);
    SX_HIP(c, hipSetDevice(c->cfg.device));
    std::vector<float> bh(c->NAp, 0.f);
    for (size_t j = 0; j < NA; ++j) bh[j] = (float)b[j];
    SX_HIP(c, c->b.ensure(sizeof(float) * c->NAp));
    SX_HIP(c, hipMemcpyAsync(c->b.p, bh.data(), sizeof(float) * c->NAp, hipMemcpyHostToDevice, c->stream));
    SX_HIP(c, hipStreamSynchronize(c->stream));
    c->b_h = std::move(bh);
    c->have_b = true;
    return SX_OK;
}

}  // extern "C"

namespace {

bool needs_fft_rows(const sx_ctx *c) {
    return c->cfg.dsp_type == SX_DSP_AUTOCORRELATE && c->plan_ok && c->plan.N1 > 1;
}

int check_compute_args(sx_ctx *c, size_t nshells, const double *qv, const double *qlen, const double *u, size_t NM,
                       const double *b_shell, double *fqt, double *fq, double *fq2) {
    if (!c) return SX_ERR_ARG;
    if (!c->staged) return fail(c, SX_ERR_STATE, "compute: no staged coordinates");
    if (!b_shell && !c->have_b) return fail(c, SX_ERR_STATE, "compute: scattering factors not set");
    if (nshells == 0 || NM == 0 || (!qv && (!qlen || !u)) || !fqt || !fq || !fq2)
        return fail(c, SX_ERR_ARG, "compute: empty shell list or null output");
    if (NM > (size_t)1 << 20 || nshells > (size_t)1 << 20)
        return fail(c, SX_ERR_ARG, "compute: too many vectors (NM=%zu, shells=%zu)", NM, nshells);
    const bool needs_fft = c->cfg.dsp_type == SX_DSP_AUTOCORRELATE;
    if (needs_fft && !c->plan_ok)
        return fail(c, SX_ERR_UNSUPPORTED, "compute: NF=%zu exceeds the supported correlation length", c->NF);
    if (qv) {
        for (size_t i = 0; i < 3 * nshells * NM; ++i)
            if (!std::isfinite(qv[i])) return fail(c, SX_ERR_ARG, "compute: non-finite q vector component");
    } else {
        for (size_t i = 0; i < nshells; ++i)
            if (!std::isfinite(qlen[i])) return fail(c, SX_ERR_ARG, "compute: non-finite |q|");
        for (size_t i = 0; i < 3 * NM; ++i)
            if (!std::isfinite(u[i])) return fail(c, SX_ERR_ARG, "compute: non-finite orientation component");
    }
    return SX_OK;
}

int compute_impl(sx_ctx *c, size_t nshells, const double *qv, const double *qlen, const double *u, size_t NM,
                 const double *b_shell, double *fqt, double *fq, double *fq2) {
    int rc = check_compute_args(c, nshells, qv, qlen, u, NM, b_shell, fqt, fq, fq2);
    if (rc != SX_OK) return rc;
    SX_HIP(c, hipSetDevice(c->cfg.device));
    c->timing = sx_timing{};
    c->timing.pad_length = c->plan_ok ? (size_t)c->plan.L : 0;
    // batch shells so that the amplitude buffer stays within budget
    size_t per = nshells;
    if (c->cfg.scattering_type == SX_SCATTERING_ALL) {
        const size_t per_shell = sizeof(float2) * NM * c->NFp * 16;  // worst-case split
        per = std::max<size_t>(1, std::min(nshells, kAmpBudgetBytes / std::max<size_t>(per_shell, 1)));
    } else if (needs_fft_rows(c)) {
        // per-(pair, half, row) parts of a of a long plan within kSplitBudgetBytes
        const size_t per_shell = sizeof(double2) * 2 * c->plan.N1 * c->NA * NM;
        per = std::max<size_t>(1, std::min(nshells, kSplitBudgetBytes / std::max<size_t>(per_shell, 1)));
    }
    for (size_t s0 = 0; s0 < nshells; s0 += per) {
        const int ns = (int)std::min(per, nshells - s0);
        rc = run_batch(c, ns, NM, qv ? qv + 3 * NM * s0 : nullptr, qlen ? qlen + s0 : nullptr, u,
                       b_shell ? b_shell + c->NA * s0 : nullptr, fqt + 2 * c->NF * s0, fq + 2 * s0, fq2 + 2 * s0);
        if (rc != SX_OK) return rc;
    }
    return SX_OK;
}

// Device bytes run_batch takes for ns oriented coherent shells in an
// allocation-free call (each buffer 256-byte aligned in the arena): the
// tile shape and atom ranges are those run_batch picks without autotuning
// and without round-filling ranges.
size_t noalloc_batch_bytes(sx_ctx *c, int ns, const double *qlen, size_t NM, bool per_shell) {
    std::vector<double> ql(ns);
    for (int s = 0; s < ns; ++s) ql[s] = qlen[s] / kTwoPi;
    sx::ShellGrid g;
    int mode = sx::shell_progression(ql.data(), ns, c->rmax, &g);
    if (mode == 2 && per_shell) mode = 0;
    const int saved = c->q_mode;
    c->q_mode = mode;
    const int shape = c->cfg.tile_shape > 0 ? c->cfg.tile_shape - 1 : default_shape(c, ns, per_shell);
    c->q_mode = saved;
    const size_t V = (size_t)ns * NM;
    const size_t split = (size_t)pick_split_blocks(c, sx::oriented_base_blocks(c->NFp, (int)NM, ns, shape));
    const size_t chunks = std::min<size_t>(NM, std::max<size_t>(1, (1024 + ns - 1) / ns));
    const bool fft = c->cfg.dsp_type == SX_DSP_AUTOCORRELATE;
    const size_t parts[] = {
        sizeof(double) * 3 * V,                                   // qrev
        sizeof(double) * ns,                                      // qlen
        sizeof(double) * 3 * NM,                                  // uvec
        mode == 2 ? sx::shell_table_bytes(ns) : 0,                // qres (runs and residuals)
        per_shell ? sizeof(float) * ns * c->NAp : 0,              // bshell
        per_shell ? sizeof(int) * V : 0,                          // brow
        sizeof(float2) * V * c->NFp * split,                      // amp
        sizeof(double) * 4 * ns * chunks,                         // apart
        fft ? sizeof(double) * ns * chunks * c->plan.L : 0,       // ppart
        sizeof(double) * 2 * ns * c->NF,                          // fqt
        sizeof(double) * 2 * ns, sizeof(double) * 2 * ns,         // fq, fq2
    };
    size_t total = 0;
    for (size_t b : parts) total += round_up(b, 256) + 256;
    return total;
}

}  // namespace

extern "C" {

int sx_compute_shells_oriented_noalloc(sx_ctx *c, size_t nshells, const double *qlen, const double *u, size_t NM,
                                       const double *b_shell, double *fqt, double *fq, double *fq2) {
    if (!c) return SX_ERR_ARG;
    if (c->cfg.scattering_type != SX_SCATTERING_ALL)
        return fail(c, SX_ERR_UNSUPPORTED, "sx_compute_shells_oriented_noalloc: coherent (all) scattering only");
    if (!qlen || !u) return fail(c, SX_ERR_ARG, "sx_compute_shells_oriented_noalloc: null |q| or orientations");
    int rc = check_compute_args(c, nshells, nullptr, qlen, u, NM, b_shell, fqt, fq, fq2);
    if (rc != SX_OK) return rc;
    if (!c->scratch) return fail(c, SX_ERR_STATE, "sx_compute_shells_oriented_noalloc: no scratch on this device");
    SX_HIP(c, hipSetDevice(c->cfg.device));
    c->timing = sx_timing{};
    c->timing.pad_length = c->plan_ok ? (size_t)c->plan.L : 0;
    // the call's work buffers become slices of this device's scratch; the
    // context's own buffers are set aside and restored afterwards
    Arena arena;
    arena.base = static_cast<char *>(c->scratch);
    arena.cap = kScratchBytes;
    DevBuf *work[] = {&c->qrev, &c->qlen, &c->uvec, &c->qres, &c->bshell, &c->brow,
                      &c->amp,  &c->apart, &c->ppart, &c->fqt, &c->fq, &c->fq2};
    constexpr size_t kWork = sizeof(work) / sizeof(work[0]);
    DevBuf saved[kWork];
    for (size_t i = 0; i < kWork; ++i) {
        saved[i] = *work[i];
        *work[i] = DevBuf{};
        work[i]->arena = &arena;
    }
    c->noalloc = true;
    const bool per_shell = b_shell != nullptr;
    size_t s0 = 0;
    while (rc == SX_OK && s0 < nshells) {
        // the most shells from s0 whose buffers fit the scratch
        size_t ns = nshells - s0;
        while (ns > 1 && noalloc_batch_bytes(c, (int)ns, qlen + s0, NM, per_shell) > kScratchBytes) ns = (ns + 1) / 2;
        const size_t need = noalloc_batch_bytes(c, (int)ns, qlen + s0, NM, per_shell);
        if (need > kScratchBytes) {
            rc = fail(c, SX_ERR_NOMEM,
                      "sx_compute_shells_oriented_noalloc: one shell of %zu orientations x %zu frames needs %zu "
                      "bytes of work memory, above the %zu MiB scratch maximum",
                      NM, c->NF, need, kScratchBytes >> 20);
            break;
        }
        arena.used = 0;
        arena.exhausted = false;
        for (size_t i = 0; i < kWork; ++i) {
            work[i]->p = nullptr;
            work[i]->cap = 0;
        }
        rc = run_batch(c, (int)ns, NM, nullptr, qlen + s0, u, per_shell ? b_shell + c->NA * s0 : nullptr,
                       fqt + 2 * c->NF * s0, fq + 2 * s0, fq2 + 2 * s0);
        if (rc != SX_OK && arena.exhausted)
            rc = fail(c, SX_ERR_NOMEM, "sx_compute_shells_oriented_noalloc: work memory above the %zu MiB scratch",
                      kScratchBytes >> 20);
        s0 += ns;
    }
    c->noalloc = false;
    for (size_t i = 0; i < kWork; ++i) *work[i] = saved[i];
    return rc;
}


int sx_compute_shells(sx_ctx *c, size_t nshells, const double *qv, size_t NM, const double *b_shell, double *fqt,
                      double *fq, double *fq2) {
    if (c && !qv) return fail(c, SX_ERR_ARG, "sx_compute_shells: null q vectors");
    return compute_impl(c, nshells, qv, nullptr, nullptr, NM, b_shell, fqt, fq, fq2);
}

int sx_compute_shells_oriented(sx_ctx *c, size_t nshells, const double *qlen, const double *u, size_t NM,
                               const double *b_shell, double *fqt, double *fq, double *fq2) {
    if (c && (!qlen || !u)) return fail(c, SX_ERR_ARG, "sx_compute_shells_oriented: null |q| or orientations");
    return compute_impl(c, nshells, nullptr, qlen, u, NM, b_shell, fqt, fq, fq2);
}

int sx_compute_shell(sx_ctx *c, const double *qv, size_t NM, double *fqt, double *fq, double *fq2) {
    return sx_compute_shells(c, 1, qv, NM, nullptr, fqt, fq, fq2);
}

int sx_compute_multipole(sx_ctx *c, int kind, size_t nshells, const double *qv, const double *axis,
                         const int *moments, size_t nmo, const double *b_shell, double *fqt, double *fq,
                         double *fq2) {
    if (!c) return SX_ERR_ARG;
    if (c->cfg.scattering_type != SX_SCATTERING_ALL)
        return fail(c, SX_ERR_UNSUPPORTED, "multipole averaging is defined for scattering.type=all");
    if (kind != SX_MULTIPOLE_SPHERE && kind != SX_MULTIPOLE_CYLINDER)
        return fail(c, SX_ERR_ARG, "unknown multipole type %d", kind);
    if (!c->staged) return fail(c, SX_ERR_STATE, "compute: no staged coordinates");
    if (!b_shell && !c->have_b) return fail(c, SX_ERR_STATE, "compute: scattering factors not set");
    if (nshells == 0 || nshells > 65535 || !qv || !moments || nmo == 0 || nmo > 100000 || !fqt || !fq || !fq2)
        return fail(c, SX_ERR_ARG, "compute: bad shell list, moment list or output");
    if (kind == SX_MULTIPOLE_CYLINDER && !axis) return fail(c, SX_ERR_ARG, "cylinder multipole needs an axis");
    // moment checks of parameters.cpp:1078-1119
    for (size_t k = 0; k < nmo; ++k) {
        const int l = moments[2 * k], m = moments[2 * k + 1];
        if (l < 0 || l > 200) return fail(c, SX_ERR_ARG, "moment %zu: major %d out of range", k, l);
        if (kind == SX_MULTIPOLE_SPHERE && (m < -l || m > l))
            return fail(c, SX_ERR_ARG, "moment %zu: minor %d not in [-%d, %d]", k, m, l, l);
        if (kind == SX_MULTIPOLE_CYLINDER && (m < 0 || m > 3 || (l == 0 && m != 0)))
            return fail(c, SX_ERR_ARG, "moment %zu: cylinder minor %d invalid for major %d", k, m, l);
    }
    if (c->cfg.dsp_type == SX_DSP_AUTOCORRELATE && !c->plan_ok)
        return fail(c, SX_ERR_UNSUPPORTED, "compute: NF=%zu exceeds the supported correlation length", c->NF);
    for (size_t i = 0; i < 3 * nshells; ++i)
        if (!std::isfinite(qv[i])) return fail(c, SX_ERR_ARG, "compute: non-finite q vector component");
    SX_HIP(c, hipSetDevice(c->cfg.device));
    c->timing = sx_timing{};
    c->timing.pad_length = c->plan_ok ? (size_t)c->plan.L : 0;
    MPSpec mp;
    mp.kind = kind;
    mp.moments.assign(moments, moments + 2 * nmo);
    if (kind == SX_MULTIPOLE_SPHERE) {
        mp.ynorm.resize(2 * nmo);
        for (size_t k = 0; k < nmo; ++k) {
            const int l = moments[2 * k], am = std::abs(moments[2 * k + 1]);
            double cm = (2.0 * am + 1.0) / (4.0 * M_PI);
            for (int i = 1; i <= am; ++i) cm *= (2.0 * i - 1.0) / (2.0 * i);
            const double lf = 0.5 * (std::log((2.0 * l + 1.0) / (2.0 * am + 1.0)) + std::lgamma(l - am + 1.0) +
                                     std::lgamma(2.0 * am + 1.0) - std::lgamma(l + am + 1.0));
            mp.ynorm[2 * k] = std::sqrt(cm);
            mp.ynorm[2 * k + 1] = 4.0 * M_PI * std::exp(lf);
        }
        int lmax = 0;
        for (size_t k = 0; k < nmo; ++k) lmax = std::max(lmax, moments[2 * k]);
        mp.plm_l1 = lmax + 1;
        const size_t base = mp.ynorm.size();
        mp.ynorm.resize(base + 2 * (size_t)mp.plm_l1 * mp.plm_l1, 0.0);
        for (int ll = 2; ll <= lmax; ++ll)
            for (int m = 0; m + 2 <= ll; ++m) {
                const size_t i = base + 2 * ((size_t)ll * mp.plm_l1 + m);
                mp.ynorm[i] = (2.0 * ll - 1.0) / (double)(ll - m);
                mp.ynorm[i + 1] = (ll + m - 1.0) / (double)(ll - m);
            }
    }
    for (double &x : mp.base) x = 0.0;
    if (kind == SX_MULTIPOLE_CYLINDER) {
        // CartesianVectorBase(axis) (coor3d.cpp:278-296)
        const double al = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
        if (!(al > 0)) return fail(c, SX_ERR_ARG, "cylinder axis has zero length");
        const double ez[3] = {axis[0] / al, axis[1] / al, axis[2] / al};
        double ek[3] = {-ez[1], ez[0], 0.0};  // (0,0,1) x ez
        double len = std::sqrt(ek[0] * ek[0] + ek[1] * ek[1] + ek[2] * ek[2]);
        if (len == 0) {
            ek[0] = ez[2];  // (0,1,0) x ez
            ek[1] = 0.0;
            ek[2] = -ez[0];
            len = std::sqrt(ek[0] * ek[0] + ek[1] * ek[1] + ek[2] * ek[2]);
        }
        const double er[3] = {ek[0] / len, ek[1] / len, ek[2] / len};
        const double ezer[3] = {ez[1] * er[2] - ez[2] * er[1], ez[2] * er[0] - ez[0] * er[2],
                                ez[0] * er[1] - ez[1] * er[0]};
        const double l2 = std::sqrt(ezer[0] * ezer[0] + ezer[1] * ezer[1] + ezer[2] * ezer[2]);
        for (int k = 0; k < 3; ++k) {
            mp.base[k] = er[k];
            mp.base[3 + k] = ezer[k] / l2;
            mp.base[6 + k] = ez[k];
        }
    }
    const size_t per = std::max<size_t>(1, kAmpBudgetBytes / (sizeof(float2) * nmo * c->NFp));
    for (size_t s0 = 0; s0 < nshells; s0 += per) {
        const int ns = (int)std::min(per, nshells - s0);
        mp.qinfo.clear();
        for (int s = 0; s < ns; ++s) {
            const double *q = qv + 3 * (s0 + s);
            if (kind == SX_MULTIPOLE_SPHERE) {
                mp.qinfo.push_back(std::sqrt(std::pow(q[0], 2) + std::pow(q[1], 2) + std::pow(q[2], 2)));
            } else {
                // q projected on the frame, as CylinderCoor3D (coor3d.cpp:116-140)
                const double px = q[0] * mp.base[0] + q[1] * mp.base[1] + q[2] * mp.base[2];
                const double py = q[0] * mp.base[3] + q[1] * mp.base[4] + q[2] * mp.base[5];
                const double pz = q[0] * mp.base[6] + q[1] * mp.base[7] + q[2] * mp.base[8];
                double phi;
                if (px != 0.0) {
                    phi = std::atan(py / px);
                    if (px < 0.0) phi = (double)(py < 0.0 ? -(float)M_PI : (float)M_PI) + phi;
                } else if (py != 0.0) {
                    phi = (double)(py < 0.0 ? -(float)(M_PI / 2) : (float)(M_PI / 2));
                } else {
                    phi = 0.0;
                }
                phi = phi < 0 ? 2 * M_PI + phi : phi;
                mp.qinfo.push_back(std::sqrt(std::pow(px, 2) + std::pow(py, 2)));
                mp.qinfo.push_back(phi);
                mp.qinfo.push_back(pz);
            }
        }
        int rc = run_batch(c, ns, nmo, nullptr, nullptr, nullptr, b_shell ? b_shell + c->NA * s0 : nullptr,
                           fqt + 2 * c->NF * s0, fq + 2 * s0, fq2 + 2 * s0, &mp);
        if (rc != SX_OK) return rc;
    }
    return SX_OK;
}

int sx_get_timing(const sx_ctx *c, sx_timing *out) {
    if (!c || !out) return SX_ERR_ARG;
    *out = c->timing;
    return SX_OK;
}

}  // extern "C"
