// sx_amp.hip -- staging transposes and the coherent amplitude kernel.
//
// Amplitude ("all" scattering, reference loop all_vectors_scatter_device.cpp:418-438):
//   A_v(t) = sum_j b_j exp(i q_v . r_j(t))
// Each lane of a wavefront owns one frame t and each wavefront a block of VB
// q vectors (wave-uniform, held in SGPRs), so the atom sum needs no cross-lane
// reduction and every coordinate load is a coalesced 256-byte row of the
// atom-major SoA block.  The phase is formed in FP64 in turns (q pre-divided
// by 2 pi), reduced to [0, 1) turns in FP64 (v_fract_f64), and evaluated with
// the hardware v_sin_f32 / v_cos_f32, which take turns (max abs error 1.3e-7
// over the reduced range, measured against FP64 on gfx950).  Sums run in FP32
// over 64 atoms and are folded into FP64 totals.
//
// When the vectors of all shells are one orientation set scaled by |q_s|
// (orientation types file and sphere, abstract_vectors_scatter_device.cpp:115-129),
// amp_oriented_kernel forms d = u_m . r once per atom and orientation and the
// phase of each shell as the single FP64 product (|q_s| / 2 pi) d, or, with
// linearly spaced shells, steps the phasor from shell to shell (sx_amp_device.h).
#include <algorithm>
#include <cmath>

#include "sx_amp_device.h"

namespace sx {

// ---------------------------------------------------------------- staging --

// src: frames [nf][NA][3] (frames F0.. of the host block); tile 64 atoms x 64 frames.
__global__ __launch_bounds__(256) void stage_frames_kernel(const float *__restrict__ src, int NA,
                                                           int F0, int nf, float *__restrict__ coords,
                                                           size_t NAp, size_t NFp) {
    __shared__ float tile[3][64][65];
    const int a0 = blockIdx.x * 64, f0 = blockIdx.y * 64;
    const int tid = threadIdx.x;
    for (int tf = 0; tf < 64; ++tf) {
        const int f = f0 + tf;
        if (tid < 192 && f < nf) {
            const int ja = tid / 3, c = tid % 3;
            const int a = a0 + ja;
            if (a < NA) tile[c][ja][tf] = src[((size_t)f * NA + a) * 3 + c];
        }
    }
    __syncthreads();
    const int tf = tid & 63;
    for (int row = tid >> 6; row < 192; row += 4) {
        const int c = row / 64, ja = row % 64;
        const int a = a0 + ja, f = f0 + tf;
        if (a < NA && f < nf) coords[(size_t)c * NAp * NFp + (size_t)a * NFp + F0 + f] = tile[c][ja][tf];
    }
}

// src: atoms [na][NF][3] (atoms A0.. of the host block).
__global__ __launch_bounds__(256) void stage_atoms_kernel(const float *__restrict__ src, int A0, int na,
                                                          int NF, float *__restrict__ coords, size_t NAp,
                                                          size_t NFp) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int ja = blockIdx.y;
    if (t >= NF || ja >= na) return;
    const float *p = src + ((size_t)ja * NF + t) * 3;
    const size_t row = (size_t)(A0 + ja) * NFp + t;
    coords[row] = p[0];
    coords[NAp * NFp + row] = p[1];
    coords[2 * NAp * NFp + row] = p[2];
}

hipError_t launch_stage_frames(const float *src, size_t NA, size_t F0, size_t nf, float *coords, size_t NAp,
                               size_t NFp, hipStream_t stream) {
    dim3 grid((unsigned)((NA + 63) / 64), (unsigned)((nf + 63) / 64));
    hipLaunchKernelGGL(stage_frames_kernel, grid, dim3(256), 0, stream, src, (int)NA, (int)F0, (int)nf, coords,
                       NAp, NFp);
    return hipGetLastError();
}

hipError_t launch_stage_atoms(const float *src, size_t a0, size_t na, size_t NF, float *coords, size_t NAp,
                              size_t NFp, hipStream_t stream) {
    dim3 grid((unsigned)((NF + 255) / 256), (unsigned)na);
    hipLaunchKernelGGL(stage_atoms_kernel, grid, dim3(256), 0, stream, src, (int)a0, (int)na, (int)NF, coords,
                       NAp, NFp);
    return hipGetLastError();
}

// Largest |x| over n floats, as the bit pattern of a non-negative float
// (ordered like unsigned integers) in *out, which the caller zeroes first.
__global__ __launch_bounds__(256) void absmax_kernel(const float *__restrict__ x, size_t n,
                                                     unsigned *__restrict__ out) {
    float m = 0.f;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        m = fmaxf(m, fabsf(x[i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

__global__ void zero_u32_kernel(unsigned *p) { *p = 0u; }

hipError_t launch_absmax(const float *x, size_t n, unsigned *out, hipStream_t stream) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, stream, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t blocks = std::min<size_t>(4096, std::max<size_t>(1, (n + 255) / 256));
    hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, x, n, out);
    return hipGetLastError();
}

// -------------------------------------------------------------- amplitude --

struct AmpKernelArgs {
    const float *X, *Y, *Z;
    size_t NFp;
    const double *qrev;
    const float *b;
    const int *brow;
    size_t NAp;
    int V, ftiles, vgroups, split, atoms_per_split;
    float2 *amp;
};

template <int VB, bool PER_ROW_B>
__global__ __launch_bounds__(256) void amp_all_kernel(AmpKernelArgs a) {
    // XCD-aware work order: hardware hands consecutive workgroups to the 8
    // XCDs round-robin; give each XCD a contiguous run of items so that the
    // wavefronts sharing one frame tile and atom range share that XCD's L2.
    const int nitems = a.ftiles * a.split * a.vgroups;
    const int per = (nitems + 7) >> 3;
    const int item = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (item >= nitems) return;
    const int vg = item % a.vgroups;
    const int rest = item / a.vgroups;
    const int sp = rest % a.split;
    const int ft = rest / a.split;

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int vbase = (vg * 4 + wave) * VB;
    const size_t t = (size_t)ft * kWave + lane;

    double qx[VB], qy[VB], qz[VB];
    int brow[VB];
#pragma unroll
    for (int i = 0; i < VB; ++i) {
        const int v = min(vbase + i, a.V - 1);
        qx[i] = a.qrev[3 * v];
        qy[i] = a.qrev[3 * v + 1];
        qz[i] = a.qrev[3 * v + 2];
        brow[i] = PER_ROW_B ? a.brow[v] : 0;
    }

    // FP32 sums over kFold atoms, folded into FP64 totals
    constexpr int kFold = 64;
    double tr[VB], ti[VB];
#pragma unroll
    for (int i = 0; i < VB; ++i) tr[i] = ti[i] = 0.0;

    const size_t j0 = (size_t)sp * a.atoms_per_split;
    const size_t j1 = min(a.NAp, j0 + (size_t)a.atoms_per_split);
    const float *__restrict__ X = a.X + t;
    const float *__restrict__ Y = a.Y + t;
    const float *__restrict__ Z = a.Z + t;

    for (size_t jf = j0; jf < j1; jf += kFold) {
        const size_t jend = min(j1, jf + (size_t)kFold);
        float cr[VB], ci[VB];
#pragma unroll
        for (int i = 0; i < VB; ++i) cr[i] = ci[i] = 0.f;
#pragma unroll 1
        for (size_t j = jf; j < jend; j += kAtomChunk) {
            float xs[kAtomChunk], ys[kAtomChunk], zs[kAtomChunk];
#pragma unroll
            for (int c = 0; c < kAtomChunk; ++c) {
                xs[c] = X[(j + c) * a.NFp];
                ys[c] = Y[(j + c) * a.NFp];
                zs[c] = Z[(j + c) * a.NFp];
            }
#pragma unroll
            for (int c = 0; c < kAtomChunk; ++c) {
                const double xd = xs[c], yd = ys[c], zd = zs[c];
                const float b0 = a.b[j + c];
#pragma unroll
                for (int i = 0; i < VB; ++i) {
                    const float bj = PER_ROW_B ? a.b[(size_t)brow[i] * a.NAp + j + c] : b0;
                    double p = qx[i] * xd;
                    p = fma(qy[i], yd, p);
                    p = fma(qz[i], zd, p);
                    const float f = (float)__builtin_amdgcn_fract(p);
                    cr[i] = fmaf(bj, __builtin_amdgcn_cosf(f), cr[i]);
                    ci[i] = fmaf(bj, __builtin_amdgcn_sinf(f), ci[i]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < VB; ++i) {
            tr[i] += (double)cr[i];
            ti[i] += (double)ci[i];
        }
    }

#pragma unroll
    for (int i = 0; i < VB; ++i) {
        const int v = vbase + i;
        if (v < a.V) a.amp[((size_t)sp * a.V + v) * a.NFp + t] = make_float2((float)tr[i], (float)ti[i]);
    }
}


// Work item of an oriented launch: XCD x = blockIdx & 7 takes an equal
// share of the (frame tile, atom range, orientation group) items of every
// shell tile, the shell tiles one after another.  At any time the CUs then
// run tiles of one shell group: with the shell tile varying fastest, the
// dispatch order tied each CU to the items of one shell tile, so tiles of
// unequal cost (exactly stepped and corrected ones, a partial last tile) left
// some CUs idle while others finished theirs.
struct OrientItem {
    int ft, sp, og, sg;
};

__device__ __forceinline__ bool orient_item(const OrientKernelArgs &a, OrientItem &w) {
    const int nsub = a.ftiles * a.split * a.ogroups;
    const int per = (nsub + 7) >> 3;
    const int local = (int)(blockIdx.x >> 3);
    w.sg = local / per;
    const int r = (int)(blockIdx.x & 7) * per + (local - w.sg * per);
    if (w.sg >= a.sgroups || r >= nsub) return false;
    w.og = r % a.ogroups;
    const int rest = r / a.ogroups;
    w.sp = rest % a.split;
    w.ft = rest / a.split;
    return true;
}

static int oriented_grid(int ftiles, int split, int ogroups, int sgroups) {
    return ((ftiles * split * ogroups + 7) / 8) * 8 * sgroups;
}

template <int VM, int VS, bool PER_ROW_B>
__global__ __launch_bounds__(256, 2) void amp_oriented_kernel(OrientKernelArgs a) {
    OrientItem w;
    if (!orient_item(a, w)) return;
    oriented_tile<VM, VS, PER_ROW_B>(a, w.ft, w.sp, w.og, w.sg * VS);
}

// Shells that are an arithmetic progression up to small residuals (modes 2-5
// of oriented_tile_body, by the tile line's flags); one factor per atom.
// The FP64 totals take the dynamic LDS (kResidualLds<VM, VS> bytes).
template <int VM, int VS>
__global__ __launch_bounds__(256, 2) void amp_oriented_res_kernel(OrientKernelArgs a) {
    extern __shared__ double tot_lds[];
    OrientItem w;
    if (!orient_item(a, w)) return;
    residual_tile<VM, VS>(a, w.ft, w.sp, w.og, w.sg * VS, tot_lds);
}

// Oriented tile shapes (orientations x shells per wavefront), picked by the
// host autotuner: 0: 2x8, 1: 4x4, 2: 1x8, 3: 1x16, 4: 1x20.
static const int kShapeVM[kOrientShapes] = {2, 4, 1, 1, 1};
static const int kShapeVS[kOrientShapes] = {8, 4, 8, 16, 20};

int orient_shape_vm(int shape) { return kShapeVM[shape]; }
int orient_shape_vs(int shape) { return kShapeVS[shape]; }

template <int VM, int VS>
static hipError_t launch_oriented_shape(const OrientKernelArgs &k, bool per_row, int blocks, hipStream_t stream) {
    if (k.ap == 2) {
        if (per_row) return hipErrorInvalidValue;
        const void *fn = (const void *)amp_oriented_res_kernel<VM, VS>;
        const size_t lds = kResidualLds<VM, VS>;
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((amp_oriented_res_kernel<VM, VS>), dim3(blocks), dim3(256), lds, stream, k);
    } else if (per_row)
        hipLaunchKernelGGL((amp_oriented_kernel<VM, VS, true>), dim3(blocks), dim3(256), 0, stream, k);
    else
        hipLaunchKernelGGL((amp_oriented_kernel<VM, VS, false>), dim3(blocks), dim3(256), 0, stream, k);
    return hipGetLastError();
}

OrientKernelArgs make_orient_args(const OrientArgs &a) {
    constexpr int CH = 16;
    const int VM = kShapeVM[a.shape], VS = kShapeVS[a.shape];
    OrientKernelArgs k;
    k.X = a.coords;
    k.Y = a.coords + a.NAp * a.NFp;
    k.Z = a.coords + 2 * a.NAp * a.NFp;
    k.NFp = a.NFp;
    k.qlen = a.qlen;
    k.u = a.u;
    k.b = a.b;
    k.NAp = a.NAp;
    k.natoms = a.natoms ? std::min(a.natoms, a.NAp) : a.NAp;
    k.NM = a.NM;
    k.ns = a.ns;
    k.V = a.NM * a.ns;
    k.ftiles = (int)(a.NFp / kWave);
    k.ogroups = (a.NM + 4 * VM - 1) / (4 * VM);
    k.sgroups = (a.ns + VS - 1) / VS;
    k.split = a.split;
    size_t aps = (k.natoms + a.split - 1) / a.split;
    aps = (aps + CH - 1) / CH * CH;
    k.atoms_per_split = (int)aps;
    k.per_shell_b = a.per_shell_b;
    k.amp = a.amp;
    k.ap = a.ap;
    k.qstep = a.qstep;
    k.q0 = a.q0;
    k.qres = a.qres;
    k.lines = a.ap == 2 ? a.lines : nullptr;
    return k;
}

int tile_lines(const double *ql, int ns, double rmax, const ShellGrid &g, const int *firsts, const int *vs,
               int ntiles, TileLine *lines, float *res) {
    double qmax = 0.0;
    for (int s = 0; s < ns; ++s) qmax = std::max(qmax, std::fabs(ql[s]));
    const double tol = 1e-12 * qmax;
    double bound = kMaxShellResidual * qmax;
    if (rmax > 0.0) bound = std::min(bound, kMaxResidualPhase / kTwoPiTurns / rmax);
    int plain = 0;
    for (int t = 0; t < ntiles; ++t) {
        const int first = firsts[t], VS = vs[t], C = VS / 2, n = std::min(VS, ns - first);
        const int mid = first + C;
        // candidate steps: the call's, and that of the exact run through the middle
        double steps[2] = {g.step, 0.0};
        int nsteps = 1;
        if (mid < ns && g.runs[mid].last > mid) steps[nsteps++] = g.runs[mid].step;
        else if (mid - 1 >= first && mid - 1 < ns && g.runs[mid - 1].last >= mid) steps[nsteps++] = g.runs[mid - 1].step;
        else if (mid >= ns && g.runs[first].last > first) steps[nsteps++] = g.runs[first].step;
        int best_flags = 4;
        double best_max = 0.0;
        TileLine best{};
        for (int c = 0; c < nsteps; ++c) {
            const double st = steps[c];
            const double qm = mid < ns ? ql[mid] : ql[first] + (double)C * st;
            int flags = 0;
            double mx = 0.0;
            for (int k = 0; k < n; ++k) {
                const double r = std::fabs(ql[first + k] - (qm + (double)(k - C) * st));
                mx = std::max(mx, r);
                if (r > tol) flags |= k > C ? kTileCorrectUp : k < C ? kTileCorrectDown : 0;
            }
            if (mx > bound) continue;
            const int pop = (flags & 1) + (flags >> 1);
            const int bpop = best_flags > 3 ? 3 : (best_flags & 1) + (best_flags >> 1);
            if (best_flags > 3 || pop < bpop || (pop == bpop && mx < best_max)) {
                best_flags = flags;
                best_max = mx;
                best = TileLine{qm, st, flags, 0};
            }
        }
        if (best_flags > 3) {
            // the call's ideal grid and its residuals (within the bounds)
            best = TileLine{g.q0 + (double)mid * g.step, g.step, kTileCorrectUp | kTileCorrectDown, 0};
        }
        lines[first] = best;
        for (int k = 0; k < n; ++k) {
            const double r = ql[first + k] - (best.qmid + (double)(k - C) * best.step);
            res[first + k] = std::fabs(r) <= tol ? 0.f : (float)(kTwoPiTurns * r);
        }
        if (!(best.flags & kTileCorrectUp)) plain += std::max(0, n - C - 1);
        if (!(best.flags & kTileCorrectDown)) plain += std::min(n, C);
        if (n > C) plain += 1;  // the middle shell lies on the line
    }
    return plain;
}

int shell_progression(const double *qlen, int ns, double rmax, ShellGrid *g) {
    *g = ShellGrid{};
    if (ns < 2) return 0;
    const double d = (qlen[ns - 1] - qlen[0]) / (double)(ns - 1);
    double qmax = 0.0, emax = 0.0;
    for (int s = 0; s < ns; ++s) qmax = std::max(qmax, std::fabs(qlen[s]));
    g->res.resize(ns);
    for (int s = 0; s < ns; ++s) {
        const double e = qlen[s] - (qlen[0] + (double)s * d);
        g->res[s] = e;
        emax = std::max(emax, std::fabs(e));
    }
    g->q0 = qlen[0];
    g->step = d;
    const double tol = 1e-12 * qmax;
    if (emax <= tol) {
        g->mode = 1;
    } else if (emax <= kMaxShellResidual * qmax && emax * rmax <= kMaxResidualPhase / kTwoPiTurns) {
        g->mode = 2;
        g->runs.resize(ns);
        for (int s = 0; s < ns; ++s) {
            // extend the run from s while every shell inside stays on the line
            // through its end points
            int last = s;
            double step = 0.0;
            for (int e = s + 1; e < std::min(ns, s + kMaxRunShells); ++e) {
                const double st = (qlen[e] - qlen[s]) / (double)(e - s);
                bool on = true;
                for (int k = s + 1; k < e && on; ++k)
                    on = std::fabs(qlen[k] - (qlen[s] + (double)(k - s) * st)) <= tol;
                if (!on) break;
                last = e;
                step = st;
            }
            g->runs[s] = ShellRun{step, last, 0};
        }
    } else {
        g->res.clear();
        g->step = 0.0;
    }
    return g->mode;
}

hipError_t launch_amp_oriented(const OrientArgs &a, hipStream_t stream) {
    if (a.shape < 0 || a.shape >= kOrientShapes) return hipErrorInvalidValue;
    const OrientKernelArgs k = make_orient_args(a);
    const int blocks = oriented_grid(k.ftiles, k.split, k.ogroups, k.sgroups);
    const bool per_row = a.per_shell_b != 0;
    switch (a.shape) {
        case 0: return launch_oriented_shape<2, 8>(k, per_row, blocks, stream);
        case 1: return launch_oriented_shape<4, 4>(k, per_row, blocks, stream);
        case 2: return launch_oriented_shape<1, 8>(k, per_row, blocks, stream);
        case 3: return launch_oriented_shape<1, 16>(k, per_row, blocks, stream);
        default: return launch_oriented_shape<1, 20>(k, per_row, blocks, stream);
    }
}

template <int VM, int VS>
static hipError_t oriented_occupancy(bool per_row, int mode, int *blocks) {
    if (mode == 2) {
        const void *k = (const void *)amp_oriented_res_kernel<VM, VS>;
        const size_t lds = kResidualLds<VM, VS>;
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k, 256, lds);
    }
    const void *k = per_row ? (const void *)amp_oriented_kernel<VM, VS, true>
                            : (const void *)amp_oriented_kernel<VM, VS, false>;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k, 256, 0);
}

hipError_t oriented_blocks_per_cu(int shape, bool per_row, int mode, int *blocks) {
    *blocks = 0;
    switch (shape) {
        case 0: return oriented_occupancy<2, 8>(per_row, mode, blocks);
        case 1: return oriented_occupancy<4, 4>(per_row, mode, blocks);
        case 2: return oriented_occupancy<1, 8>(per_row, mode, blocks);
        case 3: return oriented_occupancy<1, 16>(per_row, mode, blocks);
        case 4: return oriented_occupancy<1, 20>(per_row, mode, blocks);
        default: return hipErrorInvalidValue;
    }
}

int oriented_base_blocks(size_t NFp, int NM, int ns, int shape) {
    const int VM = kShapeVM[shape], VS = kShapeVS[shape];
    return (int)(NFp / kWave) * ((NM + 4 * VM - 1) / (4 * VM)) * ((ns + VS - 1) / VS);
}

hipError_t launch_amp_all(const AmpArgs &a, hipStream_t stream) {
    AmpKernelArgs k;
    k.X = a.coords;
    k.Y = a.coords + a.NAp * a.NFp;
    k.Z = a.coords + 2 * a.NAp * a.NFp;
    k.NFp = a.NFp;
    k.qrev = a.qrev;
    k.b = a.b;
    k.brow = a.brow;
    k.NAp = a.NAp;
    k.V = a.V;
    k.ftiles = (int)(a.NFp / kWave);
    k.vgroups = (a.V + 4 * a.vb - 1) / (4 * a.vb);
    k.split = a.split;
    size_t aps = (a.NAp + a.split - 1) / a.split;
    aps = (aps + kAtomChunk - 1) / kAtomChunk * kAtomChunk;
    k.atoms_per_split = (int)aps;
    k.amp = a.amp;
    const int nitems = k.ftiles * k.split * k.vgroups;
    const int blocks = ((nitems + 7) / 8) * 8;
    const bool per_row = a.brow != nullptr;
#define SX_AMP_CASE(VBV)                                                                         \
    case VBV:                                                                                    \
        if (per_row)                                                                             \
            hipLaunchKernelGGL((amp_all_kernel<VBV, true>), dim3(blocks), dim3(256), 0, stream, k);  \
        else                                                                                     \
            hipLaunchKernelGGL((amp_all_kernel<VBV, false>), dim3(blocks), dim3(256), 0, stream, k); \
        break;
    switch (a.vb) {
        SX_AMP_CASE(4)
        SX_AMP_CASE(8)
        SX_AMP_CASE(16)
        default:
            return hipErrorInvalidValue;
    }
#undef SX_AMP_CASE
    return hipGetLastError();
}

hipError_t all_blocks_per_cu(int vb, bool per_row, int *blocks) {
    *blocks = 0;
    const void *k = nullptr;
    switch (vb) {
        case 4: k = per_row ? (const void *)amp_all_kernel<4, true> : (const void *)amp_all_kernel<4, false>; break;
        case 8: k = per_row ? (const void *)amp_all_kernel<8, true> : (const void *)amp_all_kernel<8, false>; break;
        case 16: k = per_row ? (const void *)amp_all_kernel<16, true> : (const void *)amp_all_kernel<16, false>; break;
        default: return hipErrorInvalidValue;
    }
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k, 256, 0);
}

}  // namespace sx
