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
// over chunks of atoms and are folded into a Kahan-compensated FP32 total.
//
// When the vectors of all shells are one orientation set scaled by |q_s|
// (orientation types file and sphere, abstract_vectors_scatter_device.cpp:115-129),
// amp_oriented_kernel forms d = u_m . r once per atom and orientation and the
// phase of each shell as the single FP64 product (|q_s| / 2 pi) d.
#include "sx_internal.h"

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

    float sr[VB], si[VB], kr[VB], ki[VB];
#pragma unroll
    for (int i = 0; i < VB; ++i) sr[i] = si[i] = kr[i] = ki[i] = 0.f;

    const size_t j0 = (size_t)sp * a.atoms_per_split;
    const size_t j1 = min(a.NAp, j0 + (size_t)a.atoms_per_split);
    const float *__restrict__ X = a.X + t;
    const float *__restrict__ Y = a.Y + t;
    const float *__restrict__ Z = a.Z + t;

    for (size_t j = j0; j < j1; j += kAtomChunk) {
        float xs[kAtomChunk], ys[kAtomChunk], zs[kAtomChunk];
#pragma unroll
        for (int c = 0; c < kAtomChunk; ++c) {
            xs[c] = X[(j + c) * a.NFp];
            ys[c] = Y[(j + c) * a.NFp];
            zs[c] = Z[(j + c) * a.NFp];
        }
        float cr[VB], ci[VB];
#pragma unroll
        for (int i = 0; i < VB; ++i) cr[i] = ci[i] = 0.f;
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
#pragma unroll
        for (int i = 0; i < VB; ++i) {
            float y = cr[i] - kr[i];
            float s = sr[i] + y;
            kr[i] = (s - sr[i]) - y;
            sr[i] = s;
            y = ci[i] - ki[i];
            s = si[i] + y;
            ki[i] = (s - si[i]) - y;
            si[i] = s;
        }
    }

#pragma unroll
    for (int i = 0; i < VB; ++i) {
        const int v = vbase + i;
        if (v < a.V) a.amp[((size_t)sp * a.V + v) * a.NFp + t] = make_float2(sr[i] - kr[i], si[i] - ki[i]);
    }
}


struct OrientKernelArgs {
    const float *X, *Y, *Z;
    size_t NFp;
    const double *qlen;    // [ns] |q_s| / 2 pi
    const double *u;       // [NM][3] unit orientations
    const float *b;        // [nb][NAp]
    size_t NAp;
    int NM, ns, V, ftiles, ogroups, sgroups, split, atoms_per_split, per_shell_b;
    float2 *amp;           // [split][V][NFp], v = s * NM + m
};

template <int VM, int VS, bool PER_ROW_B>
__global__ __launch_bounds__(256) void amp_oriented_kernel(OrientKernelArgs a) {
    constexpr int CH = 16;  // atoms per FP32 chunk
    const int nitems = a.ftiles * a.split * a.ogroups * a.sgroups;
    const int per = (nitems + 7) >> 3;
    const int item = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (item >= nitems) return;
    const int sg = item % a.sgroups;
    int rest = item / a.sgroups;
    const int og = rest % a.ogroups;
    rest /= a.ogroups;
    const int sp = rest % a.split;
    const int ft = rest / a.split;

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int mbase = (og * 4 + wave) * VM;
    const int sbase = sg * VS;
    const size_t t = (size_t)ft * kWave + lane;

    double ux[VM], uy[VM], uz[VM], ql[VS];
#pragma unroll
    for (int i = 0; i < VM; ++i) {
        const int m = min(mbase + i, a.NM - 1);
        ux[i] = a.u[3 * m];
        uy[i] = a.u[3 * m + 1];
        uz[i] = a.u[3 * m + 2];
    }
#pragma unroll
    for (int k = 0; k < VS; ++k) ql[k] = a.qlen[min(sbase + k, a.ns - 1)];

    float sr[VM][VS], si[VM][VS], kr[VM][VS], ki[VM][VS];
#pragma unroll
    for (int i = 0; i < VM; ++i)
#pragma unroll
        for (int k = 0; k < VS; ++k) sr[i][k] = si[i][k] = kr[i][k] = ki[i][k] = 0.f;

    const size_t j0 = (size_t)sp * a.atoms_per_split;
    const size_t j1 = min(a.NAp, j0 + (size_t)a.atoms_per_split);
    const float *__restrict__ X = a.X + t;
    const float *__restrict__ Y = a.Y + t;
    const float *__restrict__ Z = a.Z + t;

    for (size_t j = j0; j < j1; j += CH) {
        float xs[CH], ys[CH], zs[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            xs[c] = X[(j + c) * a.NFp];
            ys[c] = Y[(j + c) * a.NFp];
            zs[c] = Z[(j + c) * a.NFp];
        }
        float cr[VM][VS], ci[VM][VS];
#pragma unroll
        for (int i = 0; i < VM; ++i)
#pragma unroll
            for (int k = 0; k < VS; ++k) cr[i][k] = ci[i][k] = 0.f;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const double xd = xs[c], yd = ys[c], zd = zs[c];
            float bk[VS];
#pragma unroll
            for (int k = 0; k < VS; ++k)
                bk[k] = PER_ROW_B ? a.b[(size_t)min(sbase + k, a.ns - 1) * a.NAp + j + c] : a.b[j + c];
#pragma unroll
            for (int i = 0; i < VM; ++i) {
                double d = ux[i] * xd;
                d = fma(uy[i], yd, d);
                d = fma(uz[i], zd, d);
#pragma unroll
                for (int k = 0; k < VS; ++k) {
                    const float f = (float)__builtin_amdgcn_fract(ql[k] * d);
                    cr[i][k] = fmaf(bk[k], __builtin_amdgcn_cosf(f), cr[i][k]);
                    ci[i][k] = fmaf(bk[k], __builtin_amdgcn_sinf(f), ci[i][k]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < VM; ++i)
#pragma unroll
            for (int k = 0; k < VS; ++k) {
                float y = cr[i][k] - kr[i][k];
                float s = sr[i][k] + y;
                kr[i][k] = (s - sr[i][k]) - y;
                sr[i][k] = s;
                y = ci[i][k] - ki[i][k];
                s = si[i][k] + y;
                ki[i][k] = (s - si[i][k]) - y;
                si[i][k] = s;
            }
    }

#pragma unroll
    for (int i = 0; i < VM; ++i)
#pragma unroll
        for (int k = 0; k < VS; ++k) {
            const int m = mbase + i, sh = sbase + k;
            if (m < a.NM && sh < a.ns) {
                const size_t v = (size_t)sh * a.NM + m;
                a.amp[((size_t)sp * a.V + v) * a.NFp + t] =
                    make_float2(sr[i][k] - kr[i][k], si[i][k] - ki[i][k]);
            }
        }
}

hipError_t launch_amp_oriented(const OrientArgs &a, hipStream_t stream) {
    constexpr int VM = 2, VS = 8, CH = 16;
    OrientKernelArgs k;
    k.X = a.coords;
    k.Y = a.coords + a.NAp * a.NFp;
    k.Z = a.coords + 2 * a.NAp * a.NFp;
    k.NFp = a.NFp;
    k.qlen = a.qlen;
    k.u = a.u;
    k.b = a.b;
    k.NAp = a.NAp;
    k.NM = a.NM;
    k.ns = a.ns;
    k.V = a.NM * a.ns;
    k.ftiles = (int)(a.NFp / kWave);
    k.ogroups = (a.NM + 4 * VM - 1) / (4 * VM);
    k.sgroups = (a.ns + VS - 1) / VS;
    k.split = a.split;
    size_t aps = (a.NAp + a.split - 1) / a.split;
    aps = (aps + CH - 1) / CH * CH;
    k.atoms_per_split = (int)aps;
    k.per_shell_b = a.per_shell_b;
    k.amp = a.amp;
    const int nitems = k.ftiles * k.split * k.ogroups * k.sgroups;
    const int blocks = ((nitems + 7) / 8) * 8;
    if (a.per_shell_b)
        hipLaunchKernelGGL((amp_oriented_kernel<VM, VS, true>), dim3(blocks), dim3(256), 0, stream, k);
    else
        hipLaunchKernelGGL((amp_oriented_kernel<VM, VS, false>), dim3(blocks), dim3(256), 0, stream, k);
    return hipGetLastError();
}

int oriented_base_blocks(size_t NFp, int NM, int ns) {
    return (int)(NFp / kWave) * ((NM + 4 * 2 - 1) / (4 * 2)) * ((ns + 8 - 1) / 8);
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

}  // namespace sx
