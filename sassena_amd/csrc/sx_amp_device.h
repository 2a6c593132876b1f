// sx_amp_device.h -- device body of the oriented coherent amplitude tile,
// shared by the per-context kernel and the persistent batched kernel.
#pragma once

#include "sx_internal.h"

namespace sx {

struct OrientKernelArgs {
    const float *X, *Y, *Z;
    size_t NFp;
    const double *qlen;    // [ns] |q_s| / 2 pi
    const double *u;       // [NM][3] unit orientations
    const float *b;        // [nb][NAp]
    size_t NAp;            // row stride of the coordinate planes and of b
    size_t natoms;         // atoms summed (<= NAp)
    int NM, ns, V, ftiles, ogroups, sgroups, split, atoms_per_split, per_shell_b;
    float2 *amp;           // [split][V][NFp], v = s * NM + m
    int ap;                // shells in arithmetic progression with step qstep (|q| / 2 pi)
    double qstep;
};

OrientKernelArgs make_orient_args(const OrientArgs &a);

// One tile of the oriented amplitude sum: frame tile ft (one frame per
// lane), atom range sp, orientations 4*VM*og.. (VM per wavefront) and
// shells VS*sg..; 256 threads.  AP: the shells of the call are in arithmetic
// progression (step a.qstep), and the tile's shells follow from its middle one
// by the three-term recurrence.
//
// Atom sums run in FP32 over sub-chunks of CH atoms (coordinates loaded
// together) and are folded into FP64 totals every kFold atoms.
// Word 3 of a raw (unswizzled, stride 0) buffer resource on CDNA: 32-bit
// float data format, no index stride.
constexpr int kBufferDword3 = 0x00020000;

template <int VM, int VS, bool PER_ROW_B, bool AP, bool BUF>
__device__ __forceinline__ void oriented_tile_body(const OrientKernelArgs &a, int ft, int sp, int og, int sg) {
    constexpr int CH = 16;     // atoms per coordinate batch
    constexpr int kFold = 64;  // atoms per FP32 partial sum
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int mbase = (og * 4 + wave) * VM;
    const int sbase = sg * VS;
    const size_t t = (size_t)ft * kWave + lane;

    double ux[VM], uy[VM], uz[VM], ql[AP ? 1 : VS];
#pragma unroll
    for (int i = 0; i < VM; ++i) {
        const int m = min(mbase + i, a.NM - 1);
        ux[i] = a.u[3 * m];
        uy[i] = a.u[3 * m + 1];
        uz[i] = a.u[3 * m + 2];
    }
    if constexpr (AP) {
        // the tile's middle shell, which need not exist in a partial last tile
        ql[0] = a.qlen[min(sbase, a.ns - 1)] + (double)(VS / 2) * a.qstep;
    } else {
#pragma unroll
        for (int k = 0; k < VS; ++k) ql[k] = a.qlen[min(sbase + k, a.ns - 1)];
    }

    double tr[VM][VS], ti[VM][VS];
#pragma unroll
    for (int i = 0; i < VM; ++i)
#pragma unroll
        for (int k = 0; k < VS; ++k) tr[i][k] = ti[i][k] = 0.0;

    const size_t j0 = (size_t)sp * a.atoms_per_split;
    const size_t j1 = min(a.natoms, j0 + (size_t)a.atoms_per_split);
    // BUF: coordinates through buffer loads, a wave-uniform resource per
    // coordinate batch (base = first atom of the batch, this frame tile), the
    // atom's row as the scalar offset and the lane's frame as the vector
    // offset, so the loads need no per-lane address arithmetic.  Else plain
    // global loads (the persistent batched kernel: measured faster there).
    const float *X = a.X + (size_t)ft * kWave;
    const float *Y = a.Y + (size_t)ft * kWave;
    const float *Z = a.Z + (size_t)ft * kWave;
    const unsigned voff = (unsigned)lane * 4u;
    const int rows = (int)(CH * a.NFp * sizeof(float));

    for (size_t jf = j0; jf < j1; jf += kFold) {
        const size_t jend = min(j1, jf + (size_t)kFold);
        float cr[VM][VS], ci[VM][VS];
#pragma unroll
        for (int i = 0; i < VM; ++i)
#pragma unroll
            for (int k = 0; k < VS; ++k) cr[i][k] = ci[i][k] = 0.f;
#pragma unroll 1
        for (size_t j = jf; j < jend; j += CH) {
            float xs[CH], ys[CH], zs[CH];
            if constexpr (BUF) {
                const __amdgpu_buffer_rsrc_t rx =
                    __builtin_amdgcn_make_buffer_rsrc((void *)(X + j * a.NFp), (short)0, rows, kBufferDword3);
                const __amdgpu_buffer_rsrc_t ry =
                    __builtin_amdgcn_make_buffer_rsrc((void *)(Y + j * a.NFp), (short)0, rows, kBufferDword3);
                const __amdgpu_buffer_rsrc_t rz =
                    __builtin_amdgcn_make_buffer_rsrc((void *)(Z + j * a.NFp), (short)0, rows, kBufferDword3);
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    const int so = (int)(c * a.NFp * sizeof(float));
                    xs[c] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rx, voff, so, 0));
                    ys[c] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ry, voff, so, 0));
                    zs[c] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rz, voff, so, 0));
                }
            } else {
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    xs[c] = X[(j + c) * a.NFp + lane];
                    ys[c] = Y[(j + c) * a.NFp + lane];
                    zs[c] = Z[(j + c) * a.NFp + lane];
                }
            }
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
                    if constexpr (AP) {
                        // phases phi_k = phi_c + (k - c) delta about the tile's middle shell
                        // c: cos and sin of phi_c and of delta / 2 from the hardware (two pairs
                        // per orientation, atom and frame), phi_{c+1} and phi_{c-1} by one
                        // complex product each, then the three-term recurrence
                        // v_{k+-1} = 2 cos(delta) v_k - v_{k-+1} up and down the tile (one FMA
                        // each per shell on the cosine and the sine); starting in the middle
                        // halves the number of steps over which rounding errors grow.  The
                        // recurrence multiplies an error in cos(delta) by up to (k - c)^2
                        // where cos(delta) is near +-1, the hardware cosine is biased there by
                        // about an ulp, and with phi_c a multiple of delta (shells starting at
                        // a multiple of the step) such errors add up coherently over atoms.
                        // The step phasor is therefore built from the half angle (s, c):
                        // cos(delta) = 1 - 2 s^2 / n or 2 c^2 / n - 1, whichever squares the
                        // smaller factor, sin(delta) = 2 s c / n, with n = s^2 + c^2 removing
                        // the magnitude error of the hardware pair (1 / n = 2 - n to first
                        // order, n being within 1e-6 of 1).
                        constexpr int C = VS / 2;
                        const float fc = (float)__builtin_amdgcn_fract(ql[0] * d);
                        const float fh = 0.5f * (float)__builtin_amdgcn_fract(a.qstep * d);
                        const float sh = __builtin_amdgcn_sinf(fh), ch = __builtin_amdgcn_cosf(fh);
                        const float rn = 2.f - fmaf(sh, sh, ch * ch);
                        const bool near0 = sh < 0.70710678f;
                        const float ph = near0 ? sh : ch;
                        const float rh = fmaf(-2.f * (ph * ph), rn, 1.f);
                        const float cw = near0 ? rh : -rh;
                        const float sw = ((sh + sh) * ch) * rn;
                        const float c2 = cw + cw;
                        const float xc = __builtin_amdgcn_cosf(fc), yc = __builtin_amdgcn_sinf(fc);
                        cr[i][C] = fmaf(bk[C], xc, cr[i][C]);
                        ci[i][C] = fmaf(bk[C], yc, ci[i][C]);
                        {  // shells above the middle
                            float xp = xc, yp = yc;
                            float xk = fmaf(xc, cw, -yc * sw), yk = fmaf(yc, cw, xc * sw);
#pragma unroll
                            for (int k = C + 1; k < VS; ++k) {
                                cr[i][k] = fmaf(bk[k], xk, cr[i][k]);
                                ci[i][k] = fmaf(bk[k], yk, ci[i][k]);
                                if (k + 1 < VS) {
                                    const float xn = fmaf(c2, xk, -xp), yn = fmaf(c2, yk, -yp);
                                    xp = xk;
                                    yp = yk;
                                    xk = xn;
                                    yk = yn;
                                }
                            }
                        }
                        {  // shells below the middle
                            float xp = xc, yp = yc;
                            float xk = fmaf(xc, cw, yc * sw), yk = fmaf(yc, cw, -xc * sw);
#pragma unroll
                            for (int k = C - 1; k >= 0; --k) {
                                cr[i][k] = fmaf(bk[k], xk, cr[i][k]);
                                ci[i][k] = fmaf(bk[k], yk, ci[i][k]);
                                if (k > 0) {
                                    const float xn = fmaf(c2, xk, -xp), yn = fmaf(c2, yk, -yp);
                                    xp = xk;
                                    yp = yk;
                                    xk = xn;
                                    yk = yn;
                                }
                            }
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < VS; ++k) {
                            const float f = (float)__builtin_amdgcn_fract(ql[k] * d);
                            cr[i][k] = fmaf(bk[k], __builtin_amdgcn_cosf(f), cr[i][k]);
                            ci[i][k] = fmaf(bk[k], __builtin_amdgcn_sinf(f), ci[i][k]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < VM; ++i)
#pragma unroll
            for (int k = 0; k < VS; ++k) {
                tr[i][k] += (double)cr[i][k];
                ti[i][k] += (double)ci[i][k];
            }
    }

#pragma unroll
    for (int i = 0; i < VM; ++i)
#pragma unroll
        for (int k = 0; k < VS; ++k) {
            const int m = mbase + i, sh = sbase + k;
            if (m < a.NM && sh < a.ns) {
                const size_t v = (size_t)sh * a.NM + m;
                a.amp[((size_t)sp * a.V + v) * a.NFp + t] = make_float2((float)tr[i][k], (float)ti[i][k]);
            }
        }
}

template <int VM, int VS, bool PER_ROW_B, bool BUF = true>
__device__ __forceinline__ void oriented_tile(const OrientKernelArgs &a, int ft, int sp, int og, int sg) {
    if (a.ap)
        oriented_tile_body<VM, VS, PER_ROW_B, true, BUF>(a, ft, sp, og, sg);
    else
        oriented_tile_body<VM, VS, PER_ROW_B, false, BUF>(a, ft, sp, og, sg);
}

}  // namespace sx
