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
    size_t NAp;
    int NM, ns, V, ftiles, ogroups, sgroups, split, atoms_per_split, per_shell_b;
    float2 *amp;           // [split][V][NFp], v = s * NM + m
    int ap;                // shells in arithmetic progression with step qstep (|q| / 2 pi)
    double qstep;
};

OrientKernelArgs make_orient_args(const OrientArgs &a);

// One tile of the oriented amplitude sum: frame tile ft (one frame per
// lane), atom range sp, orientations 4*VM*og.. (VM per wavefront) and
// shells VS*sg..; 256 threads.
template <int VM, int VS, bool PER_ROW_B>
__device__ __forceinline__ void oriented_tile(const OrientKernelArgs &a, int ft, int sp, int og, int sg) {
    constexpr int CH = 16;  // atoms per FP32 chunk
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
                if (a.ap) {
                    // phases phi_k = phi_0 + k delta: cos and sin of phi_0 and delta from
                    // the hardware (two pairs per orientation, atom and frame), phi_1 by one
                    // complex product, then the three-term recurrence
                    // v_{k+1} = 2 cos(delta) v_k - v_{k-1} on the cosine and the sine
                    // (one FMA each per shell)
                    const float f0 = (float)__builtin_amdgcn_fract(ql[0] * d);
                    const float fs = (float)__builtin_amdgcn_fract(a.qstep * d);
                    const float cw = __builtin_amdgcn_cosf(fs), sw = __builtin_amdgcn_sinf(fs);
                    const float c2 = cw + cw;
                    float xp = __builtin_amdgcn_cosf(f0), yp = __builtin_amdgcn_sinf(f0);
                    float xk = fmaf(xp, cw, -yp * sw), yk = fmaf(yp, cw, xp * sw);
                    cr[i][0] = fmaf(bk[0], xp, cr[i][0]);
                    ci[i][0] = fmaf(bk[0], yp, ci[i][0]);
#pragma unroll
                    for (int k = 1; k < VS; ++k) {
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

}  // namespace sx
