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
    // Shells stepped from phasor to phasor: 0 no (one phase per shell), 1 an
    // exact arithmetic progression (step qstep in |q| / 2 pi), 2 a progression
    // up to small residuals: shell s lies at q0 + s qstep plus qres[s] (radians
    // per Angstrom), corrected to first order in qres[s] d.
    int ap;
    double qstep;
    double q0;             // |q_0| / 2 pi of the ideal grid (ap == 2)
    const float *qres;     // [ns] residuals of the shells against the ideal grid (ap == 2)
    const TileLine *lines; // [ns] lines of the tiles at their first shells (ap == 2)
};

// Workgroup LDS of the residual-corrected tile: the FP64 totals of the 4
// wavefronts, [wave][VM * VS][2][64] doubles.
template <int VM, int VS>
constexpr size_t kResidualLds = sizeof(double) * 4 * VM * VS * 2 * kWave;

OrientKernelArgs make_orient_args(const OrientArgs &a);

// One tile of the oriented amplitude sum: frame tile ft (one frame per
// lane), atom range sp, orientations 4*VM*og.. (VM per wavefront) and
// shells sbase..sbase+VS-1; 256 threads.  MODE 1: the shells of the call are in
// arithmetic progression (step a.qstep), and the tile's shells follow from its
// middle one by the three-term recurrence.  MODE 2: as 1 on the ideal grid
// q0 + s qstep, with each shell's residual r_s = q_s - (q0 + s qstep) (the
// single-precision rounding of a reference q scan, parameters.cpp:1151)
// applied to first order: b e^{i q_s d} = b e^{i q~_s d} (1 + i r_s d), the
// neglected term (r_s d)^2 / 2 being bounded by the host.  The correction
// sums g_s = sum_j b_j d_j e^{i q~_s d_j} run beside the amplitude sums and
// enter them, scaled by r_s, at every fold; the FP64 totals then live in LDS
// (`tot`, kResidualLds bytes) to leave the registers to g_s.  Modes 2-5 step
// on the tile's own line (a.lines[sbase]: the middle shell's |q| and a step),
// the residuals being those against it; the middle shell lies on it, and the
// walks up and down from it carry corrections or not by the mode: 2 both,
// 3 neither, 4 the up walk only, 5 the down walk only.  All four keep their
// totals in LDS, so that they share one kernel's registers.
//
// Atom sums run in FP32 over sub-chunks of CH atoms (coordinates loaded
// together) and are folded into FP64 totals every kFold atoms.

// Wave-uniform, read-only kernel data (orientations, shells, residuals,
// scattering factors) seen through the constant address space, so that it is
// read by scalar loads into SGPRs.  A kernel that takes its pointers as
// arguments gets this by itself; the persistent kernel reads its arguments
// from memory, and through such generic pointers the compiler issues vector
// flat loads that it cannot separate from the LDS totals, waits for every
// load in flight after each one, and re-reads the factors inside the atom
// loop.  The data must not be written while the kernel runs.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *uniform_ptr(const T *p) {
    return (const __attribute__((address_space(4))) T *)p;
}

// The walks of a mode-2 tile that carry corrections (TileLine flags).
__device__ __forceinline__ int tile_flags(const OrientKernelArgs &a, int sbase) {
    return uniform_ptr(a.lines)[sbase].flags;
}

template <int VM, int VS, bool PER_ROW_B, int MODE, bool BUF>
__device__ __forceinline__ void oriented_tile_body(const OrientKernelArgs &a, int ft, int sp, int og, int sbase,
                                                   double *__restrict__ tot) {
    constexpr bool AP = MODE >= 1;
    constexpr bool LT = MODE >= 2;  // FP64 totals in LDS, the tile's own line
    constexpr bool UPC = MODE == 2 || MODE == 4, DNC = MODE == 2 || MODE == 5;
    constexpr bool RES = UPC || DNC;
    static_assert(!(LT && PER_ROW_B), "residual-corrected tiles take one factor per atom");
    constexpr int CH = 16;     // atoms per coordinate batch
    constexpr int kFold = 64;  // atoms per FP32 partial sum
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int mbase = (og * 4 + wave) * VM;
    const size_t t = (size_t)ft * kWave + lane;

    const auto U = uniform_ptr(a.u);
    const auto QL = uniform_ptr(a.qlen);
    const auto QR = uniform_ptr(a.qres);
    const auto B = uniform_ptr(a.b);
    const double qstep = LT ? uniform_ptr(a.lines)[sbase].step : a.qstep;
    double ux[VM], uy[VM], uz[VM], ql[AP ? 1 : VS];
#pragma unroll
    for (int i = 0; i < VM; ++i) {
        const int m = min(mbase + i, a.NM - 1);
        ux[i] = U[3 * m];
        uy[i] = U[3 * m + 1];
        uz[i] = U[3 * m + 2];
    }
    float qr[RES ? VS : 1];
    if constexpr (LT) {
        // the tile's middle shell on its line, and the residuals against it
        ql[0] = uniform_ptr(a.lines)[sbase].qmid;
        if constexpr (RES) {
#pragma unroll
            for (int k = 0; k < VS; ++k) qr[k] = QR[min(sbase + k, a.ns - 1)];
        }
    } else if constexpr (AP) {
        // the tile's middle shell, which need not exist in a partial last tile
        ql[0] = QL[min(sbase, a.ns - 1)] + (double)(VS / 2) * qstep;
    } else {
#pragma unroll
        for (int k = 0; k < VS; ++k) ql[k] = QL[min(sbase + k, a.ns - 1)];
    }

    // FP64 totals: registers, or this wavefront's slice of LDS ([VM*VS][2][64])
    double tr[LT ? 1 : VM][LT ? 1 : VS], ti[LT ? 1 : VM][LT ? 1 : VS];
    double *wt = LT ? tot + (size_t)wave * (VM * VS * 2 * kWave) + lane : nullptr;
#pragma unroll
    for (int i = 0; i < VM; ++i)
#pragma unroll
        for (int k = 0; k < VS; ++k) {
            if constexpr (LT) {
                wt[((i * VS + k) * 2) * kWave] = 0.0;
                wt[((i * VS + k) * 2 + 1) * kWave] = 0.0;
            } else {
                tr[i][k] = ti[i][k] = 0.0;
            }
        }

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
        float gr[RES ? VM : 1][RES ? VS : 1], gi[RES ? VM : 1][RES ? VS : 1];
#pragma unroll
        for (int i = 0; i < VM; ++i)
#pragma unroll
            for (int k = 0; k < VS; ++k) {
                cr[i][k] = ci[i][k] = 0.f;
                if constexpr (RES) gr[i][k] = gi[i][k] = 0.f;
            }
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
                // global (not generic) pointers: loads that only count on
                // vmcnt, so the batch is consumed as it arrives
                typedef const __attribute__((address_space(1))) float *gptr;
                const gptr gx = (gptr)X, gy = (gptr)Y, gz = (gptr)Z;
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    xs[c] = gx[(j + c) * a.NFp + lane];
                    ys[c] = gy[(j + c) * a.NFp + lane];
                    zs[c] = gz[(j + c) * a.NFp + lane];
                }
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const double xd = xs[c], yd = ys[c], zd = zs[c];
                float bk[VS];
#pragma unroll
                for (int k = 0; k < VS; ++k)
                    bk[k] = PER_ROW_B ? B[(size_t)min(sbase + k, a.ns - 1) * a.NAp + j + c] : B[j + c];
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
                        const float fh = 0.5f * (float)__builtin_amdgcn_fract(qstep * d);
                        const float sh = __builtin_amdgcn_sinf(fh), ch = __builtin_amdgcn_cosf(fh);
                        const float rn = 2.f - fmaf(sh, sh, ch * ch);
                        const bool near0 = sh < 0.70710678f;
                        const float ph = near0 ? sh : ch;
                        const float rh = fmaf(-2.f * (ph * ph), rn, 1.f);
                        const float cw = near0 ? rh : -rh;
                        const float sw = ((sh + sh) * ch) * rn;
                        const float c2 = cw + cw;
                        const float xc = __builtin_amdgcn_cosf(fc), yc = __builtin_amdgcn_sinf(fc);
                        // b_j d_j of the correction sums (one factor per atom)
                        const float bd = RES ? bk[0] * (float)d : 0.f;
                        cr[i][C] = fmaf(bk[C], xc, cr[i][C]);
                        ci[i][C] = fmaf(bk[C], yc, ci[i][C]);
                        {  // shells above the middle
                            float xp = xc, yp = yc;
                            float xk = fmaf(xc, cw, -yc * sw), yk = fmaf(yc, cw, xc * sw);
#pragma unroll
                            for (int k = C + 1; k < VS; ++k) {
                                cr[i][k] = fmaf(bk[k], xk, cr[i][k]);
                                ci[i][k] = fmaf(bk[k], yk, ci[i][k]);
                                if constexpr (UPC) {
                                    gr[i][k] = fmaf(bd, yk, gr[i][k]);
                                    gi[i][k] = fmaf(bd, xk, gi[i][k]);
                                }
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
                                if constexpr (DNC) {
                                    gr[i][k] = fmaf(bd, yk, gr[i][k]);
                                    gi[i][k] = fmaf(bd, xk, gi[i][k]);
                                }
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
                constexpr int C = VS / 2;
                if constexpr (RES) {
                    if ((k > C && UPC) || (k < C && DNC)) {
                        // sum b e^{i q~ d} (1 + i r d) = (cr - r gr) + i (ci + r gi)
                        wt[((i * VS + k) * 2) * kWave] += (double)fmaf(-qr[k], gr[i][k], cr[i][k]);
                        wt[((i * VS + k) * 2 + 1) * kWave] += (double)fmaf(qr[k], gi[i][k], ci[i][k]);
                    } else {
                        wt[((i * VS + k) * 2) * kWave] += (double)cr[i][k];
                        wt[((i * VS + k) * 2 + 1) * kWave] += (double)ci[i][k];
                    }
                } else if constexpr (LT) {
                    wt[((i * VS + k) * 2) * kWave] += (double)cr[i][k];
                    wt[((i * VS + k) * 2 + 1) * kWave] += (double)ci[i][k];
                } else {
                    tr[i][k] += (double)cr[i][k];
                    ti[i][k] += (double)ci[i][k];
                }
            }
    }

#pragma unroll
    for (int i = 0; i < VM; ++i)
#pragma unroll
        for (int k = 0; k < VS; ++k) {
            const int m = mbase + i, sh = sbase + k;
            if (m < a.NM && sh < a.ns) {
                const size_t v = (size_t)sh * a.NM + m;
                double xr, xi;
                if constexpr (LT) {
                    xr = wt[((i * VS + k) * 2) * kWave];
                    xi = wt[((i * VS + k) * 2 + 1) * kWave];
                } else {
                    xr = tr[i][k];
                    xi = ti[i][k];
                }
                a.amp[((size_t)sp * a.V + v) * a.NFp + t] = make_float2((float)xr, (float)xi);
            }
        }
}

// A tile of a residual-corrected call, in the mode its line asks for.
template <int VM, int VS>
__device__ __forceinline__ void residual_tile(const OrientKernelArgs &a, int ft, int sp, int og, int sbase,
                                              double *__restrict__ tot) {
    switch (tile_flags(a, sbase)) {
        case 0: oriented_tile_body<VM, VS, false, 3, true>(a, ft, sp, og, sbase, tot); break;
        case kTileCorrectUp: oriented_tile_body<VM, VS, false, 4, true>(a, ft, sp, og, sbase, tot); break;
        case kTileCorrectDown: oriented_tile_body<VM, VS, false, 5, true>(a, ft, sp, og, sbase, tot); break;
        default: oriented_tile_body<VM, VS, false, 2, true>(a, ft, sp, og, sbase, tot); break;
    }
}

// Modes 0 and 1 (the mode is uniform over a launch or a trajectory); modes
// 2-5 have kernels of their own, since they need `tot` in LDS.
template <int VM, int VS, bool PER_ROW_B, bool BUF = true>
__device__ __forceinline__ void oriented_tile(const OrientKernelArgs &a, int ft, int sp, int og, int sbase) {
    if (a.ap)
        oriented_tile_body<VM, VS, PER_ROW_B, 1, BUF>(a, ft, sp, og, sbase, nullptr);
    else
        oriented_tile_body<VM, VS, PER_ROW_B, 0, BUF>(a, ft, sp, og, sbase, nullptr);
}

}  // namespace sx
