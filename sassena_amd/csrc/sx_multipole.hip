// sx_multipole.hip -- amplitudes of the multipole orientational averages.
//
// Sphere (MPSphereScatterDevice::scatter, multipole_scatter_device.cpp:428-498):
//   A_lm(t) = sum_j 4 pi i^l b_j j_l(|q| r_j) conj(Y_lm(theta_j, phi_j))
// Cylinder (MPCylinderScatterDevice::scatter, :900-987): with q and r_j in the
// cylinder frame of the axis (r, phi, z),
//   A_lm(t) = sqrt(2 pi) sum_j b_j e^{i s z q_z} F_lm(r q_r, phi - phi_q)
// where F_lm is J_0, or sqrt(1/2) 2 (-1)^l J_2l cos/sin(2l psi) (m = 0, 1), or
// i sqrt(1/2) 2 (-1)^(l-1) J_(2l-1) cos/sin((2l-1) psi) (m = 2, 3).
// The staged Cartesian float coordinates are converted per atom and frame in
// FP64 and the spherical / cylinder components rounded to float, as the
// reference stages them (coordinate_sets.cpp:359-361, data_stager.cpp:111-113).
// Lane = frame; a wavefront accumulates up to kMoments moments of one
// (shell, moment block) in FP64; the amplitudes feed the same correlation
// and finalisation kernels as the vector averages.
#include <algorithm>
#include <cmath>

#include "sx_internal.h"

namespace sx {

constexpr int kMoments = 8;
constexpr double kPi = 3.14159265358979323846;

struct MultipoleKernelArgs {
    const float *X, *Y, *Z;
    size_t NFp, NAp;
    int kind;               // 0 sphere, 1 cylinder
    const double *qinfo;    // per shell: sphere |q|; cylinder q_r, phi_q, q_z (3 doubles)
    const double *base;     // cylinder frame er, ephi, ez (9 doubles)
    const int *moments;     // [nmo][2] (l, m)
    const double *ynorm;    // sphere: [nmo][2] start of the Legendre recurrence c_|m| and final 4 pi f_l|m|
    const double *plm_ab;   // sphere: [l1][l1][2] coefficients of the Legendre recurrence (plm_scaled)
    int plm_l1;
    int nmo, ns, mblocks;
    const float *b;         // [nb][NAp]
    int per_shell_b;
    float2 *amp;            // [split][ns * nmo][NFp]
    int split, zgroups;     // atom ranges; groups of 4 moment blocks per range
    size_t range;           // atoms per range (the last may be shorter)
    const float *srf;       // sphere terms per (atom, frame), or nullptr: r (float-rounded),
    const double *sct, *sst, *ssp, *scp;  // cos theta, sin theta, sin phi, cos phi
    const float *cphi, *cz;  // cylinder terms (r in srf), float-rounded as the reference stages them
    const double2 *cj01;     // j_0, j_1 (sphere) or J_0, J_1 (cylinder) per (shell, atom, frame), or nullptr
    const double *jtab;      // j_l (sphere) or J_n (cylinder) per (shell, degree or order, atom, frame), or nullptr
    int jl1;
    const double2 *etab;     // cylinder with jtab: (cos, sin) of sgn z q_z and of psi per (shell, 2, atom, frame)
};

// j_l(x) for the moment degrees ls[0..kMoments) of a wavefront in one
// recurrence: upward from j_0, j_1 to lhi where that is stable (x > lhi),
// else Miller's downward recurrence from above lhi normalised to j_0 (or j_1
// near a zero of j_0); each degree is captured as the recurrence passes it.
// SORTED: ls is non-decreasing (the reference's moment lists), so the
// recurrence runs in segments between consecutive degrees with no per-step
// comparisons; otherwise every step compares with every degree.
template <bool SORTED>
__device__ __forceinline__ void sph_bessel_moments(const int (&ls)[kMoments], int lhi, double x, double ix,
                                                   double j0, double j1, double (&jv)[kMoments]) {
    if (x == 0.0) {
#pragma unroll
        for (int k = 0; k < kMoments; ++k) jv[k] = ls[k] == 0 ? 1.0 : 0.0;
        return;
    }
#pragma unroll
    for (int k = 0; k < kMoments; ++k) jv[k] = ls[k] == 0 ? j0 : j1;  // degrees >= 2 are set below
    if (lhi < 2) return;
    if (x > (double)lhi) {
        double a = j0, bb = j1;  // j_{n-1}, j_n
        int n = 1;
        if constexpr (SORTED) {
#pragma unroll
            for (int k = 0; k < kMoments; ++k) {
                for (; n < ls[k]; ++n) {
                    const double nx = (2.0 * n + 1.0) * ix * bb - a;
                    a = bb;
                    bb = nx;
                }
                if (ls[k] >= 2) jv[k] = bb;
            }
        } else {
            for (; n < lhi; ++n) {
                const double nx = (2.0 * n + 1.0) * ix * bb - a;
                a = bb;
                bb = nx;  // j_{n+1}
#pragma unroll
                for (int k = 0; k < kMoments; ++k)
                    if (ls[k] == n + 1) jv[k] = bb;
            }
        }
        return;
    }
    // x <= lhi here, so lhi + 32 + 2 sqrt(2 lhi) bounds lhi + 32 + 2 sqrt(lhi + x)
    // for every lane: one start for the wavefront, whose step counter is then
    // uniform (scalar loop control, no per-lane trip counts)
    const int N = __builtin_amdgcn_readfirstlane(lhi + 32 + (int)(2.0 * sqrt(2.0 * (double)lhi)));
    double fp1 = 0.0, f = 1e-280;  // f_{n+1}, f_n
    double cap[kMoments];
#pragma unroll
    for (int k = 0; k < kMoments; ++k) cap[k] = 0.0;
    int n = N;
    // one downward step: f_n -> f_{n-1}, rescaling everything captured so far
    auto step = [&]() {
        const double fm1 = (2.0 * n + 1.0) * ix * f - fp1;
        fp1 = f;
        f = fm1;
        --n;
        if (fabs(f) > 1e250) {
            f *= 1e-250;
            fp1 *= 1e-250;
#pragma unroll
            for (int k = 0; k < kMoments; ++k) cap[k] *= 1e-250;
        }
    };
    if constexpr (SORTED) {
#pragma unroll
        for (int k = kMoments - 1; k >= 0; --k) {
            while (n > ls[k]) step();
            cap[k] = f;  // f_{ls[k]}
        }
    } else {
        while (n >= 2) {
#pragma unroll
            for (int k = 0; k < kMoments; ++k)
                if (ls[k] == n) cap[k] = f;
            step();
        }
    }
    while (n > 0) step();
    // now f = f_0 and fp1 = f_1, at the same scale as every capture
    const double sc = fabs(j0) >= fabs(j1) ? j0 / f : j1 / fp1;
#pragma unroll
    for (int k = 0; k < kMoments; ++k)
        if (ls[k] >= 2) jv[k] = cap[k] * sc;
}

// J_n(x), x >= 0, for the orders ns[0..kMoments) of a wavefront in one
// recurrence: orders 0 and 1 from J0, J1; the others upward from J0, J1 where
// x exceeds the largest order nhi, else by Miller's downward recurrence from
// above nhi, normalised by J_0 + 2 sum_k J_2k = 1 and captured on the way.
__device__ __forceinline__ void cyl_bessel_orders(const int (&ns)[kMoments], int nhi, double x, double J0,
                                                  double J1, double (&jv)[kMoments]) {
#pragma unroll
    for (int q = 0; q < kMoments; ++q) jv[q] = ns[q] == 0 ? J0 : (ns[q] == 1 ? J1 : 0.0);
    // orders >= 2 vanish at x = 0 (below 1e-20 they are below 1e-40 of J_0)
    if (nhi < 2 || !(x > 1e-20)) return;
    const double ix = 1.0 / x;
    // the lowest captured order >= 2: steps outside [nlo, nhi] compare with none
    int nlo = nhi;
#pragma unroll
    for (int q = 0; q < kMoments; ++q)
        if (ns[q] >= 2) nlo = min(nlo, ns[q]);
    if (x > (double)nhi) {
        double a = J0, bb = J1;  // J_{k-1}, J_k
        int k = 1;
        for (; k + 1 < nlo; ++k) {
            const double nx = (2.0 * k) * ix * bb - a;
            a = bb;
            bb = nx;  // J_{k+1}
        }
        for (; k < nhi; ++k) {
            const double nx = (2.0 * k) * ix * bb - a;
            a = bb;
            bb = nx;  // J_{k+1}
#pragma unroll
            for (int q = 0; q < kMoments; ++q)
                if (ns[q] == k + 1) jv[q] = bb;
        }
        return;
    }
    // Miller's downward recurrence.  x <= nhi here, so the start
    // nhi + 32 + 2 sqrt(2 nhi) bounds nhi + 32 + 2 sqrt(nhi + x) for every
    // lane: one even start for the wavefront, whose loop counter is then
    // uniform, and the recurrence runs in three segments, above the captured
    // orders, through them (the only steps that compare) and below them.
    int N = nhi + 32 + (int)(2.0 * sqrt(2.0 * (double)nhi));
    N = __builtin_amdgcn_readfirstlane(N + (N & 1));  // even: the normalisation sums even orders
    double fp1 = 0.0, f = 1e-280, norm = 0.0;
    double cap[kMoments];
#pragma unroll
    for (int q = 0; q < kMoments; ++q) cap[q] = 0.0;
    // one step: f <- J_{k-1} (unnormalised), 2 J_{k-1} into the sum when k - 1 is even and positive
    auto step = [&](int k) {
        const double fm1 = (2.0 * k) * ix * f - fp1;
        fp1 = f;
        f = fm1;
        if (((k - 1) & 1) == 0 && k > 1) norm += 2.0 * f;
        if (fabs(f) > 1e250) {
            f *= 1e-250;
            fp1 *= 1e-250;
            norm *= 1e-250;
#pragma unroll
            for (int q = 0; q < kMoments; ++q) cap[q] *= 1e-250;
        }
    };
    for (int k = N; k > nhi + 1; --k) step(k);
    for (int k = nhi + 1; k > nlo; --k) {
        step(k);
#pragma unroll
        for (int q = 0; q < kMoments; ++q)
            if (ns[q] == k - 1) cap[q] = f;
    }
    for (int k = nlo; k > 0; --k) step(k);
    norm += f;  // J_0
    const double inorm = 1.0 / norm;
#pragma unroll
    for (int q = 0; q < kMoments; ++q)
        if (ns[q] >= 2) jv[q] = cap[q] * inorm;
}

// N_mm P_l^m(x) with the Condon-Shortley phase, m >= 0, somx2 = sqrt(1 - x^2):
// the upward recurrence in l started from N_mm P_m^m = cm (-somx2)^m, with
// cm = sqrt((2m + 1) / (4 pi) prod_{i<=m} (2i - 1) / (2i)) (host table), so
// that no factorial is formed; the caller multiplies the finished sums by
// f_lm = N_lm / N_mm = sqrt((2l+1)/(2m+1) (l-m)! (2m)! / (l+m)!) (host table).
// Every intermediate stays finite for l <= 200.
// ab[ll * l1 + m] = ((2 ll - 1) / (ll - m), (ll + m - 1) / (ll - m)): the
// coefficients of the recurrence in the degree, built on the host and read
// through the constant address space (scalar loads: l and m are uniform
// over the wavefront), so that a step is two products and an FMA.
__device__ double plm_scaled(int l, int m, double x, double somx2, double cm, const double *ab, int l1) {
    const auto AB = (const __attribute__((address_space(4))) double *)ab;
    double pmm = cm;
    for (int i = 1; i <= m; ++i) pmm *= -somx2;
    if (l == m) return pmm;
    double pmmp1 = x * (2.0 * m + 1.0) * pmm;
    double plm = pmmp1;
    for (int ll = m + 2; ll <= l; ++ll) {
        const int i = 2 * (ll * l1 + m);
        const double pll = fma(AB[i] * x, pmmp1, -AB[i + 1] * pmm);
        pmm = pmmp1;
        pmmp1 = pll;
        plm = pll;
    }
    return plm;
}

// coor3d.cpp:168-202
__device__ void to_spherical(double x, double y, double z, double &r, double &phi, double &theta) {
    r = sqrt(x * x + y * y + z * z);
    if (r == 0) {
        theta = 0;
        phi = 0;
        return;
    }
    theta = acos(z / r);
    double p;
    if (x != 0.0) {
        p = atan(y / x);
        if (x < 0.0) p += kPi;
        else if (y < 0.0) p += 2 * kPi;
    } else if (y != 0.0) {
        p = y > 0 ? kPi / 2 : 3 * kPi / 2;
    } else {
        p = 0.0;
    }
    phi = p;
}

// The per-(atom, frame) terms of the sphere average: r, theta, phi rounded to
// float as the reference stages them, then cos(theta), sin(theta), sin/cos(phi)
// in FP64.
__device__ __forceinline__ void sphere_terms(double x, double y, double z, double &rf, double &ct, double &st,
                                             double &sp, double &cp) {
    double r, phi, theta;
    to_spherical(x, y, z, r, phi, theta);
    rf = (float)r;
    const double phif = (float)phi, thetaf = (float)theta;
    ct = cos(thetaf);
    st = sqrt((1.0 - ct) * (1.0 + ct));
    sincos(phif, &sp, &cp);
}

// One thread per (atom, frame): the sphere terms, formed once per call
// instead of once per (shell, moment block) wavefront.
__global__ __launch_bounds__(256) void sphere_terms_kernel(const float *__restrict__ X, const float *__restrict__ Y,
                                                           const float *__restrict__ Z, size_t n, float *srf,
                                                           double *sct, double *sst, double *ssp, double *scp) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    double rf, ct, st, sp, cp;
    sphere_terms(X[e], Y[e], Z[e], rf, ct, st, sp, cp);
    srf[e] = (float)rf;
    sct[e] = ct;
    sst[e] = st;
    ssp[e] = sp;
    scp[e] = cp;
}

// coor3d.cpp:116-140 (the pi offsets pass through a float-valued sign())
__device__ void to_cylinder(double x, double y, double z, double &r, double &phi, double &zz) {
    r = sqrt(x * x + y * y);
    double p;
    if (x != 0.0) {
        p = atan(y / x);
        if (x < 0.0) p = (double)(y < 0.0 ? -(float)kPi : (float)kPi) + p;
    } else if (y != 0.0) {
        p = (double)(y < 0.0 ? -(float)(kPi / 2) : (float)(kPi / 2));
    } else {
        p = 0.0;
    }
    phi = p < 0 ? 2 * kPi + p : p;
    zz = z;
}

// The cylinder-frame coordinates of every (atom, frame) (the frame of
// MultipoleKernelArgs::base), rounded to float as the reference stages them:
// one thread per (atom, frame), once per call.
__global__ __launch_bounds__(256) void cylinder_terms_kernel(const float *__restrict__ X, const float *__restrict__ Y,
                                                             const float *__restrict__ Z, size_t n,
                                                             const double *__restrict__ base, float *rr, float *pp,
                                                             float *zz) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double x = X[e], y = Y[e], z = Z[e];
    const double px = x * base[0] + y * base[1] + z * base[2];
    const double py = x * base[3] + y * base[4] + z * base[5];
    const double pz = x * base[6] + y * base[7] + z * base[8];
    double r, phi, zc;
    to_cylinder(px, py, pz, r, phi, zc);
    rr[e] = (float)r;
    pp[e] = (float)phi;
    zz[e] = (float)zc;
}

// j_0(x), j_1(x) and 1/x (0, 1, 0 at x = 0) from the argument's sin and cos.
__device__ __forceinline__ void sph_bessel01(double x, double &j0, double &j1, double &ix) {
    double sx = 0.0, cx = 1.0;
    ix = 0.0;
    j0 = 1.0;
    j1 = 0.0;
    if (x != 0.0) {
        sincos(x, &sx, &cx);
        ix = 1.0 / x;
        j0 = sx * ix;
        j1 = (j0 - cx) * ix;
    }
}

// j_0 and j_1 of |q| r for every (shell, atom, frame) of a sphere call, from
// the sphere terms (r rounded to float, as the amplitude kernel forms its
// argument): formed once per call instead of once per moment block.
__global__ __launch_bounds__(256) void sphere_bessel01_kernel(const float *__restrict__ rr, size_t n,
                                                              const double *__restrict__ qinfo, int ns,
                                                              double2 *__restrict__ out) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double rf = rr[e];
    for (int s = 0; s < ns; ++s) {
        double j0, j1, ix;
        sph_bessel01(qinfo[s] * rf, j0, j1, ix);
        out[(size_t)s * n + e] = make_double2(j0, j1);
    }
}

// j_0 .. j_L (L < kBesselTableL1) of x in one recurrence, as
// sph_bessel_moments forms each degree: j_0 and j_1 as given, upward to L
// where x > L, else Miller's recurrence from L + 32 + 2 sqrt(2 L) normalised
// to j_0 (or j_1 near a zero of j_0).  Loops over the table's fixed degree
// count with guards on L keep every index a constant.
__device__ __forceinline__ void sph_bessel_all(int L, double x, double ix, double j0, double j1,
                                               double (&jv)[kBesselTableL1]) {
#pragma unroll
    for (int l = 0; l < kBesselTableL1; ++l) jv[l] = 0.0;
    if (x == 0.0) {
        jv[0] = 1.0;
        return;
    }
    jv[0] = j0;
    jv[1] = j1;
    if (L < 2) return;
    if (x > (double)L) {
        double a = j0, bb = j1;  // j_{n-1}, j_n
#pragma unroll
        for (int n = 1; n + 1 < kBesselTableL1; ++n)
            if (n < L) {
                const double nx = (2.0 * n + 1.0) * ix * bb - a;
                a = bb;
                bb = nx;
                jv[n + 1] = bb;
            }
        return;
    }
    int n = L + 32 + (int)(2.0 * sqrt(2.0 * (double)L));
    double fp1 = 0.0, f = 1e-280;  // f_{n+1}, f_n
    double cap[kBesselTableL1];
#pragma unroll
    for (int l = 0; l < kBesselTableL1; ++l) cap[l] = 0.0;
    // f_n -> f_{n-1}, rescaling everything captured so far
    auto step = [&]() {
        const double fm1 = (2.0 * n + 1.0) * ix * f - fp1;
        fp1 = f;
        f = fm1;
        --n;
        if (fabs(f) > 1e250) {
            f *= 1e-250;
            fp1 *= 1e-250;
#pragma unroll
            for (int l = 0; l < kBesselTableL1; ++l) cap[l] *= 1e-250;
        }
    };
    while (n > L) step();
#pragma unroll
    for (int l = kBesselTableL1 - 1; l >= 0; --l)
        if (l <= L) {
            cap[l] = f;  // f_l
            if (l > 0) step();
        }
    // f = f_0 and fp1 = f_1, at the scale of every capture
    const double sc = fabs(j0) >= fabs(j1) ? j0 / f : j1 / fp1;
#pragma unroll
    for (int l = 2; l < kBesselTableL1; ++l)
        if (l <= L) jv[l] = cap[l] * sc;
}

// j_0 .. j_{L1-1} of |q_s| r for every (shell, atom, frame) of a sphere call
// (out[(s L1 + l) n + e], e the (atom, frame) index), from the sphere terms:
// formed once per call instead of once per moment block.
__global__ __launch_bounds__(256) void sphere_bessel_table_kernel(const float *__restrict__ rr, size_t n,
                                                                  const double *__restrict__ qinfo, int ns, int L1,
                                                                  double *__restrict__ out) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double rf = rr[e];
    for (int s = 0; s < ns; ++s) {
        double j0, j1, ix;
        sph_bessel01(qinfo[s] * rf, j0, j1, ix);
        double jv[kBesselTableL1];
        sph_bessel_all(L1 - 1, qinfo[s] * rf, ix, j0, j1, jv);
#pragma unroll
        for (int l = 0; l < kBesselTableL1; ++l)
            if (l < L1) out[((size_t)s * L1 + l) * n + e] = jv[l];
    }
}

// J_0 and J_1 of q_r r for every (shell, atom, frame) of a cylinder call, from
// the cylinder terms (r rounded to float, as the amplitude kernel forms p):
// formed once per call instead of once per moment block (the library
// Bessel functions took a third of the amplitude kernel's time).
__global__ __launch_bounds__(256) void cylinder_bessel01_kernel(const float *__restrict__ rr, size_t n,
                                                                const double *__restrict__ qinfo, int ns,
                                                                double2 *__restrict__ out) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double rf = rr[e];
    for (int s = 0; s < ns; ++s) {
        const double p = rf * qinfo[3 * s];
        out[(size_t)s * n + e] = make_double2(j0(p), j1(p));
    }
}

// J_0 .. J_nhi (nhi < NT) of x >= 0 in one recurrence, as
// cyl_bessel_orders forms each order: J_0 and J_1 as given, upward where
// x > nhi, else Miller's recurrence from an even start above nhi normalised
// by J_0 + 2 sum_k J_2k = 1.  Loops over the table's fixed order count with
// guards on nhi keep every index a constant.
template <int NT>
__device__ __forceinline__ void cyl_bessel_all(int nhi, double x, double J0, double J1, double (&jv)[NT]) {
#pragma unroll
    for (int q = 0; q < NT; ++q) jv[q] = 0.0;
    jv[0] = J0;
    jv[1] = J1;
    if (nhi < 2 || !(x > 1e-20)) return;
    const double ix = 1.0 / x;
    if (x > (double)nhi) {
        double a = J0, bb = J1;  // J_{k-1}, J_k
#pragma unroll
        for (int k = 1; k + 1 < NT; ++k)
            if (k < nhi) {
                const double nx = (2.0 * k) * ix * bb - a;
                a = bb;
                bb = nx;
                jv[k + 1] = bb;
            }
        return;
    }
    int N = nhi + 32 + (int)(2.0 * sqrt(2.0 * (double)nhi));
    N += N & 1;  // even: the normalisation sums even orders
    double fp1 = 0.0, f = 1e-280, norm = 0.0;
    // f <- J_{k-1} (unnormalised), 2 J_{k-1} into the sum when k - 1 is even and positive
    auto step = [&](int k) {
        const double fm1 = (2.0 * k) * ix * f - fp1;
        fp1 = f;
        f = fm1;
        if (((k - 1) & 1) == 0 && k > 1) norm += 2.0 * f;
        if (fabs(f) > 1e250) {
            f *= 1e-250;
            fp1 *= 1e-250;
            norm *= 1e-250;
#pragma unroll
            for (int q = 0; q < NT; ++q) jv[q] *= 1e-250;
        }
    };
    int k = N;
    for (; k > nhi + 1; --k) step(k);
    // k = nhi + 1: each step now yields an order to keep
#pragma unroll
    for (int o = NT - 1; o >= 0; --o)
        if (o <= nhi) {
            step(o + 1);  // f = J_o
            jv[o] = f;
        }
    norm += f;  // J_0
    const double inorm = 1.0 / norm;
#pragma unroll
    for (int q = 2; q < NT; ++q)
        if (q <= nhi) jv[q] *= inorm;
    jv[0] = J0;
    jv[1] = J1;
}

// J_0 .. J_{N1-1} of q_r r for every (shell, atom, frame) of a cylinder call
// (out[(s N1 + n) NAp NFp + e]), from the cylinder terms (r rounded to float):
// formed once per call instead of once per moment block.
// NT: orders the thread holds (at least N1; 21 covers the moments up to l = 10).
// Also the two phases of every (shell, atom, frame), by the amplitude
// kernel's own expressions: ex[(2 s) n + e] = (cos, sin)(sgn z q_z),
// ex[(2 s + 1) n + e] = (cos, sin)(phi - q_phi).
template <int NT>
__global__ __launch_bounds__(256, NT <= 21 ? 3 : 2) void cylinder_bessel_table_kernel(
    const float *__restrict__ rr, const float *__restrict__ cphi, const float *__restrict__ cz, size_t n,
    const double *__restrict__ qinfo, int ns, int N1, double *__restrict__ out, double2 *__restrict__ ex) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double rf = rr[e], phif = cphi[e], zf = cz[e];
    for (int s = 0; s < ns; ++s) {
        const double qphi = qinfo[3 * s + 1], qz = qinfo[3 * s + 2];
        double sgn = 1.0;
        if ((zf != 0) && (qz != 0)) sgn = (zf * qz) / (fabs(zf) * fabs(qz));
        double exs, exc, sp1, cp1;
        sincos(sgn * zf * qz, &exs, &exc);
        sincos(phif - qphi, &sp1, &cp1);
        ex[(size_t)(2 * s) * n + e] = make_double2(exc, exs);
        ex[(size_t)(2 * s + 1) * n + e] = make_double2(cp1, sp1);
        const double p = rf * qinfo[3 * s];
        double jv[NT];
        cyl_bessel_all<NT>(N1 - 1, p, j0(p), j1(p), jv);
#pragma unroll
        for (int q = 0; q < NT; ++q)
            if (q < N1) out[((size_t)s * N1 + q) * n + e] = jv[q];
    }
}

// KIND 0 sphere, 1 cylinder, each instantiated on its own.  Both are bound by
// the latency of serial FP64 recurrences, so they are held to 3 (sphere) and
// 2 (cylinder) wavefronts per SIMD, the bounds that measured fastest despite
// some spilled registers (1 wavefront per SIMD took 1.6x (sphere) and 1.8x
// (cylinder) as long).
// SRC 0: the spherical or cylinder coordinates of every (atom, frame) are
// formed here; 1: they come from sphere_terms_kernel / cylinder_terms_kernel
// (a.srf ...); 2 (cylinder): as 1, and J_0, J_1 of every (shell, atom, frame)
// come from cylinder_bessel01_kernel (a.cj01).  Each is its own
// instantiation, so the acos, atan, sincos and Bessel code of the paths a
// kernel does not take adds nothing to its registers or code.
// NSH: shells per block (sphere only): the Legendre functions and e^{i m phi}
// of an (atom, frame) serve all of them; each keeps its own Bessel functions
// and sums.
template <int KIND, int SRC, int NSH = 1>
__global__ __launch_bounds__(256, KIND == 0 ? 3 : 2) void multipole_kernel(MultipoleKernelArgs a) {
    static_assert(NSH >= 1 && (KIND == 0 || NSH == 1), "several shells per block on the sphere only");
    constexpr int kind = KIND;
    constexpr bool PRE = SRC >= 1;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // block = (frame tile, shell, group of 4 moment blocks + zgroups x atom range)
    const int ftile = blockIdx.x;
    const int s = blockIdx.y * NSH;  // first shell of the block
    const int zg = blockIdx.z % a.zgroups;
    const int sp = blockIdx.z / a.zgroups;
    const int mb = zg * 4 + wave;
    if (mb >= a.mblocks) return;  // whole wavefront exits; the kernel has no barriers
    const int k0 = mb * kMoments;
    const int nk = min(kMoments, a.nmo - k0);
    const size_t t = (size_t)ftile * 64 + lane;
    double ar[NSH][kMoments], ai[NSH][kMoments];
#pragma unroll
    for (int h = 0; h < NSH; ++h)
#pragma unroll
        for (int k = 0; k < kMoments; ++k) ar[h][k] = ai[h][k] = 0.0;
    // shells past the last (an odd count) read the last one's data with zero weight
    int sh[NSH];
    const float *brow[NSH];
#pragma unroll
    for (int h = 0; h < NSH; ++h) {
        sh[h] = min(s + h, a.ns - 1);
        brow[h] = a.b + (a.per_shell_b ? (size_t)sh[h] * a.NAp : 0);
    }
    const size_t jend = min(a.NAp, (size_t)(sp + 1) * a.range);
    for (size_t j = (size_t)sp * a.range; j < jend; ++j) {
        float bjs[NSH];
        bool any = false;
#pragma unroll
        for (int h = 0; h < NSH; ++h) {
            bjs[h] = s + h < a.ns ? brow[h][j] : 0.f;
            any = any || bjs[h] != 0.f;
        }
        if (!any) continue;  // padding atoms and zero lengths add nothing
        const float bj = bjs[0];
        // The moments' parameters, read again in every atom iteration through
        // an index the compiler cannot take for a constant: held across the
        // atom loop they fill the SGPRs and spill into VGPR lanes, with a
        // v_readlane at every use.  The reads are scalar loads of a few cached
        // words.
        int zo = 0;
        asm volatile("" : "+s"(zo));
        int ls[kMoments], ms[kMoments];
        double cms[kMoments];  // sphere: start of each moment's Legendre recurrence
#pragma unroll
        for (int k = 0; k < kMoments; ++k) {
            const int kk = min(k0 + k, a.nmo - 1) + zo;
            ls[k] = a.moments[2 * kk];
            ms[k] = a.moments[2 * kk + 1];
            cms[k] = kind == 0 ? a.ynorm[2 * kk] : 0.0;
        }
        int lhi = 0;
        bool sorted = true;  // degrees non-decreasing over the wavefront's moments
#pragma unroll
        for (int k = 0; k < kMoments; ++k) {
            lhi = max(lhi, ls[k]);
            if (k > 0 && ls[k] < ls[k - 1]) sorted = false;
        }
        double b9[9];
        // cylinder: Bessel order per moment (J_0 for (0,0), J_2l for m = 0, 1,
        // J_2l-1 for m = 2, 3) and the largest
        int cn[kMoments];
        int nhi = 0;
        if (kind == 1) {
            if (!PRE)
                for (int k = 0; k < 9; ++k) b9[k] = a.base[k + zo];
#pragma unroll
            for (int k = 0; k < kMoments; ++k) {
                const int l = ls[k], m = ms[k];
                cn[k] = (l == 0 || m > 3 || m < 0) ? 0 : (m <= 1 ? 2 * l : 2 * l - 1);
                nhi = max(nhi, cn[k]);
            }
        }
        double x = 0.0, y = 0.0, z = 0.0;
        if constexpr (!PRE) {
            x = a.X[j * a.NFp + t];
            y = a.Y[j * a.NFp + t];
            z = a.Z[j * a.NFp + t];
        }
        if (kind == 0) {
            double rf, ct, somx2, sp, cp;
            if constexpr (PRE) {
                const size_t e = j * a.NFp + t;
                rf = a.srf[e];
                ct = a.sct[e];
                somx2 = a.sst[e];
                sp = a.ssp[e];
                cp = a.scp[e];
            } else {
                sphere_terms(x, y, z, rf, ct, somx2, sp, cp);
            }
            // per shell: the moments' j_l from the table, or j_0, j_1 and 1/x
            // of the argument once for all moments and the recurrence
            double jv[NSH][kMoments];
            if constexpr (SRC == 3) {
                const size_t plane = a.NAp * a.NFp, e = j * a.NFp + t;
#pragma unroll
                for (int h = 0; h < NSH; ++h) {
                    double cur = 0.0;
                    int lp = -1;
#pragma unroll
                    for (int k = 0; k < kMoments; ++k) {
                        if (ls[k] != lp) {  // one load per distinct degree
                            lp = ls[k];
                            cur = a.jtab[((size_t)sh[h] * a.jl1 + lp) * plane + e];
                        }
                        jv[h][k] = cur;
                    }
                }
            } else {
#pragma unroll
            for (int h = 0; h < NSH; ++h) {
                const double xq = a.qinfo[sh[h]] * rf;
                double ix = 0.0, j0 = 1.0, j1 = 0.0;
                if constexpr (SRC == 2) {
                    const double2 jj = a.cj01[((size_t)sh[h] * a.NAp + j) * a.NFp + t];
                    j0 = jj.x;
                    j1 = jj.y;
                    if (xq != 0.0) ix = 1.0 / xq;
                } else {
                    sph_bessel01(xq, j0, j1, ix);
                }
                if (sorted)
                    sph_bessel_moments<true>(ls, lhi, xq, ix, j0, j1, jv[h]);
                else
                    sph_bessel_moments<false>(ls, lhi, xq, ix, j0, j1, jv[h]);
            }
            }
            // e^{i am phi} carried from moment to moment: one product by
            // e^{+-i phi} when |m| steps by one (the reference's lists),
            // repeated products from 1 otherwise
            double er = 1.0, ei = 0.0;
            int aprev = 0;
#pragma unroll
            for (int k = 0; k < kMoments; ++k) {
                if (k >= nk) break;
                const int l = ls[k], m = ms[k], am = m < 0 ? -m : m;
                // the Y_lm normalisation and 4 pi scale the finished sums
                const double th = plm_scaled(l, am, ct, somx2, cms[k], a.plm_ab, a.plm_l1);
                if (am == aprev + 1) {
                    const double nr = er * cp - ei * sp;
                    ei = er * sp + ei * cp;
                    er = nr;
                } else if (am == aprev - 1) {
                    const double nr = er * cp + ei * sp;
                    ei = ei * cp - er * sp;
                    er = nr;
                } else if (am != aprev) {
                    er = 1.0;
                    ei = 0.0;
                    for (int q = 0; q < am; ++q) {
                        const double nr = er * cp - ei * sp;
                        ei = er * sp + ei * cp;
                        er = nr;
                    }
                }
                aprev = am;
                double yr = th * er, yi = th * ei;
                if (m < 0) {
                    const double sg = (am & 1) ? -1.0 : 1.0;
                    yr = sg * yr;
                    yi = -sg * yi;
                }
                // 4 pi i^l b j_l conj(Y)
#pragma unroll
                for (int h = 0; h < NSH; ++h) {
                    const double f = (double)bjs[h] * jv[h][k];
                    const double cr = f * yr, ci = -f * yi;
                    switch (l & 3) {
                        case 0: ar[h][k] += cr; ai[h][k] += ci; break;
                        case 1: ar[h][k] -= ci; ai[h][k] += cr; break;
                        case 2: ar[h][k] -= cr; ai[h][k] -= ci; break;
                        default: ar[h][k] += ci; ai[h][k] -= cr; break;
                    }
                }
            }
        } else {
            const double qr = a.qinfo[3 * s], qphi = a.qinfo[3 * s + 1], qz = a.qinfo[3 * s + 2];
            double rf, phif, zf;
            if constexpr (PRE) {
                const size_t e = j * a.NFp + t;
                rf = a.srf[e];
                phif = a.cphi[e];
                zf = a.cz[e];
            } else {
                const double px = x * b9[0] + y * b9[1] + z * b9[2];
                const double py = x * b9[3] + y * b9[4] + z * b9[5];
                const double pz = x * b9[6] + y * b9[7] + z * b9[8];
                double r, phi, zz;
                to_cylinder(px, py, pz, r, phi, zz);
                rf = (float)r;
                phif = (float)phi;
                zf = (float)zz;
            }
            const size_t plane = a.NAp * a.NFp, e = j * a.NFp + t;
            double exs, exc;
            if constexpr (SRC == 3) {
                const double2 ee = a.etab[(size_t)(2 * s) * plane + e];
                exc = ee.x;
                exs = ee.y;
            } else {
                double sgn = 1.0;
                if ((zf != 0) && (qz != 0)) sgn = (zf * qz) / (fabs(zf) * fabs(qz));
                sincos(sgn * zf * qz, &exs, &exc);
            }
            const double p = rf * qr, psi = phif - qphi;
            double jn[kMoments];
            if constexpr (SRC == 3) {
                // the moments' orders from the table, one load per distinct order
                double cur = 0.0;
                int np = -1;
#pragma unroll
                for (int k = 0; k < kMoments; ++k) {
                    if (cn[k] != np) {
                        np = cn[k];
                        cur = a.jtab[((size_t)s * a.jl1 + np) * plane + e];
                    }
                    jn[k] = cur;
                }
            } else {
                double J0, J1;
                if constexpr (SRC == 2) {
                    const double2 jj = a.cj01[((size_t)s * a.NAp + j) * a.NFp + t];
                    J0 = jj.x;
                    J1 = jj.y;
                } else {
                    J0 = j0(p);
                    J1 = j1(p);
                }
                cyl_bessel_orders(cn, nhi, p, J0, J1, jn);
            }
            double sn = 0.0, cs = 1.0, sp1, cp1;
            if constexpr (SRC == 3) {
                const double2 pp = a.etab[(size_t)(2 * s + 1) * plane + e];
                cp1 = pp.x;
                sp1 = pp.y;
            } else {
                sincos(psi, &sp1, &cp1);  // e^{i psi}
            }
#pragma unroll
            for (int k = 0; k < kMoments; ++k) {
                if (k >= nk) break;
                const int l = ls[k], m = ms[k];
                // cos and sin of n psi per distinct order (m = 0, 1 and m = 2, 3
                // share theirs): from the previous order by up to four products
                // with e^{+-i psi} (the orders of a wavefront step by small
                // amounts, 2l, 2l - 1, 2l + 2, 2l + 1), else by sincos
                if (k == 0 || cn[k] != cn[k - 1]) {
                    const int d = k == 0 ? 0 : cn[k] - cn[k - 1];
                    if (k == 0 || d > 4 || d < -4) {
                        sincos(cn[k] * psi, &sn, &cs);
                    } else {
                        const double s1 = d > 0 ? sp1 : -sp1;
                        for (int r = d > 0 ? d : -d; r > 0; --r) {
                            const double nc = cs * cp1 - sn * s1;
                            sn = sn * cp1 + cs * s1;
                            cs = nc;
                        }
                    }
                }
                double cr = 0.0, ci = 0.0;
                const double sgl = (l & 1) ? -1.0 : 1.0;
                if (m == 0 && l == 0) {
                    cr = jn[k];
                } else if (m == 0) {
                    cr = sqrt(0.5) * 2.0 * sgl * jn[k] * cs;
                } else if (m == 1) {
                    cr = sqrt(0.5) * 2.0 * sgl * jn[k] * sn;
                } else if (m == 2) {
                    ci = sqrt(0.5) * 2.0 * (-sgl) * jn[k] * cs;
                } else if (m == 3) {
                    ci = sqrt(0.5) * 2.0 * (-sgl) * jn[k] * sn;
                }
                ar[0][k] += (double)bj * (cr * exc - ci * exs);
                ai[0][k] += (double)bj * (cr * exs + ci * exc);
            }
        }
    }
#pragma unroll
    for (int h = 0; h < NSH; ++h) {
        if (s + h >= a.ns) break;
#pragma unroll
        for (int k = 0; k < kMoments; ++k) {
            if (k >= nk) break;
            const double norm = kind == 1 ? sqrt(2.0 * kPi) : a.ynorm[2 * (k0 + k) + 1];
            const size_t v = ((size_t)sp * a.ns + s + h) * a.nmo + k0 + k;
            a.amp[v * a.NFp + t] = make_float2((float)(norm * ar[h][k]), (float)(norm * ai[h][k]));
        }
    }
}

hipError_t launch_multipole(const MultipoleArgs &m, hipStream_t stream) {
    MultipoleKernelArgs k;
    k.X = m.coords;
    k.Y = m.coords + m.NAp * m.NFp;
    k.Z = m.coords + 2 * m.NAp * m.NFp;
    k.NFp = m.NFp;
    k.NAp = m.NAp;
    k.kind = m.kind;
    k.qinfo = m.qinfo;
    k.base = m.base;
    k.moments = m.moments;
    k.ynorm = m.ynorm;
    // the recurrence coefficients follow the per-moment factors
    k.plm_l1 = m.plm_l1;
    k.plm_ab = m.ynorm ? m.ynorm + 2 * m.nmo : nullptr;
    k.nmo = m.nmo;
    k.ns = m.ns;
    k.mblocks = (m.nmo + kMoments - 1) / kMoments;
    k.b = m.b;
    k.per_shell_b = m.per_shell_b;
    k.amp = m.amp;
    k.split = std::max(m.split, 1);
    k.zgroups = (k.mblocks + 3) / 4;
    k.range = (m.NAp + k.split - 1) / k.split;
    const size_t n = m.NAp * m.NFp;
    k.srf = nullptr;
    k.sct = k.sst = k.ssp = k.scp = nullptr;
    k.cphi = k.cz = nullptr;
    k.cj01 = static_cast<const double2 *>(m.cj01);
    k.jtab = static_cast<const double *>(m.jtab);
    k.jl1 = m.jl1;
    k.etab = m.kind == 1 && m.jtab ? reinterpret_cast<const double2 *>(k.jtab + (size_t)m.ns * m.jl1 * m.NAp * m.NFp)
                                   : nullptr;
    if (m.sph && m.kind == 0) {
        k.srf = static_cast<const float *>(m.sph);
        const double *d = reinterpret_cast<const double *>(static_cast<const char *>(m.sph) + sizeof(float) * n);
        k.sct = d;
        k.sst = d + n;
        k.ssp = d + 2 * n;
        k.scp = d + 3 * n;
    } else if (m.sph) {
        k.srf = static_cast<const float *>(m.sph);
        k.cphi = k.srf + n;
        k.cz = k.srf + 2 * n;
    }
    if ((size_t)k.zgroups * k.split > 65535) return hipErrorInvalidValue;
    dim3 grid((unsigned)(m.NFp / 64), (unsigned)m.ns, (unsigned)(k.zgroups * k.split));
    if (m.kind == 0 && k.srf && k.jtab) {
        grid.y = (unsigned)((m.ns + kSphereShells - 1) / kSphereShells);
        hipLaunchKernelGGL((multipole_kernel<0, 3, kSphereShells>), grid, dim3(256), 0, stream, k);
    } else if (m.kind == 0 && k.srf && k.cj01) {
        grid.y = (unsigned)((m.ns + kSphereShells - 1) / kSphereShells);
        hipLaunchKernelGGL((multipole_kernel<0, 2, kSphereShells>), grid, dim3(256), 0, stream, k);
    } else if (m.kind == 0 && k.srf)
        hipLaunchKernelGGL((multipole_kernel<0, 1>), grid, dim3(256), 0, stream, k);
    else if (m.kind == 0)
        hipLaunchKernelGGL((multipole_kernel<0, 0>), grid, dim3(256), 0, stream, k);
    else if (k.srf && k.jtab)
        hipLaunchKernelGGL((multipole_kernel<1, 3>), grid, dim3(256), 0, stream, k);
    else if (k.srf && k.cj01)
        hipLaunchKernelGGL((multipole_kernel<1, 2>), grid, dim3(256), 0, stream, k);
    else if (k.srf)
        hipLaunchKernelGGL((multipole_kernel<1, 1>), grid, dim3(256), 0, stream, k);
    else
        hipLaunchKernelGGL((multipole_kernel<1, 0>), grid, dim3(256), 0, stream, k);
    return hipGetLastError();
}

hipError_t multipole_blocks_per_cu(int kind, int *blocks) {
    return kind == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, multipole_kernel<0, 2, kSphereShells>, 256, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, multipole_kernel<1, 2>, 256, 0);
}

int multipole_base_blocks(size_t NFp, int ns, int nmo, int kind) {
    const int mblocks = (nmo + kMoments - 1) / kMoments;
    const int rows = kind == 0 ? (ns + kSphereShells - 1) / kSphereShells : ns;
    return (int)(NFp / 64) * rows * ((mblocks + 3) / 4);
}

hipError_t launch_multipole_sphere_terms(const MultipoleArgs &m, void *sph, hipStream_t stream) {
    const size_t n = m.NAp * m.NFp;
    float *srf = static_cast<float *>(sph);
    double *d = reinterpret_cast<double *>(static_cast<char *>(sph) + sizeof(float) * n);
    const float *X = m.coords, *Y = m.coords + n, *Z = m.coords + 2 * n;
    hipLaunchKernelGGL(sphere_terms_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, X, Y, Z, n, srf,
                       d, d + n, d + 2 * n, d + 3 * n);
    return hipGetLastError();
}

hipError_t launch_multipole_cylinder_terms(const MultipoleArgs &m, void *terms, hipStream_t stream) {
    const size_t n = m.NAp * m.NFp;
    float *rr = static_cast<float *>(terms);
    const float *X = m.coords, *Y = m.coords + n, *Z = m.coords + 2 * n;
    hipLaunchKernelGGL(cylinder_terms_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, X, Y, Z, n,
                       m.base, rr, rr + n, rr + 2 * n);
    return hipGetLastError();
}

hipError_t launch_multipole_bessel_table(const MultipoleArgs &m, int L1, void *out, hipStream_t stream) {
    if (L1 < 1 || L1 > (m.kind == 0 ? kBesselTableL1 : kCylTableN1)) return hipErrorInvalidValue;
    const size_t n = m.NAp * m.NFp;
    if (m.kind == 0)
        hipLaunchKernelGGL(sphere_bessel_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                           static_cast<const float *>(m.sph), n, m.qinfo, m.ns, L1, static_cast<double *>(out));
    else {
        // the phases follow the orders: [ns][2][n] double2
        const float *rr = static_cast<const float *>(m.sph);
        double *J = static_cast<double *>(out);
        double2 *ex = reinterpret_cast<double2 *>(J + (size_t)m.ns * L1 * n);
        if (L1 <= 21)
            hipLaunchKernelGGL(cylinder_bessel_table_kernel<21>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                               stream, rr, rr + n, rr + 2 * n, n, m.qinfo, m.ns, L1, J, ex);
        else
            hipLaunchKernelGGL(cylinder_bessel_table_kernel<kCylTableN1>, dim3((unsigned)((n + 255) / 256)),
                               dim3(256), 0, stream, rr, rr + n, rr + 2 * n, n, m.qinfo, m.ns, L1, J, ex);
    }
    return hipGetLastError();
}

hipError_t launch_multipole_bessel01(const MultipoleArgs &m, void *out, hipStream_t stream) {
    const size_t n = m.NAp * m.NFp;
    if (m.kind == 0)
        hipLaunchKernelGGL(sphere_bessel01_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                           static_cast<const float *>(m.sph), n, m.qinfo, m.ns, static_cast<double2 *>(out));
    else
        hipLaunchKernelGGL(cylinder_bessel01_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                           static_cast<const float *>(m.sph), n, m.qinfo, m.ns, static_cast<double2 *>(out));
    return hipGetLastError();
}

}  // namespace sx
