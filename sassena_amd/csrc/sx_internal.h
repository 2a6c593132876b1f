// sx_internal.h -- shared declarations of the gfx950 hot-path kernels and the
// C-ABI host layer.  Device data layout (all in HBM, one context):
//
//   coords  float  [3][NAp][NFp]   SoA x/y/z rows, atom-major; NFp = NF rounded
//                                  up to 64 (one wavefront of frames), NAp = NA
//                                  rounded up to the atom chunk; padding is 0
//   b       float  [nb][NAp]       scattering factors (padding atoms have b = 0)
//   qrev    double [V][3]          q vectors divided by 2*pi (phase in turns)
//   amp     float2 [split][V][NFp] coherent amplitudes A_v(t) per atom range
//   ppart   double [nshells][chunks][L]   partial power spectra sum |X_f|^2
//   apart   double [nshells][chunks][4]   partial sum a, sum |a|^2 per chunk
//   out     double fqt [nshells][NF][2], fq/fq2 [nshells][2]
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace sx {

// Word 3 of a raw (unswizzled, stride 0) buffer resource on CDNA: 32-bit
// float data format, no index stride.
constexpr int kBufferDword3 = 0x00020000;

constexpr int kWave = 64;         // CDNA wavefront width
constexpr int kAtomChunk = 8;     // atoms per coordinate batch of amp_all_kernel
constexpr int kMaxPasses = 16;

// Stockham plan of one half-length FFT (Lh points); the full correlation
// length is L = 2*N1*Lh >= 2*NF.  Radices are ordered powers of two first and at
// most one odd radix last, so every stride before the last pass is a power of
// two.  N1 = 1: each half of a padded series is one Lh-point FFT in one
// workgroup's LDS.  N1 > 1 (series longer than the largest such FFT): each
// half is N1 * Lh points, transformed as N1 rows of Lh points whose inputs a
// small N1-point DFT forms on the fly (row k1 of half h holds the bins
// 2 (k1 + N1 k2) + h, k2 < Lh).
struct FftPlan {
    int Lh = 0;
    int L = 0;
    int N1 = 1;
    int T = 0;      // threads per block
    int E = 0;      // elements per thread per half
    int npass = 0;
    int radix[kMaxPasses] = {};
    int ns[kMaxPasses] = {};
    // Packed twiddles of pass p >= 1 start at toff[p]: entry toff[p] + kk*R + q
    // is exp(-2 pi i q kk / (Ns R)), kk < Ns, q < R (one contiguous run per
    // thread and pass).  ntw is the table length.
    int toff[kMaxPasses] = {};
    int ntw = 0;
};

// Returns false if NF cannot be served by a supported FFT length.
bool make_fft_plan(size_t NF, FftPlan *plan);
// Largest frame count a plan serves (16 rows of the largest one-block FFT).
constexpr int kMaxRows = 16;
constexpr size_t kMaxCorrFrames = (size_t)kMaxRows * 10240;

// ---- launchers (all asynchronous on `stream`; return hipError_t) ----

hipError_t launch_stage_frames(const float *src, size_t NA, size_t NF0, size_t nf,
                               float *coords, size_t NAp, size_t NFp, hipStream_t stream);
hipError_t launch_stage_atoms(const float *src, size_t a0, size_t na, size_t NF,
                              float *coords, size_t NAp, size_t NFp, hipStream_t stream);

// Largest |x| over n floats into *out (bits of a non-negative float); a
// fill kernel clears *out first, so the pair may be captured in a graph.
hipError_t launch_absmax(const float *x, size_t n, unsigned *out, hipStream_t stream);

struct AmpArgs {
    const float *coords;   // [3][NAp][NFp]
    size_t NAp, NFp;
    int NF;
    const double *qrev;    // [V][3]
    const float *b;        // [nb][NAp]
    const int *brow;       // [V] row of b per vector, or nullptr (row 0)
    int V;
    int vb;                // vectors per wavefront: 4, 8 or 16
    int split;             // atom ranges
    float2 *amp;           // [split][V][NFp]
};
hipError_t launch_amp_all(const AmpArgs &a, hipStream_t stream);
// Workgroups of amp_all_kernel with vb vectors per wavefront that fit one CU.
hipError_t all_blocks_per_cu(int vb, bool per_row, int *blocks);

// An exact run of shells from one shell s: shells s..last (at most
// kMaxRunShells) form an arithmetic progression to 1e-12 relative with this
// step.  A reference scan's powf points form such runs wherever the rounding
// of j / (points - 1) keeps one offset (for 40 points from 0.1 to 4.0:
// shells 10-19 and 20-39).
struct ShellRun {
    double step;   // turns per Angstrom
    int last;      // last shell of the run from this one
    int pad;
};
constexpr int kMaxRunShells = 32;

// The line a residual-corrected tile steps on: shells first + k (k < VS) at
// qmid + (k - VS/2) step plus residuals r_k, the middle shell (k = VS/2)
// exactly when it exists.  The line runs through the middle shell with the
// step of the exact run around it, or with the call's step, whichever leaves
// fewer of the two walks from the middle (up: k > VS/2, down: k < VS/2) with
// residuals; flags bit 0: the up walk carries first-order corrections, bit 1:
// the down walk does.  Stored at the tile's first shell.
struct TileLine {
    double qmid;   // turns per Angstrom
    double step;
    int flags;
    int pad;
};
constexpr int kTileCorrectUp = 1, kTileCorrectDown = 2;

// Oriented shells: v = s * NM + m has q_v = |q_s| u_m.  NAp must be a multiple
// of 16 (it is a multiple of 64).
struct OrientArgs {
    const float *coords;   // [3][NAp][NFp]
    size_t NAp, NFp;
    const double *qlen;    // [ns] |q_s| / 2 pi
    const double *u;       // [NM][3]
    const float *b;        // [ns][NAp] if per_shell_b else [NAp]
    int NM, ns, split, per_shell_b;
    int shape = 0;         // tile shape index (< kOrientShapes)
    float2 *amp;           // [split][ns*NM][NFp]
    // Shells in arithmetic progression (qlen[s] = qlen[0] + s qstep): the
    // tile evaluates the phasor of one shell and steps to the others by the
    // three-term recurrence with exp(2 pi i qstep d).  ap = 2: the same on the
    // ideal grid q0 + s qstep, each shell's residual qres[s] (radians per
    // Angstrom, |q| units) corrected to first order; one factor per atom.
    int ap = 0;
    double qstep = 0.0;
    double q0 = 0.0;
    // ap = 2: tile lines at the tiles' first shells and residuals against them
    const float *qres = nullptr;
    const TileLine *lines = nullptr;
    // Atoms summed (the first natoms of the NAp rows); 0: all NAp.  The
    // autotuner times shapes on a prefix of the atoms.
    size_t natoms = 0;
};
// How shells |q_s| (in |q| / 2 pi, turns per Angstrom) may be stepped by the
// oriented kernel.  mode 1: an arithmetic progression to 1e-12 relative;
// mode 2: a progression up to residuals r_s = q_s - (q_0 + s step) of at most
// kMaxShellResidual relative, which the kernel corrects to first order in
// 2 pi r_s d, provided that phase stays below kMaxResidualPhase for every
// coordinate (|d| <= rmax, Angstrom), so that the neglected second-order term
// is below 5e-7; else mode 0 (one phase per shell).  The single-precision
// powf of a reference q scan (parameters.cpp:1151) leaves residuals of about
// 6e-8 relative.
struct ShellGrid {
    int mode = 0;
    double q0 = 0.0, step = 0.0;    // turns per Angstrom
    std::vector<double> res;        // [ns], turns per Angstrom (mode 2)
    std::vector<ShellRun> runs;     // [ns] exact runs from each shell (mode 2)
};
// Device bytes of the mode-2 shell tables: [ns] TileLine, then [ns] float residuals.
inline size_t shell_table_bytes(int ns) { return (sizeof(TileLine) + sizeof(float)) * (size_t)ns; }
// Tile lines and residuals (radians per Angstrom) of the mode-2 tiles that
// start at firsts[t] with vs[t] shells: lines[firsts[t]], res[shells of the
// tile].  Returns the shells stepped without corrections (the middle shells
// and the shells of the walks without residuals).
int tile_lines(const double *qlen, int ns, double rmax, const ShellGrid &g, const int *firsts, const int *vs,
               int ntiles, TileLine *lines, float *res);
constexpr double kMaxShellResidual = 1e-6;
constexpr double kMaxResidualPhase = 1e-3;
constexpr double kTwoPiTurns = 6.283185307179586476925286766559;
int shell_progression(const double *qlen, int ns, double rmax, ShellGrid *g);
constexpr int kOrientShapes = 5;
hipError_t launch_amp_oriented(const OrientArgs &a, hipStream_t stream);
int oriented_base_blocks(size_t NFp, int NM, int ns, int shape);
// Workgroups of the oriented kernel of this shape (and stepping mode) that fit one CU at once.
hipError_t oriented_blocks_per_cu(int shape, bool per_row, int mode, int *blocks);
int orient_shape_vm(int shape);
int orient_shape_vs(int shape);

struct CorrArgs {
    FftPlan plan;
    int NF, NM, nshells, chunks;
    const float2 *twL;     // [L] exp(-2 pi i e / L), float (pre-twiddle of the odd bins)
    const double2 *twL64;  // [L] exp(-2 pi i e / L), double
    const float2 *twp;     // [plan.ntw] packed pass twiddles, float
    const double2 *twp64;  // [plan.ntw] packed pass twiddles, double
    const double2 *w;      // [L] weights of a = mean_tau C(tau), in last-pass order (launch_weights)
    const int *aidx;       // "self": atoms with a nonzero length (nullptr: all NA atoms)
    double *ppart;         // [nshells][chunks][L]
    double *apart;         // [nshells][chunks][4]
    // "all" inputs
    const float2 *amp; size_t NFp; int split; size_t split_stride;
    // "self" inputs
    const float *coords; size_t NAp; int NA;
    const double *qrev; const float *b; int brow_per_shell;
};
hipError_t launch_corr_all(const CorrArgs &a, hipStream_t stream);
// Plans of N1 > 1 rows: one block per (shell, chunk of series, half, row);
// each wavefront's part of a per series goes to apart_rows[shell][series][2 N1][wavefronts],
// which self_a_reduce sums per series (wpp = 2 N1 x wavefronts).  self: the
// series are (atom, vector) pairs with phases formed in the kernel, else the
// coherent amplitudes of the vectors.
hipError_t launch_corr_rows(const CorrArgs &a, bool self, double2 *apart_rows, hipStream_t stream);
inline int corr_rows_parts(const FftPlan &P) { return 2 * P.N1 * (P.T / 64); }
// Dynamic LDS of one coherent correlation block for plan P (the FP64
// transform data; the block reduction reuses its start).
size_t corr_all_lds_bytes(const FftPlan &P);
// halves: 1 transforms the two halves of a padded series one after the other
// (half the LDS, twice the resident blocks), 2 both at once.
hipError_t launch_corr_self(const CorrArgs &a, int halves, hipStream_t stream);
// The two halves of every padded series in different blocks (half the LDS
// of a block, more resident blocks): each block writes its half's bins of the
// power-spectrum rows and, per pair, each wavefront's part of a into
// apair[shell][pair][half][wavefront]; self_a_reduce then forms sum a and
// sum |a|^2 per shell into apart[shell][rblocks][4] (wpp = 2 x wavefronts).
hipError_t launch_corr_self_split(const CorrArgs &a, double2 *apair, hipStream_t stream);
// Workgroups of the split kernel for plan P that fit one CU at once.
hipError_t self_split_blocks_per_cu(const FftPlan &P, int *blocks);
// Whether the split kernel serves plan P (every one-block plan of 256
// threads or fewer, and the 512-thread ones with a fixed plan).
bool self_split_supported(const FftPlan &P);
// Direct FP64 sums S(tau) = sum_t conj(x_t) x_{t+tau} of the last min(64, NF)
// lags of every (atom, vector) pair of the self path, summed over pairs into
// tail[shell][tblocks][64] (one wavefront per pair).  The FP32 transforms
// divide their rounding error by NF - tau, which grows to NF at the last lag.
constexpr int kTailLags = 64;
hipError_t launch_self_tail(const CorrArgs &a, double2 *tail, int tblocks, hipStream_t stream);
hipError_t launch_self_a_reduce(const double2 *apair, long npairs, int wpp, int nshells, int rblocks,
                                double *apart, hipStream_t stream);

// w_f = (1/NF) sum_tau e^{+2 pi i f tau / L} / (L (NF - tau)), stored in the
// order in which the last FFT pass visits the bins: for f = 2 (k1 + N1 k2) + h,
// k2 = j + q Lh/R (R the last radix of the plan), at (h N1 + k1) Lh + q Lh/R + j.
// tw64: [L] exp(+2 pi i e / L).
hipError_t launch_weights(const FftPlan &P, int NF, const double2 *tw64, double2 *w, hipStream_t stream);

struct FinalArgs {
    int L, NF, NM, nshells, chunks;
    int achunks = 0;       // entries of apart per shell (0: chunks)
    // "self": the last tail_k lags summed directly in FP64 (tail[shell][tail_blocks][tail_k],
    // entry k = lag NF-1-k) replace the transform's values there
    const double2 *tail = nullptr;
    int tail_k = 0, tail_blocks = 0;
    double norm = 0.0;     // final factor (1/NM for vector averages); 0 means 1/NM
    const double *ppart;   // nullptr: fqt already written by the series kernel
    const double *apart;
    const double2 *tw64;   // [L] exp(+2 pi i e / L), double
    int conj;              // 1 for the direct method (conjugate result)
    double *fqt, *fq, *fq2;
};
hipError_t launch_final(const FinalArgs &a, hipStream_t stream);
// Bins of the summed power spectrum one block of final_kernel holds in LDS at a time.
constexpr int kFinalChunk = 8192;
inline size_t final_lds_bytes(int L) { return sizeof(double) * (size_t)(L < kFinalChunk ? L : kFinalChunk); }

// dsp.type square (mode 1) / plain (mode 2) on coherent amplitudes; norm is
// the final factor of fqt (1/NM for vector averages)
hipError_t launch_series_all(int mode, const float2 *amp, size_t NFp, int split,
                             size_t split_stride, int NF, int NM, int nshells, int chunks,
                             double *fqt, double *apart, double norm, hipStream_t stream);

// Multipole amplitudes (kind 0 sphere, 1 cylinder), one series per (shell, moment)
struct MultipoleArgs {
    const float *coords;
    size_t NAp, NFp;
    int kind;
    const double *qinfo;   // sphere: |q| per shell; cylinder: (q_r, phi_q, q_z) per shell
    const double *base;    // cylinder frame, 9 doubles
    const int *moments;    // [nmo][2]
    const double *ynorm = nullptr;  // sphere: [nmo][2] (c_|m|, 4 pi f_l|m|) of plm_scaled, then
                                    // [plm_l1][plm_l1][2] the coefficients of its recurrence
    int plm_l1 = 0;
    int nmo, ns;
    const float *b;
    int per_shell_b;
    float2 *amp;           // [split][ns * nmo][NFp]: one partial set per atom range
    int split = 1;
    // per (atom, frame) terms, or nullptr to form them in the kernel: sphere
    // from launch_multipole_sphere_terms ([NAp][NFp] planes: r as float, then
    // cos(theta), sin(theta), sin(phi), cos(phi) as double), cylinder from
    // launch_multipole_cylinder_terms (r, phi, z in the axis frame as float)
    const void *sph = nullptr;
    // with the terms: j_0, j_1 of |q| r (sphere) or J_0, J_1 of q_r r
    // (cylinder) per (shell, atom, frame) from launch_multipole_bessel01, or
    // nullptr to form them in the kernel
    const void *cj01 = nullptr;
    // sphere: j_l for l < jl1 per (shell, degree, atom, frame) from
    // launch_multipole_bessel_table, or nullptr
    const void *jtab = nullptr;
    int jl1 = 0;
};
hipError_t launch_multipole(const MultipoleArgs &m, hipStream_t stream);
// Degrees the sphere's Bessel table holds at most (l <= 16), and orders the
// cylinder's (n <= 32, the orders 2l and 2l - 1 of l <= 16).
constexpr int kBesselTableL1 = 17;
constexpr int kCylTableN1 = 33;
// j_0 .. j_{L1-1} (sphere) or J_0 .. J_{L1-1} (cylinder) of every (shell,
// atom, frame) of a call from its terms (m.sph), into
// out[(s L1 + l) NAp NFp + atom NFp + frame] (FP64); the cylinder's table
// then holds (cos, sin) of sgn z q_z and of phi - q_phi per (shell, 2, atom,
// frame) as double2 (32 B per (shell, atom, frame) more).
hipError_t launch_multipole_bessel_table(const MultipoleArgs &m, int L1, void *out, hipStream_t stream);
// Workgroups of the multipole kernel that fit one CU, and its workgroups per
// atom range (kind 0, sphere: kSphereShells shells per workgroup).
constexpr int kSphereShells = 2;
hipError_t multipole_blocks_per_cu(int kind, int *blocks);
int multipole_base_blocks(size_t NFp, int ns, int nmo, int kind);
// Bytes of the sphere-term planes for NAp x NFp (atom, frame) pairs.
constexpr size_t kSphereTermBytes = sizeof(float) + 4 * sizeof(double);
hipError_t launch_multipole_sphere_terms(const MultipoleArgs &m, void *sph, hipStream_t stream);
constexpr size_t kCylinderTermBytes = 3 * sizeof(float);
hipError_t launch_multipole_cylinder_terms(const MultipoleArgs &m, void *terms, hipStream_t stream);
// Needs m.sph set to the terms (r first); out: [ns][NAp][NFp] double2.
hipError_t launch_multipole_bessel01(const MultipoleArgs &m, void *out, hipStream_t stream);

// "self" with dsp.type square (mode 1) / plain (mode 2), frames in tiles of
// kMaxSeriesFrames; with more than one tile each tile's part of a per pair
// goes to parts[shell][pair][tile] (series_tiles(NF) parts per pair) for
// launch_self_a_reduce, else a.apart receives the chunk sums directly.
constexpr int kMaxSeriesFrames = 8192;
inline int series_tiles(int NF) { return (NF + kMaxSeriesFrames - 1) / kMaxSeriesFrames; }
hipError_t launch_series_self(int mode, const CorrArgs &a, double2 *rows, double *fqt, double2 *parts,
                              hipStream_t stream);

}  // namespace sx
