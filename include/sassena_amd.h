/*
 * sassena_amd.h -- C ABI of the MI355X (gfx950) scattering-device hot path.
 *
 * One context (sx_ctx) holds one staged trajectory block in HBM and computes
 * orientationally averaged intermediate scattering functions for q shells.
 * Plain C types only: no C++ types or exceptions cross this boundary.
 * Every entry point returns SX_OK (0) or a negative SX_ERR_* code; the text
 * of the last error is available from sx_last_error().  A context is driven by
 * one host thread and runs its kernels on its own HIP stream; it is not
 * thread-safe.
 *
 * Reference interfaces replaced (paths relative to the Sassena source tree):
 *   sx_create/sx_destroy   ScatterDeviceFactory::create + device destructor
 *                          (include/scatter_devices/scatter_device_factory.hpp:28-33,
 *                           src/scatter_devices/scatter_device_factory.cpp:145-200,
 *                           src/scatter_devices/all_vectors_scatter_device.cpp:32-55,134-145)
 *   sx_stage               DataStagerByFrame::stage / DataStagerByAtom::stage
 *                          (include/stager/data_stager.hpp:42-100,
 *                           src/stager/data_stager.cpp:72-118, 214-338)
 *   sx_set_factors         ScatterFactors::update (src/scatter_devices/scatter_factors.cpp:56-78)
 *   sx_compute_shell(s)    AllVectorsScatterDevice::compute / SelfVectorsScatterDevice::compute
 *                          (src/scatter_devices/all_vectors_scatter_device.cpp:238-361,
 *                           src/scatter_devices/self_vectors_scatter_device.cpp:145-239);
 *                          outputs in the layout of HDF5WriterClient::write
 *                          (include/services/file_writer_service.hpp:64)
 */
#ifndef SASSENA_AMD_H
#define SASSENA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SX_OK 0
#define SX_ERR_ARG (-1)         /* invalid argument or size */
#define SX_ERR_HIP (-2)         /* a HIP runtime call failed */
#define SX_ERR_STATE (-3)       /* call out of order (e.g. compute before stage) */
#define SX_ERR_NOMEM (-4)       /* device or host allocation failed */
#define SX_ERR_UNSUPPORTED (-5) /* configuration not supported by this build */

/* scattering.type (parameters.cpp:372, 400-402) */
#define SX_SCATTERING_ALL 0
#define SX_SCATTERING_SELF 1
/* scattering.dsp.type (parameters.cpp:375, 502-511) */
#define SX_DSP_AUTOCORRELATE 0
#define SX_DSP_SQUARE 1
#define SX_DSP_PLAIN 2
/* scattering.dsp.method (parameters.cpp:376): "fftw" or "direct".  The
   direct method's result is the complex conjugate of the fftw method's
   (smath.cpp:51-76 vs 141-156); both are reproduced. */
#define SX_METHOD_FFT 0
#define SX_METHOD_DIRECT 1
/* host coordinate layouts accepted by sx_stage (float32, Angstrom) */
#define SX_LAYOUT_FRAMES_AOS 0 /* [NF][NA][3]: DataStagerByFrame (data_stager.cpp:111-113) */
#define SX_LAYOUT_ATOMS_AOS 1  /* [NA][NF][3]: DataStagerByAtom (data_stager.cpp:327-330) */
#define SX_LAYOUT_SOA 2        /* [3][NA][NF]: the device layout, copied as is */

typedef struct sx_ctx sx_ctx;

typedef struct sx_config {
    int device;          /* HIP device ordinal */
    int scattering_type; /* SX_SCATTERING_* */
    int dsp_type;        /* SX_DSP_* */
    int dsp_method;      /* SX_METHOD_* */
    int profile;         /* nonzero: time every kernel with HIP events */
    int vector_block;    /* 0: automatic; else q vectors per wavefront (4, 8 or 16) */
    int atom_split;      /* 0: automatic; else atom ranges per amplitude tile */
    int tile_shape;      /* oriented shells: 0 autotuned, 1..5 fixed (2x8, 4x4, 1x8, 1x16,
                            1x20 orientations x shells per wavefront) */
    int fft_halves;      /* "self" autocorrelation, the two halves of a padded series: 3 in
                            different blocks (every one-block FFT length, up to 10,240 frames),
                            2 both in one block's sweep, 1 one after the other in one block;
                            0 picks 3 where memory allows, else 2 */
    int multipole_terms; /* multipole: the spherical (sphere, 36 B) or cylinder-frame
                            (cylinder, 12 B) coordinates of every (atom, frame) 1 formed once
                            per call in a pre-pass, 2 formed by the amplitude kernel itself;
                            0 picks 1 up to 16 GB of terms, else 2.  With 0 and 1 the Bessel
                            functions of every (shell, atom, frame) also come from a pre-pass
                            (8 B each; sphere j_0..j_lmax for lmax <= 16, cylinder J_0..J_nmax for
                            orders nmax <= 32) while the table stays within 96 GB and leaves
                            4 GB of device memory free; 3 is 1 without that table (the two
                            lowest only, the rest by recurrence in the kernel) */
    int reserved[6];
} sx_config;

/* Per-kernel timing of the last sx_compute_* call (profile != 0). */
typedef struct sx_timing {
    double amplitude_ms;   /* amplitude kernel ("all") or fused amplitude+FFT ("self") */
    double correlate_ms;   /* per-vector FFT power spectrum kernel ("all") */
    double finalize_ms;    /* shell transform, normalisation and fq/fq2 */
    double total_ms;       /* first launch to last completion, device clock */
    size_t amplitude_launches;
    size_t correlate_launches;
    size_t terms;          /* atom*frame*vector terms of the call */
    size_t pad_length;     /* FFT length L used for the time correlation */
    int tile_shape;        /* oriented shells: tile shape of the last amplitude launch (1..5), else 0 */
    int shell_progression; /* oriented shells: 1 if the last launch stepped phasors across shells in
                              arithmetic progression, else 0 */
    int shell_residuals;   /* oriented shells: 1 if that progression held only up to small residuals
                              (a reference q scan rounded by single-precision powf,
                              parameters.cpp:1151), corrected to first order per shell, else 0 */
    int shell_exact_runs;  /* oriented shells with residuals: shells of the last launch stepped without
                              first-order corrections, because they lie on their tile's line (an exact
                              progression run of the scan through the tile's middle shell) */
} sx_timing;

const char *sx_version(void);

int sx_create(const sx_config *cfg, sx_ctx **out);
void sx_destroy(sx_ctx *ctx);
const char *sx_last_error(const sx_ctx *ctx);

/* Copies NA atoms x NF frames of float32 coordinates into device HBM as SoA
   x/y/z rows, atom-major.  The caller keeps ownership of xyz.  Replaces any
   previously staged block. */
int sx_stage(sx_ctx *ctx, const float *xyz, size_t NA, size_t NF, int layout);

/* Scattering factor b_j of every staged atom (same for all shells). */
int sx_set_factors(sx_ctx *ctx, const double *b, size_t NA);

/* One q shell: qv holds NM orientation-scaled vectors [NM][3] (the output of
   AbstractVectorsScatterDevice::init_subvectors).  fqt receives NF complex
   values interleaved (re, im), fq and fq2 one complex value each: exactly the
   arguments of HDF5WriterClient::write (fq0 is fqt[0]). */
int sx_compute_shell(sx_ctx *ctx, const double *qv, size_t NM,
                     double *fqt, double *fq, double *fq2);

/* nshells shells of NM vectors each in one batched launch sequence.
   qv: [nshells][NM][3]; b_shell: NULL (use sx_set_factors) or
   [nshells][NA] factors per shell (X-ray form factors depend on |q|);
   fqt: [nshells][NF][2]; fq, fq2: [nshells][2]. */
int sx_compute_shells(sx_ctx *ctx, size_t nshells, const double *qv, size_t NM,
                      const double *b_shell, double *fqt, double *fq, double *fq2);

/* Shells that share one orientation set, the reference's orientation types
   "file" and "sphere" (abstract_vectors_scatter_device.cpp:115-129): vector m
   of shell s is q_{s,m} = qlen[s] * u[m].  u: [NM][3] unit rows (normalised
   as parameters.cpp:932-944 does), qlen: [nshells] (|q| of each shell);
   other arguments as sx_compute_shells.  Same results as sx_compute_shells
   on the expanded vectors, with one FP64 product per term for the phase. */
int sx_compute_shells_oriented(sx_ctx *ctx, size_t nshells, const double *qlen, const double *u, size_t NM,
                               const double *b_shell, double *fqt, double *fq, double *fq2);

/* Allocation-free variant of sx_compute_shells_oriented for coherent ("all")
   contexts.  No counterpart in the reference: it is for callers that cannot
   allocate device memory during the call (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  The call's work
   buffers (q tables, partial amplitudes, power spectra, outputs before the
   copy to the host) are slices of one file-scope device array of 256 MiB per
   GPU, held until the process ends; the staged coordinates, factors and FFT
   tables are those sx_stage / sx_set_factors set up.  Shells are processed
   in as many passes as that maximum requires; a call whose single shell needs
   more (NM x NF x 8 bytes of amplitudes and more) fails with SX_ERR_NOMEM
   before anything runs.  Tile shapes are not autotuned (cfg.tile_shape, else
   the default rule).  Every call on one GPU shares that GPU's array: calls
   that overlap on one GPU (two host threads, two contexts) race.  Arguments
   and results as sx_compute_shells_oriented. */
int sx_compute_shells_oriented_noalloc(sx_ctx *ctx, size_t nshells, const double *qlen, const double *u, size_t NM,
                                       const double *b_shell, double *fqt, double *fq, double *fq2);

/* Multipole orientational averaging (scattering.average.orientation.type =
   multipole), coherent "all" scattering only: MPSphereScatterDevice and
   MPCylinderScatterDevice (src/scatter_devices/multipole_scatter_device.cpp:
   428-498, 900-987).  qv: [nshells][3] q vectors; moments: [nmo][2] (major,
   minor) pairs as the moments parameters give them (parameters.cpp:1037-1122);
   axis: the orientation axis (cylinder only).  Outputs as sx_compute_shells;
   the final factor is 1/(4 pi) (sphere) or 1/(2 pi) (cylinder). */
#define SX_MULTIPOLE_SPHERE 0
#define SX_MULTIPOLE_CYLINDER 1
int sx_compute_multipole(sx_ctx *ctx, int kind, size_t nshells, const double *qv, const double *axis,
                         const int *moments, size_t nmo, const double *b_shell, double *fqt, double *fq,
                         double *fq2);

/* Timing of the last compute call (zeros unless cfg.profile was set). */
int sx_get_timing(const sx_ctx *ctx, sx_timing *out);

/* FFT length the context uses for NF frames (>= 2NF), 0 if unsupported. */
size_t sx_pad_length(size_t NF);

/* ------------------------------------------------------------------------
   Batched trajectories.  Many independent coherent "all" trajectories of
   uneven size (atoms, frames, shells, orientations) run as one captured HIP
   graph: a persistent amplitude kernel draining a per-XCD work queue of
   (trajectory, frame tile, orientation/shell group, atom range) tiles, then
   the correlation and finalisation of every trajectory's shells.  Each
   trajectory follows sx_compute_shells_oriented semantics.  There is no
   single counterpart in the reference, which runs one trajectory per
   invocation (AbstractScatterDevice::run, abstract_scatter_device.cpp:105-143);
   per trajectory the results are those of AllVectorsScatterDevice::compute. */

typedef struct sx_batch sx_batch;

typedef struct sx_trajectory {
    const float *xyz;     /* host coordinates, layout as sx_stage */
    size_t NA, NF;
    int layout;
    const double *b;      /* [NA] scattering factors */
    const double *qlen;   /* [nshells] |q| */
    size_t nshells;
    const double *u;      /* [NM][3] unit orientations */
    size_t NM;
    double *fqt;          /* outputs: [nshells][NF][2], [nshells][2], [nshells][2] */
    double *fq;
    double *fq2;
} sx_trajectory;

typedef struct sx_batch_timing {
    double total_ms;      /* graph launch to completion (HIP events) */
    double amplitude_ms;  /* persistent amplitude kernel (last sx_batch_time_amplitude) */
    size_t terms;         /* sum over trajectories of shells*NM*NA*NF */
    size_t tiles;         /* amplitude work items in the queue */
    int tile_shape;       /* oriented tile shape used: 0..4 the fixed shapes of cfg.tile_shape 1..5,
                             5 mixed (1x8, 1x16 and 1x20 tiles chosen per trajectory) */
    size_t terms_stepped;  /* terms of trajectories whose shells step phasors (exact progression) */
    size_t terms_residual; /* terms of trajectories stepped with residual corrections */
    size_t terms_exact_runs; /* of those, terms of tiles whose shells all lie on their tile's line
                                (an exact run of the scan), stepped without corrections */
} sx_batch_timing;

int sx_batch_create(const sx_config *cfg, sx_batch **out);
void sx_batch_destroy(sx_batch *batch);
const char *sx_batch_last_error(const sx_batch *batch);
/* Stages every trajectory, plans the amplitude tiles and captures the batch
   graph.  cfg.tile_shape 1..5 runs every trajectory with that fixed shape;
   0 covers each trajectory's shells by the cheapest mix of 1x8, 1x16 and
   1x20 tiles (one persistent kernel runs all three).  The trajectory array (and its output
   buffers) must stay valid until the last sx_batch_run* call. */
int sx_batch_stage(sx_batch *batch, const sx_trajectory *trajs, size_t ntraj);
/* One launch of the whole batch; outputs are copied to every trajectory. */
int sx_batch_run(sx_batch *batch);
/* Trajectory k alone through the same kernels (its own captured graph);
   *ms receives its device latency. */
int sx_batch_run_one(sx_batch *batch, size_t k, double *ms);
int sx_batch_get_timing(const sx_batch *batch, sx_batch_timing *out);
/* The persistent amplitude kernel of the whole batch alone, outside the
   graph (its amplitudes are recomputed, nothing else runs); *ms receives its
   device time.  For per-kernel rooflines. */
int sx_batch_time_amplitude(sx_batch *batch, double *ms);

/* ------------------------------------------------------------------------
   The network of the learned control variate.  No counterpart in the
   reference.  A 3-layer MLP, 16 inputs (zero padded), two hidden tanh layers
   of width 128, one output; W1 and W2 are BF16 MFMA operands with FP32
   accumulation, the rest is FP32.  Parameters are one flat FP32 array:
   W1[16][128], b1[128], W2[128][128], b2[128], W3[128], b3 (row-major, input
   index first).  Training minimises the mean squared error with Adam
   (beta 0.9 / 0.999, cosine-decayed rate) on batches drawn on the device. */

typedef struct sx_mlp sx_mlp;

size_t sx_mlp_param_count(void);
int sx_mlp_create(int device, unsigned long long seed, sx_mlp **out);
void sx_mlp_destroy(sx_mlp *net);
const char *sx_mlp_last_error(const sx_mlp *net);
int sx_mlp_set_params(sx_mlp *net, const float *params, size_t count);
int sx_mlp_get_params(sx_mlp *net, float *params, size_t count);
/* x: [rows][16], t: [rows]; grads: [param_count] of the mean squared error
   over the rows; *loss receives that error. */
int sx_mlp_gradients(sx_mlp *net, const float *x, const float *t, size_t rows, float *grads, float *loss);
int sx_mlp_train(sx_mlp *net, const float *x, const float *t, size_t rows, int steps, int batch, float lr,
                 unsigned long long seed, float *final_loss);
int sx_mlp_predict(sx_mlp *net, const float *x, size_t rows, float *y);

#ifdef __cplusplus
}
#endif

#endif /* SASSENA_AMD_H */
