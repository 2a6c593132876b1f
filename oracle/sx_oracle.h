/*
 * sx_oracle.h -- FP64 CPU restatement of Sassena's scattering-device hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/ as the parity checker, by
 * __graft_entry__.smoke() as the checker of one small GPU call, and by
 * bench.py's cpu_baseline leg as the timed host path.
 *
 * Every function follows the reference arithmetic (float coordinates promoted
 * to double, FP64 libm sin/cos, FP64 accumulators, zero padding to 2*NF and an
 * unnormalised forward/backward DFT pair) and cites the reference file:line it
 * restates.  Complex numbers are interleaved (re, im) doubles, the layout of
 * fftw_complex.
 */
#ifndef SX_ORACLE_H
#define SX_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { SXO_ALL = 0, SXO_SELF = 1 };
enum { SXO_DSP_AUTOCORRELATE = 0, SXO_DSP_SQUARE = 1, SXO_DSP_PLAIN = 2 };
enum { SXO_METHOD_FFT = 0, SXO_METHOD_DIRECT = 1 };

/* In-place unnormalised DFT of n complex values (any n >= 1).
   sign = -1 is FFTW_FORWARD (exp(-2 pi i k n / N)), +1 is FFTW_BACKWARD. */
int sxo_fft(double *data, size_t n, int sign);

/* smath.cpp:141-156: data holds 2*NF complex values (NF signal, NF zeros). */
int sxo_autocorrelate_fft(double *data, size_t NF);
/* smath.cpp:51-76: O(NF^2) correlation, with the reference's sign of the
   imaginary part (sum_k a_k * conj(a_{k+tau}) / (NF - tau)). */
int sxo_autocorrelate_direct(double *data, size_t NF);
/* smath.cpp:168-177 */
void sxo_square_elements(double *data, size_t NF);
/* all_vectors_scatter_device.cpp:232: a = (1/NF) sum_{i<NF} at[i]. */
void sxo_store_one(const double *at, size_t NF, double *a_out);

/* all_vectors_scatter_device.cpp:418-438: coords frames-major [NF][NA][3]. */
void sxo_scatter_all(const float *coords, size_t NA, size_t NF,
                     const double *b, const double q[3], double *at);
/* self_vectors_scatter_device.cpp:288-322: coords atom-major [NA][NF][3];
   writes 2*NF complex values, the second half zero. */
void sxo_scatter_self(const float *coords, size_t NF, size_t ai, double s,
                      const double q[3], double *at);

/* all_vectors_scatter_device.cpp:238-361 (single partition): one q shell.
   qv: NM orientation-scaled vectors [NM][3]; fqt: [NF][2]; fq, fq2: [2]. */
int sxo_compute_all(const float *coords_frames, size_t NA, size_t NF,
                    const double *b, const double *qv, size_t NM,
                    int dsp_type, int dsp_method, int nthreads,
                    double *fqt, double *fq, double *fq2);
/* self_vectors_scatter_device.cpp:145-239: coords atom-major [NA][NF][3]. */
int sxo_compute_self(const float *coords_atoms, size_t NA, size_t NF,
                     const double *b, const double *qv, size_t NM,
                     int dsp_type, int dsp_method, int nthreads,
                     double *fqt, double *fq, double *fq2);

/* abstract_vectors_scatter_device.cpp:112-175 for orientation types none,
   file and sphere: q_m = |q| * u_m, or q itself when no set is given.
   Returns the number of vectors written to out ([NM][3]). */
size_t sxo_init_subvectors(const double q[3], const double *u, size_t nu,
                           double *out);
/* parameters.cpp:932-944: every non-zero row is scaled to unit length. */
void sxo_normalise_rows(double *v, size_t n);
/* parameters.cpp:1142-1162: one scan of `points` vectors along `base`. */
size_t sxo_scan(double from, double to, size_t points, double exponent,
                const double base[3], double *out);

/* Multipole averaging (multipole_scatter_device.cpp): special functions,
   per-moment amplitudes and the per-shell driver (kind 0 sphere, 1 cylinder;
   moments [NMO][2] (l, m); axis only for the cylinder). */
void sxo_sph_bessel_all(int L, double x, double *j);
double sxo_ylm_theta(int l, int m, double x);
void sxo_ylm(int l, int m, double theta, double phi, double *re, double *im);
double sxo_cyl_bessel_j(int n, double x);
void sxo_cylinder_base(const double axis[3], double *base);
void sxo_scatter_mpsphere(const float *coords, size_t NA, size_t NF, const double *b, double ql, int l, int m,
                          double *at);
void sxo_scatter_mpcylinder(const float *coords, size_t NA, size_t NF, const double *b, const double q[3],
                            const double axis[3], int l, int m, double *at);
int sxo_compute_multipole(int kind, const float *coords_frames, size_t NA, size_t NF, const double *b,
                          const double q[3], const double axis[3], const int *moments, size_t NMO, int dsp_type,
                          int dsp_method, double *fqt, double *fq, double *fq2);

#ifdef __cplusplus
}
#endif
#endif
