/*
 * sx_oracle.c -- FP64 CPU restatement of Sassena's scattering-device hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
 * timed host path of the benchmark's CPU baseline.  The product library never
 * links or calls it.
 *
 * Parity pins (tests/test_oracle.py): sxo_fft against a naive DFT and numpy's
 * pocketfft; sxo_autocorrelate_fft against sxo_autocorrelate_direct (which
 * returns the complex conjugate, exactly as the reference's two methods do);
 * the compute drivers against closed forms (static atoms, single atoms,
 * Gaussian random walk, Debye limit) and against golden vectors written by an
 * independent numpy implementation (tests/golden/make_golden.py).
 *
 * Compiled with -ffp-contract=off so that q.r and the accumulations round as
 * the reference's x86-64 baseline build does.
 */
#include "sx_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define SXO_PI 3.14159265358979323846

/* ---------------------------------------------------------------- DFT --- */

static size_t smallest_factor(size_t n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (size_t p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

static size_t largest_factor(size_t n) {
    size_t best = 1;
    while (n > 1) {
        size_t p = smallest_factor(n);
        if (p > best) best = p;
        n /= p;
    }
    return best;
}

typedef struct {
    const double *tw; /* exp(sign 2 pi i e / N), e < N, interleaved */
    size_t N;
    double *tmp;      /* 4 * max radix doubles */
} fft_plan_t;

/* out[k] = sum_j in[j*istride] * W_n^{jk}; out is contiguous, n complex. */
static void fft_rec(const fft_plan_t *pl, const double *in, size_t istride,
                    double *out, size_t n) {
    if (n == 1) {
        out[0] = in[0];
        out[1] = in[1];
        return;
    }
    size_t p = smallest_factor(n);
    size_t m = n / p;
    for (size_t r = 0; r < p; ++r)
        fft_rec(pl, in + 2 * r * istride, istride * p, out + 2 * r * m, m);
    size_t step = pl->N / n; /* W_n^e = W_N^{e*step} */
    double *y = pl->tmp, *x = pl->tmp + 2 * p;
    for (size_t k = 0; k < m; ++k) {
        for (size_t r = 0; r < p; ++r) {
            y[2 * r] = out[2 * (r * m + k)];
            y[2 * r + 1] = out[2 * (r * m + k) + 1];
        }
        for (size_t s = 0; s < p; ++s) {
            double ar = 0.0, ai = 0.0;
            size_t kk = k + s * m;
            for (size_t r = 0; r < p; ++r) {
                size_t e = ((r * kk) % n) * step;
                double wr = pl->tw[2 * e], wi = pl->tw[2 * e + 1];
                ar += wr * y[2 * r] - wi * y[2 * r + 1];
                ai += wr * y[2 * r + 1] + wi * y[2 * r];
            }
            x[2 * s] = ar;
            x[2 * s + 1] = ai;
        }
        for (size_t s = 0; s < p; ++s) {
            out[2 * (k + s * m)] = x[2 * s];
            out[2 * (k + s * m) + 1] = x[2 * s + 1];
        }
    }
}

int sxo_fft(double *data, size_t n, int sign) {
    if (n == 0) return 0;
    double *tw = (double *)malloc(sizeof(double) * 2 * n);
    double *out = (double *)malloc(sizeof(double) * 2 * n);
    size_t pmax = largest_factor(n);
    double *tmp = (double *)malloc(sizeof(double) * 4 * (pmax + 1));
    if (!tw || !out || !tmp) {
        free(tw); free(out); free(tmp);
        return -1;
    }
    for (size_t e = 0; e < n; ++e) {
        double ang = 2.0 * SXO_PI * (double)e / (double)n;
        tw[2 * e] = cos(ang);
        tw[2 * e + 1] = (sign < 0 ? -1.0 : 1.0) * sin(ang);
    }
    fft_plan_t pl = {tw, n, tmp};
    fft_rec(&pl, data, 1, out, n);
    memcpy(data, out, sizeof(double) * 2 * n);
    free(tw); free(out); free(tmp);
    return 0;
}

/* ------------------------------------------------------------- smath --- */

/* smath.cpp:141-156: forward FFT of length 2NF, |X|^2 with the imaginary
   part zeroed, backward FFT, scale 1/(2NF (NF - i)) for i < NF. */
int sxo_autocorrelate_fft(double *data, size_t NF) {
    size_t L = 2 * NF;
    if (sxo_fft(data, L, -1)) return -1;
    for (size_t i = 0; i < L; ++i) {
        data[2 * i] = data[2 * i] * data[2 * i] + data[2 * i + 1] * data[2 * i + 1];
        data[2 * i + 1] = 0.0;
    }
    if (sxo_fft(data, L, +1)) return -1;
    for (size_t i = 0; i < NF; ++i) {
        double factor = (1.0 / (double)(2 * NF * (NF - i)));
        data[2 * i] *= factor;
        data[2 * i + 1] *= factor;
    }
    return 0;
}

/* smath.cpp:51-76: note the imaginary part is -a1r*a2i + a1i*a2r. */
int sxo_autocorrelate_direct(double *data, size_t NF) {
    double *loc = (double *)malloc(sizeof(double) * 2 * (NF ? NF : 1));
    if (!loc) return -1;
    memcpy(loc, data, sizeof(double) * 2 * NF);
    for (size_t tau = 0; tau < NF; ++tau) {
        double re = 0.0, im = 0.0;
        size_t last = NF - tau;
        for (size_t k = 0; k < last; ++k) {
            const double *a1 = loc + 2 * k, *a2 = loc + 2 * (k + tau);
            re += a1[0] * a2[0] + a1[1] * a2[1];
            im += -a1[0] * a2[1] + a1[1] * a2[0];
        }
        data[2 * tau] = re / (double)last;
        data[2 * tau + 1] = im / (double)last;
    }
    free(loc);
    return 0;
}

/* smath.cpp:168-177 */
void sxo_square_elements(double *data, size_t NF) {
    for (size_t n = 0; n < NF; ++n) {
        double r = data[2 * n] * data[2 * n] + data[2 * n + 1] * data[2 * n + 1];
        data[2 * n] = r;
        data[2 * n + 1] = 0.0;
    }
}

static int dsp(double *at, size_t NF, int dsp_type, int dsp_method) {
    /* all_vectors_scatter_device.cpp:209-229, self_vectors_scatter_device.cpp:115-135 */
    if (dsp_type == SXO_DSP_AUTOCORRELATE) {
        if (dsp_method == SXO_METHOD_DIRECT) return sxo_autocorrelate_direct(at, NF);
        if (dsp_method == SXO_METHOD_FFT) return sxo_autocorrelate_fft(at, NF);
        return -2;
    }
    if (dsp_type == SXO_DSP_SQUARE) {
        sxo_square_elements(at, NF);
        return 0;
    }
    if (dsp_type == SXO_DSP_PLAIN) return 0;
    return -2;
}

/* all_vectors_scatter_device.cpp:231-236 with smath::reduce (smath.cpp:239-247)
   and smath::add_elements (smath.cpp:209-215). */
void sxo_store_one(const double *at, size_t NF, double *a_out) {
    double sr = 0.0, si = 0.0;
    for (size_t i = 0; i < NF; ++i) {
        sr += at[2 * i];
        si += at[2 * i + 1];
    }
    a_out[0] = sr * (1.0 / (double)NF);
    a_out[1] = si * (1.0 / (double)NF);
}

/* -------------------------------------------------------- amplitudes --- */

void sxo_scatter_all(const float *coords, size_t NA, size_t NF,
                     const double *b, const double q[3], double *at) {
    double qx = q[0], qy = q[1], qz = q[2];
    for (size_t fi = 0; fi < NF; ++fi) {
        const float *p_data = coords + fi * NA * 3;
        double Ar = 0.0, Ai = 0.0;
        for (size_t j = 0; j < NA; ++j) {
            double esf = b[j];
            float x = p_data[3 * j], y = p_data[3 * j + 1], z = p_data[3 * j + 2];
            double p = x * qx + y * qy + z * qz;
            Ar += esf * cos(p);
            Ai += esf * sin(p);
        }
        at[2 * fi] = Ar;
        at[2 * fi + 1] = Ai;
    }
}

void sxo_scatter_self(const float *coords, size_t NF, size_t ai, double s,
                      const double q[3], double *at) {
    double qx = q[0], qy = q[1], qz = q[2];
    const float *p_data = coords + ai * NF * 3;
    for (size_t j = 0; j < NF; ++j) {
        float x1 = p_data[j * 3], y1 = p_data[j * 3 + 1], z1 = p_data[j * 3 + 2];
        double p1 = x1 * qx + y1 * qy + z1 * qz;
        double sp1 = sin(p1), cp1 = cos(p1);
        at[2 * j] = s * cp1;
        at[2 * j + 1] = s * sp1;
    }
    memset(at + 2 * NF, 0, sizeof(double) * 2 * NF);
}

/* ----------------------------------------------------------- drivers --- */

typedef struct {
    int kind;
    const float *coords;
    size_t NA, NF, NM;
    const double *b, *qv;
    int dsp_type, dsp_method;
    size_t first, count;     /* vectors (all) or atoms (self) of this worker */
    double *C;               /* all: per-vector C_m [NM][NF][2] */
    double *a;               /* all: per-vector a_m [NM][2] */
    double *at_sum;          /* self: thread-local sum of C [NF][2] */
    double a_sum[2], a2_sum[2];
    int err;
} work_t;

static void *worker(void *arg) {
    work_t *w = (work_t *)arg;
    size_t NF = w->NF;
    double *buf = (double *)calloc(4 * NF + 2, sizeof(double));
    if (!buf) { w->err = -1; return NULL; }
    if (w->kind == SXO_ALL) {
        for (size_t m = w->first; m < w->first + w->count; ++m) {
            /* scatter (:388-438) then alignpad (:186-207): NF values, NF zeros */
            memset(buf, 0, sizeof(double) * 4 * NF);
            sxo_scatter_all(w->coords, w->NA, NF, w->b, w->qv + 3 * m, buf);
            int e = dsp(buf, NF, w->dsp_type, w->dsp_method);
            if (e) { w->err = e; break; }
            memcpy(w->C + 2 * NF * m, buf, sizeof(double) * 2 * NF);
            sxo_store_one(buf, NF, w->a + 2 * m);
        }
    } else {
        /* self_vectors_scatter_device.cpp:167-193: atoms outer, vectors inner */
        for (size_t n = w->first; n < w->first + w->count; ++n) {
            for (size_t m = 0; m < w->NM; ++m) {
                sxo_scatter_self(w->coords, NF, n, w->b[n], w->qv + 3 * m, buf);
                int e = dsp(buf, NF, w->dsp_type, w->dsp_method);
                if (e) { w->err = e; goto done; }
                double a[2];
                sxo_store_one(buf, NF, a);
                w->a_sum[0] += a[0];
                w->a_sum[1] += a[1];
                /* a*conj(a): imaginary part is exactly zero */
                w->a2_sum[0] += a[0] * a[0] + a[1] * a[1];
                for (size_t i = 0; i < NF; ++i) {
                    w->at_sum[2 * i] += buf[2 * i];
                    w->at_sum[2 * i + 1] += buf[2 * i + 1];
                }
            }
        }
    }
done:
    free(buf);
    return NULL;
}

static int run_workers(work_t *proto, size_t nitems, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if ((size_t)nthreads > nitems) nthreads = (int)(nitems ? nitems : 1);
    work_t *ws = (work_t *)calloc((size_t)nthreads, sizeof(work_t));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    if (!ws || !th) { free(ws); free(th); return -1; }
    size_t base = nitems / (size_t)nthreads, rem = nitems % (size_t)nthreads, off = 0;
    int err = 0;
    for (int t = 0; t < nthreads; ++t) {
        ws[t] = *proto;
        ws[t].first = off;
        ws[t].count = base + ((size_t)t < rem ? 1 : 0);
        off += ws[t].count;
        if (proto->kind == SXO_SELF) {
            ws[t].at_sum = (double *)calloc(2 * proto->NF + 2, sizeof(double));
            if (!ws[t].at_sum) err = -1;
        }
    }
    if (!err) {
        for (int t = 0; t < nthreads; ++t)
            if (pthread_create(&th[t], NULL, worker, &ws[t])) {
                /* fall back to running in the caller */
                worker(&ws[t]);
                th[t] = 0;
            }
        for (int t = 0; t < nthreads; ++t)
            if (th[t]) pthread_join(th[t], NULL);
    }
    if (!err && proto->kind == SXO_SELF) {
        /* reduce thread partials in thread order */
        for (int t = 0; t < nthreads; ++t) {
            for (size_t i = 0; i < 2 * proto->NF; ++i) proto->at_sum[i] += ws[t].at_sum[i];
            proto->a_sum[0] += ws[t].a_sum[0];
            proto->a_sum[1] += ws[t].a_sum[1];
            proto->a2_sum[0] += ws[t].a2_sum[0];
            proto->a2_sum[1] += ws[t].a2_sum[1];
        }
    }
    for (int t = 0; t < nthreads; ++t) {
        if (ws[t].err) err = ws[t].err;
        free(ws[t].at_sum);
    }
    free(ws);
    free(th);
    return err;
}

int sxo_compute_all(const float *coords_frames, size_t NA, size_t NF,
                    const double *b, const double *qv, size_t NM,
                    int dsp_type, int dsp_method, int nthreads,
                    double *fqt, double *fq, double *fq2) {
    if (NF == 0 || NM == 0) return -3;
    work_t w;
    memset(&w, 0, sizeof w);
    w.kind = SXO_ALL;
    w.coords = coords_frames;
    w.NA = NA; w.NF = NF; w.NM = NM;
    w.b = b; w.qv = qv;
    w.dsp_type = dsp_type; w.dsp_method = dsp_method;
    w.C = (double *)malloc(sizeof(double) * 2 * NF * NM);
    w.a = (double *)malloc(sizeof(double) * 2 * NM);
    if (!w.C || !w.a) { free(w.C); free(w.a); return -1; }
    int err = run_workers(&w, NM, nthreads);
    if (!err) {
        /* store() in vector order (:231-236), then the 1/NM scale (:355-360) */
        double afr = 0, afi = 0, a2r = 0, a2i = 0;
        memset(fqt, 0, sizeof(double) * 2 * NF);
        for (size_t m = 0; m < NM; ++m) {
            double ar = w.a[2 * m], ai = w.a[2 * m + 1];
            afr += ar; afi += ai;
            a2r += ar * ar + ai * ai;
            const double *C = w.C + 2 * NF * m;
            for (size_t i = 0; i < 2 * NF; ++i) fqt[i] += C[i];
        }
        double factor = 1.0 / (double)NM;
        for (size_t i = 0; i < 2 * NF; ++i) fqt[i] *= factor;
        fq[0] = afr * factor; fq[1] = afi * factor;
        fq2[0] = a2r * factor; fq2[1] = a2i * factor;
    }
    free(w.C);
    free(w.a);
    return err;
}

int sxo_compute_self(const float *coords_atoms, size_t NA, size_t NF,
                     const double *b, const double *qv, size_t NM,
                     int dsp_type, int dsp_method, int nthreads,
                     double *fqt, double *fq, double *fq2) {
    if (NF == 0 || NM == 0) return -3;
    work_t w;
    memset(&w, 0, sizeof w);
    w.kind = SXO_SELF;
    w.coords = coords_atoms;
    w.NA = NA; w.NF = NF; w.NM = NM;
    w.b = b; w.qv = qv;
    w.dsp_type = dsp_type; w.dsp_method = dsp_method;
    w.at_sum = (double *)calloc(2 * NF + 2, sizeof(double));
    if (!w.at_sum) return -1;
    int err = run_workers(&w, NA, nthreads);
    if (!err) {
        /* self_vectors_scatter_device.cpp:233-238: scale by 1/NM only */
        double factor = 1.0 / (double)NM;
        for (size_t i = 0; i < 2 * NF; ++i) fqt[i] = w.at_sum[i] * factor;
        fq[0] = w.a_sum[0] * factor; fq[1] = w.a_sum[1] * factor;
        fq2[0] = w.a2_sum[0] * factor; fq2[1] = w.a2_sum[1] * factor;
    }
    free(w.at_sum);
    return err;
}

/* ------------------------------------------------------- q vectors --- */

size_t sxo_init_subvectors(const double q[3], const double *u, size_t nu,
                           double *out) {
    if (nu == 0) {
        out[0] = q[0]; out[1] = q[1]; out[2] = q[2];
        return 1;
    }
    /* CartesianCoor3D::length (coor3d.cpp:58-60) */
    double ql = sqrt(pow(q[0], 2) + pow(q[1], 2) + pow(q[2], 2));
    for (size_t i = 0; i < nu; ++i) {
        out[3 * i] = ql * u[3 * i];
        out[3 * i + 1] = ql * u[3 * i + 1];
        out[3 * i + 2] = ql * u[3 * i + 2];
    }
    return nu;
}

void sxo_normalise_rows(double *v, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        double *r = v + 3 * i;
        double ql = sqrt(pow(r[0], 2) + pow(r[1], 2) + pow(r[2], 2));
        if (ql != 0) {
            double s = 1.0 / ql;
            r[0] *= s; r[1] *= s; r[2] *= s;
        }
    }
}

size_t sxo_scan(double from, double to, size_t points, double exponent,
                const double base[3], double *out) {
    size_t n = 0;
#define PUSH(s) do { out[3*n] = (s)*base[0]; out[3*n+1] = (s)*base[1]; out[3*n+2] = (s)*base[2]; ++n; } while (0)
    if (points == 0) return 0;
    if (points == 1) { PUSH((from + to) / 2); return n; }
    if (points == 2) { PUSH(from); PUSH(to); return n; }
    PUSH(from);
    for (size_t j = 1; j < points - 1; ++j) {
        /* single-precision powf, as at parameters.cpp:1151 */
        double scal = from + (double)powf((float)(j * 1.0 / (double)(points - 1)), (float)exponent) * (to - from);
        PUSH(scal);
    }
    PUSH(to);
#undef PUSH
    return n;
}

/* ------------------------------------------------------- multipole --- */

/* Spherical Bessel functions j_0..j_L at x >= 0 (the role of
   boost::math::sph_bessel in multipole_scatter_device.cpp:482): upward
   recurrence where it is stable (x > L), otherwise Miller's downward
   recurrence normalised to j_0 (or j_1 near a zero of j_0). */
void sxo_sph_bessel_all(int L, double x, double *j) {
    if (x == 0.0) {
        j[0] = 1.0;
        for (int l = 1; l <= L; ++l) j[l] = 0.0;
        return;
    }
    const double s = sin(x), c = cos(x);
    const double j0 = s / x, j1 = s / (x * x) - c / x;
    if (x > (double)L) {
        j[0] = j0;
        if (L >= 1) j[1] = j1;
        for (int l = 1; l < L; ++l) j[l + 1] = (2.0 * l + 1.0) / x * j[l] - j[l - 1];
        return;
    }
    int N = L + 32 + (int)(2.0 * sqrt((double)L + x));
    double fp1 = 0.0, f = 1e-280;
    double scale_hist = 1.0;
    (void)scale_hist;
    for (int l = N; l >= 0; --l) {
        const double fm1 = (2.0 * l + 1.0) / x * f - fp1;  /* f_{l-1} */
        if (l <= L) j[l] = f;
        fp1 = f;
        f = fm1;
        if (fabs(f) > 1e250) {
            /* rescale the running values and those stored */
            const double r = 1e-250;
            f *= r;
            fp1 *= r;
            for (int k = l; k <= L; ++k) j[k] *= r;
        }
    }
    /* j[] now holds unnormalised values for l = 0..L */
    const double norm = fabs(j0) >= fabs(j1) ? j0 / j[0] : j1 / j[1];
    for (int l = 0; l <= L; ++l) j[l] *= norm;
}

/* Normalised associated Legendre term of Y_l^m for m >= 0 (Condon-Shortley
   phase): sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m(x). */
double sxo_ylm_theta(int l, int m, double x) {
    /* N_lm P_l^m(x), the theta part of boost::math::spherical_harmonic
       (multipole_scatter_device.cpp:480), by the upward recurrence in l from
       N_mm P_m^m = (-1)^m sqrt((2m+1)/(4 pi) prod_{i<=m} (2i-1)/(2i)) (1-x^2)^{m/2},
       then the factor sqrt((2l+1)/(2m+1) (l-m)! (2m)! / (l+m)!) from log-gamma:
       no factorial is formed, so it stays finite to high degree. */
    const double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double cm = (2.0 * m + 1.0) / (4.0 * SXO_PI);
    for (int i = 1; i <= m; ++i) cm *= (2.0 * i - 1.0) / (2.0 * i);
    double pmm = sqrt(cm);
    for (int i = 1; i <= m; ++i) pmm *= -somx2;
    double plm = pmm;
    if (l > m) {
        double pmmp1 = x * (2.0 * m + 1.0) * pmm;
        plm = pmmp1;
        for (int ll = m + 2; ll <= l; ++ll) {
            const double pll = ((2.0 * ll - 1.0) * x * pmmp1 - (ll + m - 1.0) * pmm) / (ll - m);
            pmm = pmmp1;
            pmmp1 = pll;
            plm = pll;
        }
    }
    const double lf = 0.5 * (log((2.0 * l + 1.0) / (2.0 * m + 1.0)) + lgamma(l - m + 1.0) + lgamma(2.0 * m + 1.0) -
                             lgamma(l + m + 1.0));
    return exp(lf) * plm;
}

/* Y_l^m(theta, phi) for any |m| <= l; Y_l^{-m} = (-1)^m conj(Y_l^m). */
void sxo_ylm(int l, int m, double theta, double phi, double *re, double *im) {
    const int am = m < 0 ? -m : m;
    const double t = sxo_ylm_theta(l, am, cos(theta));
    double yr = t * cos(am * phi), yi = t * sin(am * phi);
    if (m < 0) {
        const double sgn = (am & 1) ? -1.0 : 1.0;
        yr = sgn * yr;
        yi = -sgn * yi;
    }
    *re = yr;
    *im = yi;
}

/* SphericalCoor3D(CartesianCoor3D) (coor3d.cpp:168-202) */
static void to_spherical(double x, double y, double z, double *r, double *phi, double *theta) {
    *r = sqrt(pow(x, 2) + pow(y, 2) + pow(z, 2));
    if (*r == 0) {
        *theta = 0;
        *phi = 0;
        return;
    }
    *theta = acos(z / *r);
    double p;
    if (x != 0.0) {
        p = atan(y / x);
        if (x < 0.0) p += SXO_PI;
        else if (y < 0.0) p += 2 * SXO_PI;
    } else if (y != 0.0) {
        p = y > 0 ? SXO_PI / 2 : 3 * SXO_PI / 2;
    } else {
        p = 0.0;
    }
    *phi = p;
}

/* MPSphereScatterDevice::scatter (multipole_scatter_device.cpp:428-498) for
   one moment over all frames: coordinates are converted to (r, phi, theta)
   in FP64 and stored as float, as the staged spherical block does. */
void sxo_scatter_mpsphere(const float *coords, size_t NA, size_t NF, const double *b, double ql, int l, int m,
                          double *at) {
    double *jl = (double *)malloc(sizeof(double) * (size_t)(l + 1));
    for (size_t fi = 0; fi < NF; ++fi) {
        const float *p_data = coords + fi * NA * 3;
        double Ar = 0.0, Ai = 0.0;
        /* 4 pi i^l */
        const double ilr = (l % 4 == 0) ? 1.0 : (l % 4 == 2 ? -1.0 : 0.0);
        const double ili = (l % 4 == 1) ? 1.0 : (l % 4 == 3 ? -1.0 : 0.0);
        for (size_t j = 0; j < NA; ++j) {
            double r, phi, theta;
            to_spherical(p_data[3 * j], p_data[3 * j + 1], p_data[3 * j + 2], &r, &phi, &theta);
            const float rf = (float)r, phif = (float)phi, thetaf = (float)theta;
            sxo_sph_bessel_all(l, ql * (double)rf, jl);
            double yr, yi;
            sxo_ylm(l, m, (double)thetaf, (double)phif, &yr, &yi);
            /* fmpiilesf * aabess * conj(Y) */
            const double fr = 4 * SXO_PI * ilr * b[j], fi2 = 4 * SXO_PI * ili * b[j];
            const double cr = yr * jl[l], ci = -yi * jl[l];
            Ar += fr * cr - fi2 * ci;
            Ai += fr * ci + fi2 * cr;
        }
        at[2 * fi] = Ar;
        at[2 * fi + 1] = Ai;
    }
    free(jl);
}

/* Bessel function of the first kind J_n(x), n >= 0, x >= 0 (the role of
   boost::math::cyl_bessel_j, multipole_scatter_device.cpp:958-976):
   Miller's downward recurrence normalised by J_0 + 2 sum J_2k = 1, or the
   upward recurrence from J_0, J_1 when n < x. */
double sxo_cyl_bessel_j(int n, double x) {
    if (x == 0.0) return n == 0 ? 1.0 : 0.0;
    if (n == 0 || n == 1 || n < x) {
        /* J0, J1 from the Miller sum below when x is small, else upward */
        if (x > 25.0 && n < x) {
            /* asymptotic-free route: upward recurrence from accurate J0, J1 */
            double j0 = jn(0, x), j1 = jn(1, x);
            if (n == 0) return j0;
            if (n == 1) return j1;
            for (int k = 1; k < n; ++k) {
                const double j2 = 2.0 * k / x * j1 - j0;
                j0 = j1;
                j1 = j2;
            }
            return j1;
        }
    }
    return jn(n, x);
}

/* CylinderCoor3D(CartesianCoor3D) (coor3d.cpp:116-140); sign() of
   coor3d.hpp:30 takes floats, so the pi offsets are float-rounded. */
static void to_cylinder(double x, double y, double z, double *r, double *phi, double *zz) {
    *r = sqrt(pow(x, 2) + pow(y, 2));
    double p;
    if (x != 0.0) {
        p = atan(y / x);
        if (x < 0.0) p = (double)(y < 0.0 ? -(float)SXO_PI : (float)SXO_PI) + p;
    } else if (y != 0.0) {
        p = (double)(y < 0.0 ? -(float)(SXO_PI / 2) : (float)(SXO_PI / 2));
    } else {
        p = 0.0;
    }
    p = p < 0 ? 2 * SXO_PI + p : p;
    *phi = p;
    *zz = z;
}

/* CartesianVectorBase(axis) (coor3d.cpp:278-296): er, ephi, ez */
void sxo_cylinder_base(const double axis[3], double *base /*[3][3]*/) {
    const double al = sqrt(pow(axis[0], 2) + pow(axis[1], 2) + pow(axis[2], 2));
    const double ez[3] = {axis[0] / al, axis[1] / al, axis[2] / al};
    double ekez[3] = {0 * ez[2] - 1 * ez[1], 1 * ez[0] - 0 * ez[2], 0 * ez[1] - 0 * ez[0]};
    double len = sqrt(pow(ekez[0], 2) + pow(ekez[1], 2) + pow(ekez[2], 2));
    if (len == 0) {
        /* ej x ez with ej = (0,1,0) */
        ekez[0] = 1 * ez[2] - 0 * ez[1];
        ekez[1] = 0 * ez[0] - 0 * ez[2];
        ekez[2] = 0 * ez[1] - 1 * ez[0];
        len = sqrt(pow(ekez[0], 2) + pow(ekez[1], 2) + pow(ekez[2], 2));
    }
    const double er[3] = {ekez[0] / len, ekez[1] / len, ekez[2] / len};
    double ezer[3] = {ez[1] * er[2] - ez[2] * er[1], ez[2] * er[0] - ez[0] * er[2], ez[0] * er[1] - ez[1] * er[0]};
    const double l2 = sqrt(pow(ezer[0], 2) + pow(ezer[1], 2) + pow(ezer[2], 2));
    for (int k = 0; k < 3; ++k) {
        base[k] = er[k];
        base[3 + k] = ezer[k] / l2;
        base[6 + k] = ez[k];
    }
}

/* MPCylinderScatterDevice::scatter (multipole_scatter_device.cpp:900-987). */
void sxo_scatter_mpcylinder(const float *coords, size_t NA, size_t NF, const double *b, const double q[3],
                            const double axis[3], int l, int m, double *at) {
    double base[9];
    sxo_cylinder_base(axis, base);
    const double qp[3] = {q[0] * base[0] + q[1] * base[1] + q[2] * base[2],
                          q[0] * base[3] + q[1] * base[4] + q[2] * base[5],
                          q[0] * base[6] + q[1] * base[7] + q[2] * base[8]};
    double qr, qphi, qz;
    to_cylinder(qp[0], qp[1], qp[2], &qr, &qphi, &qz);
    for (size_t fi = 0; fi < NF; ++fi) {
        const float *p_data = coords + fi * NA * 3;
        double Ar = 0.0, Ai = 0.0;
        for (size_t j = 0; j < NA; ++j) {
            const double x = p_data[3 * j], y = p_data[3 * j + 1], z0 = p_data[3 * j + 2];
            const double cp[3] = {x * base[0] + y * base[1] + z0 * base[2], x * base[3] + y * base[4] + z0 * base[5],
                                  x * base[6] + y * base[7] + z0 * base[8]};
            double r, phi, z;
            to_cylinder(cp[0], cp[1], cp[2], &r, &phi, &z);
            const float rf = (float)r, phif = (float)phi, zf = (float)z;
            r = rf;
            phi = phif;
            z = zf;
            const double esf = b[j];
            double parallel_sign = 1.0;
            if ((z != 0) && (qz != 0)) parallel_sign = (z * qz) / (fabs(z) * fabs(qz));
            const double ex_r = cos(parallel_sign * z * qz), ex_i = sin(parallel_sign * z * qz);
            const double p = r * qr;
            const double psiphi = phi - qphi;
            double cr = 0.0, ci = 0.0; /* complex factor multiplying expi * esf */
            if (m == 0 && l == 0) {
                cr = sxo_cyl_bessel_j(0, p);
            } else if (m == 0) {
                cr = sqrt(0.5) * 2.0 * powf(-1.0f, (float)l) * sxo_cyl_bessel_j(2 * l, p) * cos(2 * l * psiphi);
            } else if (m == 1) {
                cr = sqrt(0.5) * 2.0 * powf(-1.0f, (float)l) * sxo_cyl_bessel_j(2 * l, p) * sin(2 * l * psiphi);
            } else if (m == 2) {
                ci = sqrt(0.5) * 2.0 * powf(-1.0f, (float)(l - 1)) * sxo_cyl_bessel_j(2 * l - 1, p) *
                     cos((2 * l - 1) * psiphi);
            } else if (m == 3) {
                ci = sqrt(0.5) * 2.0 * powf(-1.0f, (float)(l - 1)) * sxo_cyl_bessel_j(2 * l - 1, p) *
                     sin((2 * l - 1) * psiphi);
            }
            Ar += esf * (cr * ex_r - ci * ex_i);
            Ai += esf * (cr * ex_i + ci * ex_r);
        }
        const double norm = sqrt(2 * SXO_PI);
        at[2 * fi] = norm * Ar;
        at[2 * fi + 1] = norm * Ai;
    }
}

/* MPSphereScatterDevice::compute / MPCylinderScatterDevice::compute: the
   moments play the role of the vectors; final scale 1/(4 pi) (sphere,
   multipole_scatter_device.cpp:394) or 1/(2 pi) (cylinder, :866). */
int sxo_compute_multipole(int kind, const float *coords_frames, size_t NA, size_t NF, const double *b,
                          const double q[3], const double axis[3], const int *moments, size_t NMO, int dsp_type,
                          int dsp_method, double *fqt, double *fq, double *fq2) {
    if (NF == 0 || NMO == 0) return -3;
    double *buf = (double *)calloc(4 * NF + 2, sizeof(double));
    if (!buf) return -1;
    memset(fqt, 0, sizeof(double) * 2 * NF);
    double afr = 0, afi = 0, a2r = 0;
    const double ql = sqrt(pow(q[0], 2) + pow(q[1], 2) + pow(q[2], 2));
    for (size_t k = 0; k < NMO; ++k) {
        memset(buf, 0, sizeof(double) * 4 * NF);
        const int l = moments[2 * k], m = moments[2 * k + 1];
        if (kind == 0) sxo_scatter_mpsphere(coords_frames, NA, NF, b, ql, l, m, buf);
        else sxo_scatter_mpcylinder(coords_frames, NA, NF, b, q, axis, l, m, buf);
        int e = dsp(buf, NF, dsp_type, dsp_method);
        if (e) {
            free(buf);
            return e;
        }
        double a[2];
        sxo_store_one(buf, NF, a);
        afr += a[0];
        afi += a[1];
        a2r += a[0] * a[0] + a[1] * a[1];
        for (size_t i = 0; i < 2 * NF; ++i) fqt[i] += buf[i];
    }
    const double factor = kind == 0 ? 1.0 / (4 * SXO_PI) : 1.0 / (2 * SXO_PI);
    for (size_t i = 0; i < 2 * NF; ++i) fqt[i] *= factor;
    fq[0] = afr * factor;
    fq[1] = afi * factor;
    fq2[0] = a2r * factor;
    fq2[1] = 0.0;
    free(buf);
    return 0;
}
