/*
 * gexs_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; see gexs_oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math), so the
 * window-index arithmetic round(((nu-cut)-start)/dnu) is plain IEEE FP64 as in
 * Managers/CPU/OpenMPManager.cpp:399-400.
 *
 * How each function is pinned (tests/test_oracle*.py):
 *   gxo_partition      two-level closed form; mpmath sum on synthetic states;
 *                      bit-for-bit against compute_partition of the reference
 *                      compiled by `make -C oracle ref` (when buildable).
 *   gxo_select/window  brute force on tiny grids; exact k+0.5 round-half cases;
 *                      clamp at lo/hi; bit-for-bit against the reference build.
 *   gxo_humlic         scipy.special.wofz within the W4 error bound; special
 *                      cases K(x,0), K(0,y) = erfcx(y); bit-for-bit against
 *                      humlic() of compute/CPU/profiles.cpp built here.
 *   gxo_voigt916/gh11  scipy.special.wofz (1e-9 / GH error bound); GH nodes
 *                      and weights against numpy hermgauss.
 *   gxo_cross_section  Voigt: bit-for-bit (single thread) against
 *                      OpenMPManager::ComputeVoigtTP built here; integral
 *                      limits; symmetry.  Doppler: telescoping erf identity
 *                      (band strength = sum S to 1e-12) and against
 *                      OpenMPManager::ComputeDoppler's private grid times 0.5/dnu.
 *                      Gaussian / Voigt-Quad: closed forms and wofz.
 */
#include "gexs_oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* physical constants, common/defines.h:4-27 (values are part of the contract) */
#define GX_LN2 0.69314718056
#define GX_SQRTLN2 0.832554611
#define GX_ISQRTPI 0.564189584
#define GX_NA 6.0221412927e23
#define GX_CM_COEF 1.327209365e-12
#define GX_PI 3.14159265359
#define GX_PLANCK 6.6260693e-27
#define GX_VELLGT 2.99792458e+10
#define GX_BOLTZ 1.380658e-16

/* ------------------------------------------------------------------ Q(T) */
double gxo_partition(const double* E, const int32_t* g, int64_t n, double T) {
    /* exomol_functions.cpp:55-67 keeps its own copies of h, c, k; they equal
     * the defines.h values, so beta is the same expression. */
    const double planck = 6.6260693e-27, vellgt = 2.99792458e10, boltz = 1.380658e-16;
    double beta = planck * vellgt / (boltz * T);
    double q = 0.0;
    for (int64_t i = 0; i < n; ++i) q += g[i] * exp(-beta * E[i]);
    return q;
}

/* ------------------------------------------------------ selection, windows */
int gxo_select(double nu, double start, double end, double hw) {
    if (nu < start - hw) return 0;
    if (nu > end + hw) return 0; /* the reference breaks here (assumes sorted input) */
    if (nu == 0.0) return 0;
    return 1;
}

void gxo_window(double nu, double cut, double start, double dnu, int32_t lo, int32_t hi,
                int32_t* ib, int32_t* ie) {
    double b = round((nu - cut - start) / dnu);
    double e = round((nu + cut - start) / dnu);
    double blo = (double)lo, ehi = (double)hi;
    *ib = (int32_t)(b > blo ? b : blo);
    *ie = (int32_t)(e < ehi ? e : ehi);
}

void gxo_windows(int64_t n, const double* nu, double cut, double start, double dnu,
                 int32_t lo, int32_t hi, int32_t* ib, int32_t* ie) {
    for (int64_t i = 0; i < n; ++i) {
        double v = nu[i] == 0.0 ? 1e-6 : nu[i];
        gxo_window(v, cut, start, dnu, lo, hi, &ib[i], &ie[i]);
    }
}

/* ------------------------------------------------------- Voigt functions */
double gxo_humlic(double x, double y) {
    /* Humlicek (1982) JQSRT 27:437, the four-region W4 form of
     * compute/CPU/profiles.cpp:11-44, evaluated with C99 complex arithmetic in
     * the same expression order.  t = y - i x. */
    double complex t = CMPLX(y, -x);
    double s = fabs(x) + y;
    if (s >= 15) {
        double complex r = t * 0.5641896 / (0.5 + t * t);
        return creal(r);
    }
    if (s >= 5.5) {
        double complex u = t * t;
        double complex r = t * (1.410474 + u * .5641896) / (.75 + u * (3. + u));
        return creal(r);
    }
    if (y >= 0.195 * fabs(x) - 0.176) {
        double complex r = (16.4955 + t * (20.20933 + t * (11.96482 + t * (3.778987 + t * .5642236)))) /
                           (16.4955 + t * (38.82363 + t * (39.27121 + t * (21.69274 + t * (6.699398 + t)))));
        return creal(r);
    }
    double complex u = t * t;
    double complex r =
        cexp(u) - t * (36183.31 - u * (3321.9905 - u * (1540.787 - u * (219.0313 - u * (35.76683 - u * (1.320522 - u * .56419)))))) /
                      (32066.6 - u * (24322.84 - u * (9022.228 - u * (2186.181 - u * (364.2191 - u * (61.57037 - u * (1.841439 - u)))))));
    return creal(r);
}

double gxo_erfcx(double y) {
    if (y < 10.0) return exp(y * y) * erfc(y);
    /* Laplace continued fraction erfcx(y) = (1/sqrt(pi)) / (y + (1/2)/(y + 1/(y + (3/2)/(y + ...)))) */
    double f = y;
    for (int k = 60; k >= 1; --k) f = y + (0.5 * k) / f;
    return 0.56418958354775628 / f;
}

double gxo_voigt916(double x, double y) {
    /* Zaghloul & Ali (2011) ACM TOMS 38(2) Alg. 916, the series for Re w(z):
     * K = e^{-x^2} erfcx(y) cos(2xy) + (2a/pi) sin^2(xy) e^{-x^2} / y
     *     + (2a/pi) y [ -cos(2xy) s1 + (s2 + s3)/2 ],
     * s1 = sum e^{-(a^2 n^2 + x^2)}/(a^2 n^2 + y^2), s2 = sum e^{-(an+x)^2}/(..),
     * s3 = sum e^{-(an-x)^2}/(..), n >= 1.  a = 1/2 bounds the truncation error
     * near e^{-pi^2/a^2}.  (The reference's dead copy, cross_kernal.cu:122-148,
     * uses a = 1 with 4 terms.) */
    const double a = 0.5;
    x = fabs(x);
    double ex2 = exp(-x * x);
    if (x == 0.0) return gxo_erfcx(y);
    if (y == 0.0) return ex2;
    double yy = y * y;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int nmax = (int)ceil(x / a) + 16;
    for (int n = 1; n <= nmax; ++n) {
        double an = a * n;
        double f = 1.0 / (an * an + yy);
        s1 += f * exp(-(an * an + x * x));
        s2 += f * exp(-(an + x) * (an + x));
        s3 += f * exp(-(an - x) * (an - x));
    }
    double xy = x * y;
    double c2 = cos(2.0 * xy), sn = sin(xy);
    double k = ex2 * gxo_erfcx(y) * c2;
    k += (2.0 * a / M_PI) * sn * sn * ex2 / y;
    k += (2.0 * a / M_PI) * y * (-c2 * s1 + 0.5 * (s2 + s3));
    return k;
}

/* 11-point Gauss-Hermite rule (A&S 25.4.46, table 25.10): squared positive
 * nodes and weights.  Centre node t = 0 has weight GH11_W0. */
static const double GH11_T2[5] = {0.4313988071478515, 1.7597536984236963, 4.104465362828315,
                                  7.746703779542555, 13.457678352057581};
static const double GH11_W[5] = {0.42935975235612495, 0.11722787516770851, 0.011911395444911507,
                                 0.00034681946632334544, 1.4395603937142596e-06};
static const double GH11_W0 = 0.6547592869145916;

double gxo_gh11(double x, double y) {
    /* K(x,y) = (y/pi) int e^{-t^2} / ((x-t)^2 + y^2) dt ~ (y/pi) sum_k w_k/((x-t_k)^2+y^2);
     * the +-t_k pair sums to 2 (r + t^2) / ((r + t^2)^2 - 4 x^2 t^2), r = x^2+y^2. */
    double r = x * x + y * y;
    double acc = GH11_W0 / r;
    for (int k = 0; k < 5; ++k) {
        double q = r + GH11_T2[k];
        acc += GH11_W[k] * 2.0 * q / (q * q - 4.0 * x * x * GH11_T2[k]);
    }
    return acc * y / M_PI;
}

double gxo_voigt_quad(double x, double y) {
    /* selection rule of cross_kernal.cu:1144-1150 */
    return (x * x + y * y < 100.0) ? gxo_voigt916(x, y) : gxo_gh11(x, y);
}

/* ------------------------------------------------------- per-line values */
static double dpw_coef(const gxo_params* p) {
    return sqrt(2.0 * GX_LN2 * GX_BOLTZ * p->T * GX_NA / (p->mass)) / GX_VELLGT;
}

double gxo_strength(const gxo_params* p, double nu, double E, double A, int32_t g) {
    double cmcoef = 1.0 / (8.0 * GX_PI * GX_VELLGT);
    double beta = GX_PLANCK * GX_VELLGT / (GX_BOLTZ * p->T);
    double v = nu == 0.0 ? 1e-6 : nu;
    if (p->profile == GXO_VOIGT || p->profile == GXO_VOIGT_QUAD)
        return cmcoef * A * g * GX_ISQRTPI * exp(-beta * E) * (1.0 - exp(-beta * v)) / (v * v * p->Q);
    return cmcoef * A * g * exp(-beta * E) * (1.0 - exp(-beta * v)) / (v * v * p->Q);
}

double gxo_gamma_l(const gxo_params* p, double gamma, double nexp) {
    double gl = gamma * pow(p->t_ref / p->T, nexp);
    if (p->pressure_mode) gl = gl * (p->P * p->p_scale);
    return gl;
}

double gxo_inv_width(const gxo_params* p, double nu) {
    if (p->profile == GXO_GAUSSIAN) return GX_SQRTLN2 / p->hwhm;
    double v = nu == 0.0 ? 1e-6 : nu;
    if (p->width_conv) return GX_SQRTLN2 / (v * dpw_coef(p));
    return 1.0 / (v * dpw_coef(p));
}

void gxo_line_params(const gxo_params* p, int64_t n, const double* nu, const double* E,
                     const double* A, const int32_t* g, const double* gamma,
                     const double* nexp, double* S, double* gL, double* inv_w) {
    for (int64_t i = 0; i < n; ++i) {
        S[i] = gxo_strength(p, nu[i], E[i], A[i], g[i]);
        gL[i] = (gamma && nexp) ? gxo_gamma_l(p, gamma[i], nexp[i]) : 0.0;
        inv_w[i] = gxo_inv_width(p, nu[i]);
    }
}

/* ------------------------------------------------------ the line sum */
typedef struct {
    const gxo_params* p;
    double start, dnu;
    int32_t lo, hi;
    int64_t l0, l1;
    const double *nu, *E, *A, *gamma, *nexp;
    const int32_t* g;
    double* grid; /* indexed by absolute point, offset by lo */
    int64_t neval;
} gxo_job;

/* contribution of one line to point j (S, widths precomputed by caller) */
static inline double line_term(const gxo_params* p, double v, double S, double gG, double y,
                               double fj, double dnu) {
    if (p->profile == GXO_VOIGT || p->profile == GXO_VOIGT_QUAD) {
        double dfreq = v - fj;
        double x = fabs(dfreq) * gG;
        double k = p->profile == GXO_VOIGT ? gxo_humlic(x, y) : gxo_voigt_quad(x, y);
        return S * k * gG;
    }
    /* bin-averaged Gaussian: ComputeDoppler's |dnu| form (OpenMPManager.cpp:307-313) */
    double d = fabs(v - fj);
    double x0 = dnu * 0.5;
    double xp = gG * (d + x0), xm = gG * (d - x0);
    double de = erf(xp) - erf(xm);
    return S * de;
}

static void* run_job(void* arg) {
    gxo_job* jb = (gxo_job*)arg;
    const gxo_params* p = jb->p;
    int voigt = (p->profile == GXO_VOIGT || p->profile == GXO_VOIGT_QUAD);
    int64_t neval = 0;
    for (int64_t i = jb->l0; i < jb->l1; ++i) {
        double v = jb->nu[i] == 0.0 ? 1e-6 : jb->nu[i];
        double S = gxo_strength(p, jb->nu[i], jb->E[i], jb->A[i], jb->g[i]);
        int32_t ib, ie;
        gxo_window(v, p->cutoff, jb->start, jb->dnu, jb->lo, jb->hi, &ib, &ie);
        if (S == 0.0) continue;
        if (ie > ib) neval += ie - ib;
        double gG = gxo_inv_width(p, jb->nu[i]);
        double y = 0.0;
        if (voigt) y = gxo_gamma_l(p, jb->gamma[i], jb->nexp[i]) * gG;
        for (int32_t j = ib; j < ie; ++j) {
            double fj = jb->start + (double)j * jb->dnu;
            jb->grid[j - jb->lo] += line_term(p, v, S, gG, y, fj, jb->dnu);
        }
    }
    jb->neval = neval;
    return NULL;
}

int64_t gxo_cross_section(const gxo_params* p, double start, double dnu, int32_t start_idx,
                          int32_t nwin, int64_t nlines, const double* nu, const double* E,
                          const double* A, const int32_t* g, const double* gamma,
                          const double* nexp, double* out, int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nwin <= 0) return 0;
    int64_t block = (nlines + nthreads - 1) / nthreads;
    gxo_job* jobs = (gxo_job*)calloc((size_t)nthreads, sizeof(gxo_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; ++t) {
        gxo_job* jb = &jobs[t];
        jb->p = p; jb->start = start; jb->dnu = dnu;
        jb->lo = start_idx; jb->hi = start_idx + nwin;
        jb->l0 = (int64_t)t * block;
        jb->l1 = jb->l0 + block < nlines ? jb->l0 + block : nlines;
        if (jb->l0 > nlines) jb->l0 = nlines;
        jb->nu = nu; jb->E = E; jb->A = A; jb->g = g; jb->gamma = gamma; jb->nexp = nexp;
        jb->grid = (double*)calloc((size_t)nwin, sizeof(double));
    }
    if (nthreads == 1) {
        run_job(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, run_job, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    /* ordered merge (the reference merges under a mutex in arbitrary order,
     * OpenMPManager.cpp:415-420; a fixed order makes the oracle repeatable) */
    int voigt = (p->profile == GXO_VOIGT || p->profile == GXO_VOIGT_QUAD);
    double norm = 0.5 / dnu; /* bin-averaged profiles, cross_kernal_v2.cu:399 */
    int64_t neval = 0;
    for (int t = 0; t < nthreads; ++t) {
        for (int32_t j = 0; j < nwin; ++j)
            out[start_idx + j] += voigt ? jobs[t].grid[j] : jobs[t].grid[j] * norm;
        neval += jobs[t].neval;
        free(jobs[t].grid);
    }
    free(jobs);
    free(th);
    return neval;
}

static int64_t lower_bound_d(const double* a, int64_t n, double v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t m = lo + (hi - lo) / 2;
        if (a[m] < v) lo = m + 1; else hi = m;
    }
    return lo;
}

typedef struct {
    const gxo_params* p;
    double start, dnu;
    int32_t lo, hi;
    int64_t nlines;
    const double *nu, *E, *A, *gamma, *nexp;
    const int32_t* g;
    const int32_t* pts;
    double* out;
    int64_t k0, k1;
} gxo_at_job;

static void* run_at(void* arg) {
    gxo_at_job* jb = (gxo_at_job*)arg;
    const gxo_params* p = jb->p;
    int voigt = (p->profile == GXO_VOIGT || p->profile == GXO_VOIGT_QUAD);
    for (int64_t k = jb->k0; k < jb->k1; ++k) {
        int32_t j = jb->pts[k];
        double fj = jb->start + (double)j * jb->dnu;
        /* every line whose window holds j has |nu - f_j| <= cut + dnu */
        double span = p->cutoff + 2.0 * jb->dnu;
        int64_t i0 = lower_bound_d(jb->nu, jb->nlines, fj - span);
        double acc = 0.0;
        for (int64_t i = i0; i < jb->nlines && jb->nu[i] <= fj + span; ++i) {
            double v = jb->nu[i] == 0.0 ? 1e-6 : jb->nu[i];
            int32_t ib, ie;
            gxo_window(v, p->cutoff, jb->start, jb->dnu, jb->lo, jb->hi, &ib, &ie);
            if (j < ib || j >= ie) continue;
            double S = gxo_strength(p, jb->nu[i], jb->E[i], jb->A[i], jb->g[i]);
            if (S == 0.0) continue;
            double gG = gxo_inv_width(p, jb->nu[i]);
            double y = voigt ? gxo_gamma_l(p, jb->gamma[i], jb->nexp[i]) * gG : 0.0;
            acc += line_term(p, v, S, gG, y, fj, jb->dnu);
        }
        jb->out[k] = voigt ? acc : acc * (0.5 / jb->dnu);
    }
    return NULL;
}

void gxo_cross_section_at(const gxo_params* p, double start, double dnu, int32_t start_idx,
                          int32_t nwin, int64_t nlines, const double* nu, const double* E,
                          const double* A, const int32_t* g, const double* gamma,
                          const double* nexp, int64_t npts, const int32_t* pts, double* out,
                          int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    gxo_at_job* jobs = (gxo_at_job*)calloc((size_t)nthreads, sizeof(gxo_at_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    int64_t block = (npts + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        gxo_at_job* jb = &jobs[t];
        jb->p = p; jb->start = start; jb->dnu = dnu;
        jb->lo = start_idx; jb->hi = start_idx + nwin; jb->nlines = nlines;
        jb->nu = nu; jb->E = E; jb->A = A; jb->g = g; jb->gamma = gamma; jb->nexp = nexp;
        jb->pts = pts; jb->out = out;
        jb->k0 = (int64_t)t * block < npts ? (int64_t)t * block : npts;
        jb->k1 = jb->k0 + block < npts ? jb->k0 + block : npts;
    }
    if (nthreads == 1) {
        run_at(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, run_at, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    free(jobs);
    free(th);
}

int64_t gxo_neval(const gxo_params* p, double start, double dnu, int32_t start_idx,
                  int32_t nwin, int64_t nlines, const double* nu, const double* E,
                  const double* A, const int32_t* g) {
    int64_t n = 0;
    for (int64_t i = 0; i < nlines; ++i) {
        double v = nu[i] == 0.0 ? 1e-6 : nu[i];
        int32_t ib, ie;
        gxo_window(v, p->cutoff, start, dnu, start_idx, start_idx + nwin, &ib, &ie);
        if (ie <= ib) continue;
        if (gxo_strength(p, nu[i], E[i], A[i], g[i]) == 0.0) continue;
        n += ie - ib;
    }
    return n;
}

double gxo_total_intensity(double T, double Q, int64_t n, const double* nu, const double* E,
                           const double* A, const int32_t* g) {
    double beta = GX_PLANCK * GX_VELLGT / (GX_BOLTZ * T);
    double tot = 0.0;
    for (int64_t i = 0; i < n; ++i)
        tot += GX_CM_COEF * A[i] * g[i] * exp(-beta * E[i]) * (1.0 - exp(-beta * nu[i])) /
               (nu[i] * nu[i] * Q);
    return tot;
}
