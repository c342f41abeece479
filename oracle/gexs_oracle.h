/*
 * gexs_oracle.h -- FP64 CPU oracle for the line-by-line cross-section path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gexs_amd/, include/)
 * links, loads or calls this library.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * It restates, in FP64 and in the reference's operation order, the CPU path
 * of GEXS (citations are to files of the reference tree):
 *   partition function     Readers/Exomol/exomol_functions.cpp:50-73
 *   line selection         Profiles/BaseProfile.cpp:152-157 (filter on both sides)
 *   grid                   Profiles/BaseProfile.cpp:10,19
 *   Voigt line assembly    Managers/CPU/OpenMPManager.cpp:376-426 (ComputeVoigtTP)
 *   Humlicek W4            compute/CPU/profiles.cpp:11-44
 *   Doppler                Managers/CPU/OpenMPManager.cpp:280-323 plus the
 *                          0.5/dnu normalisation of compute/CUDA/cross_kernal_v2.cu:399
 *   fixed-HWHM Gaussian    compute/CUDA/cross_kernal.cu:647-716 (semantics)
 *   Voigt-Quad             compute/CUDA/cross_kernal.cu:122-217,1144-1150 (semantics,
 *                          with textbook Gauss-Hermite nodes/weights and the
 *                          Zaghloul-Ali series at a = 1/2)
 *   band strength          Profiles/BaseProfile.cpp:250-267,283-291
 *
 * Pinning status of every function is listed in the header of gexs_oracle.c.
 */
#ifndef GEXS_ORACLE_H
#define GEXS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GXO_VOIGT = 0, GXO_DOPPLER = 1, GXO_GAUSSIAN = 2, GXO_VOIGT_QUAD = 3 };

typedef struct gxo_params {
    int32_t profile;        /* GXO_* */
    int32_t width_conv;     /* 0: 1/w = 1/(nu*dpwcoef) (reference CPU); 1: sqrt(ln2)/(nu*dpwcoef) */
    int32_t pressure_mode;  /* 0: gammaL = gamma*(Tref/T)^n (reference CPU); 1: ... * (P*p_scale) */
    int32_t reserved;
    double T, P, Q;         /* temperature K, pressure, partition function */
    double mass;            /* mean molecular mass, amu */
    double t_ref, p_scale;  /* reference temperature, pressure unit scale */
    double cutoff;          /* window half-width in cm^-1 */
    double hwhm;            /* fixed HWHM of the GAUSSIAN profile, cm^-1 */
} gxo_params;

/* Q(T) = sum g_i exp(-beta E_i), sequential, exomol_functions.cpp:60-67 */
double gxo_partition(const double* E, const int32_t* g, int64_t n, double T);

/* 1 iff the line is kept: !(nu < start-hw) && !(nu > end+hw) && nu != 0 */
int gxo_select(double nu, double start, double end, double hw);

/* per-line window, OpenMPManager.cpp:399-400, clamped to [lo, hi] */
void gxo_window(double nu, double cut, double start, double dnu, int32_t lo, int32_t hi,
                int32_t* ib, int32_t* ie);
void gxo_windows(int64_t n, const double* nu, double cut, double start, double dnu,
                 int32_t lo, int32_t hi, int32_t* ib, int32_t* ie);

/* Voigt functions K(x,y) = Re w(x+iy) */
double gxo_humlic(double x, double y);      /* Humlicek W4, profiles.cpp:11-44 */
double gxo_voigt916(double x, double y);    /* Zaghloul & Ali (2011) series, a = 1/2 */
double gxo_gh11(double x, double y);        /* 11-point Gauss-Hermite quadrature */
double gxo_voigt_quad(double x, double y);  /* 916 if x^2+y^2 < 100 else GH-11 */
double gxo_erfcx(double y);

/* per-line quantities (Voigt S includes 1/sqrt(pi) as in OpenMPManager.cpp:394) */
double gxo_strength(const gxo_params* p, double nu, double E, double A, int32_t g);
double gxo_gamma_l(const gxo_params* p, double gamma, double nexp);
double gxo_inv_width(const gxo_params* p, double nu);
void gxo_line_params(const gxo_params* p, int64_t n, const double* nu, const double* E,
                     const double* A, const int32_t* g, const double* gamma,
                     const double* nexp, double* S, double* gL, double* inv_w);

/* Full cross-section: accumulates into out[start_idx .. start_idx+nwin).
 * Grid f[j] = start + j*dnu.  Lines with S == 0 are skipped.  nthreads
 * workers take contiguous line blocks into private grids merged in worker
 * order (deterministic).  Returns N_eval = sum over lines of max(0, ie-ib). */
int64_t gxo_cross_section(const gxo_params* p, double start, double dnu, int32_t start_idx,
                          int32_t nwin, int64_t nlines, const double* nu, const double* E,
                          const double* A, const int32_t* g, const double* gamma,
                          const double* nexp, double* out, int32_t nthreads);

/* Cross-section at selected grid points only; lines must be sorted by nu
 * (ascending).  For sorted input the per-point summation order equals the
 * single-thread full oracle, so values agree bit for bit. */
void gxo_cross_section_at(const gxo_params* p, double start, double dnu, int32_t start_idx,
                          int32_t nwin, int64_t nlines, const double* nu, const double* E,
                          const double* A, const int32_t* g, const double* gamma,
                          const double* nexp, int64_t npts, const int32_t* pts, double* out,
                          int32_t nthreads);

/* N_eval only (no profile evaluation). */
int64_t gxo_neval(const gxo_params* p, double start, double dnu, int32_t start_idx,
                  int32_t nwin, int64_t nlines, const double* nu, const double* E,
                  const double* A, const int32_t* g);

/* Sum over lines of CM_COEF*A*g*exp(-beta E)*(1-exp(-beta nu))/(nu^2 Q)
 * (BaseProfile.cpp:258-260). */
double gxo_total_intensity(double T, double Q, int64_t n, const double* nu, const double* E,
                           const double* A, const int32_t* g);

#ifdef __cplusplus
}
#endif
#endif
