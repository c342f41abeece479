// profiles_device.h -- profile functions evaluated on gfx950.
//
// Humlicek W4 (JQSRT 27:437, 1982): the four-region rational/exponential
// approximation of K(x,y) = Re w(x+iy) used by the reference in
// compute/CPU/profiles.cpp:11-44 (FP64) and compute/CUDA/cross_kernal_v2.cu:52-84
// (FP32).  Here complex arithmetic is spelled out on (re, im) pairs and only
// the real part of the last quotient is formed.
#pragma once
#include <hip/hip_runtime.h>

namespace gx {

template <typename R>
struct cplx {
    R r, i;
};

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <typename R>
__device__ __forceinline__ cplx<R> cadd(R a, cplx<R> b) {
    return {a + b.r, b.i};
}
template <typename R>
__device__ __forceinline__ cplx<R> csubr(R a, cplx<R> b) {  // a - b
    return {a - b.r, -b.i};
}
template <typename R>
__device__ __forceinline__ cplx<R> cscale(cplx<R> a, R s) {
    return {a.r * s, a.i * s};
}
// a + t*p and a - u*p as Horner steps
template <typename R>
__device__ __forceinline__ cplx<R> hadd(R a, cplx<R> t, cplx<R> p) {
    return cadd(a, cmul(t, p));
}
template <typename R>
__device__ __forceinline__ cplx<R> hsub(R a, cplx<R> u, cplx<R> p) {
    return csubr(a, cmul(u, p));
}

// Region III:  (16.4955 + t(20.20933 + t(11.96482 + t(3.778987 + 0.5642236 t))))
//            / (16.4955 + t(38.82363 + t(39.27121 + t(21.69274 + t(6.699398 + t)))))
template <typename R>
__device__ __forceinline__ void w4_region3(cplx<R> t, cplx<R>& num, cplx<R>& den) {
    cplx<R> p = cadd(R(3.778987), cscale(t, R(.5642236)));
    p = hadd(R(11.96482), t, p);
    p = hadd(R(20.20933), t, p);
    num = hadd(R(16.4955), t, p);
    cplx<R> q = cadd(R(6.699398), t);
    q = hadd(R(21.69274), t, q);
    q = hadd(R(39.27121), t, q);
    q = hadd(R(38.82363), t, q);
    den = hadd(R(16.4955), t, q);
}

// Region IV rational part: t (36183.31 - u(3321.9905 - u(1540.787 - u(219.0313
//   - u(35.76683 - u(1.320522 - 0.56419 u)))))) / (32066.6 - u(24322.84 - u(9022.228
//   - u(2186.181 - u(364.2191 - u(61.57037 - u(1.841439 - u)))))))
template <typename R>
__device__ __forceinline__ void w4_region4(cplx<R> t, cplx<R> u, cplx<R>& num, cplx<R>& den) {
    cplx<R> p = csubr(R(1.320522), cscale(u, R(.56419)));
    p = hsub(R(35.76683), u, p);
    p = hsub(R(219.0313), u, p);
    p = hsub(R(1540.787), u, p);
    p = hsub(R(3321.9905), u, p);
    p = hsub(R(36183.31), u, p);
    num = cmul(t, p);
    cplx<R> q = csubr(R(1.841439), u);
    q = hsub(R(61.57037), u, q);
    q = hsub(R(364.2191), u, q);
    q = hsub(R(2186.181), u, q);
    q = hsub(R(9022.228), u, q);
    q = hsub(R(24322.84), u, q);
    den = hsub(R(32066.6), u, q);
}

// Re(a / b)
template <typename R>
__device__ __forceinline__ R re_div(cplx<R> a, cplx<R> b) {
    return (a.r * b.r + a.i * b.i) / (b.r * b.r + b.i * b.i);
}
__device__ __forceinline__ float re_div_f(cplx<float> a, cplx<float> b) {
    return (a.r * b.r + a.i * b.i) * __builtin_amdgcn_rcpf(b.r * b.r + b.i * b.i);
}

// Region I only: c*y*(0.5+y^2+x^2) / ((0.5+y^2-x^2)^2 + 4x^2y^2)
template <typename R>
__device__ __forceinline__ R w4_region1(R x, R y) {
    R x2 = x * x, b = R(0.5) + y * y;
    R d = b - x2;
    return R(0.5641896) * y * (b + x2) / (d * d + R(4) * x2 * y * y);
}

// full W4 in FP64 (regions chosen on S = |x| + y exactly as the reference)
__device__ inline double w4_f64(double x, double y) {
    typedef cplx<double> C;
    C t = {y, -x};
    double s = fabs(x) + y;
    if (s >= 15.0) return w4_region1(fabs(x), y);
    C u = cmul(t, t);
    if (s >= 5.5) {
        C num = cmul(t, cadd(1.410474, cscale(u, .5641896)));
        C den = cadd(.75, cmul(u, cadd(3.0, u)));
        return re_div(num, den);
    }
    C num, den;
    if (y >= 0.195 * fabs(x) - 0.176) {
        w4_region3(t, num, den);
        return re_div(num, den);
    }
    w4_region4(t, u, num, den);
    // Re exp(u) = e^{u.r} cos(u.i)
    return exp(u.r) * cos(u.i) - re_div(num, den);
}

// Region II in real arithmetic.  Humlicek's region-II rational
// t(1.410474 + 0.5641896 u) / (0.75 + u(3 + u)), t = y - ix, u = t^2, is the
// 4-point Gauss-Hermite quadrature of w(z), (i/pi) sum_k w_k / (z - a_k),
// with nodes +-a_k the roots of H4 (a_k^2 = (3 -+ sqrt 6)/2) and weights w_k
// (1.410474 = 2.5/sqrt(pi), 0.5641896 = 1/sqrt(pi) to the published digits,
// so the two agree within 3e-8 relative).  Its real part, the nodes taken in
// +-a pairs, is
//   K = sum_k (2 w_k / pi) y B_k / (B_k^2 - 4 a_k^2 x^2),  B_k = x^2 + y^2 + a_k^2,
// where B_k^2 - 4 a_k^2 x^2 = ((x-a_k)^2 + y^2)((x+a_k)^2 + y^2) > 0 for
// x + y >= 5.5.  Both pairs share one reciprocal: ~15 VALU against ~21 for the
// complex quotient, and 4e-7 against 1.5e-6 maximum relative error in FP32
// over the region (both against the FP64 complex form).
__device__ __forceinline__ float w4_region2_f32(float x, float y) {
    const float a1s = 0.27525512860841095f;   // (3 - sqrt 6) / 2
    const float a2s = 2.7247448713915890f;    // (3 + sqrt 6) / 2
    const float c1 = 0.5124242247547685f;    // 2 w_1 / pi, w_1 = 0.80491409000551284
    const float c2 = 0.05176535879298783f;    // 2 w_2 / pi, w_2 = 0.08131283544724518
    float x2 = x * x;
    float s = __builtin_fmaf(y, y, x2);
    float b1 = s + a1s, b2 = s + a2s;
    float d1 = __builtin_fmaf(b1, b1, -4.0f * a1s * x2);
    float d2 = __builtin_fmaf(b2, b2, -4.0f * a2s * x2);
    float n = __builtin_fmaf(c1 * b1, d2, c2 * b2 * d1);
    return y * n * __builtin_amdgcn_rcpf(d1 * d2);
}

// W4 in FP32 for x >= 0 and x + y < 15 (regions II-IV; region I is
// evaluated inline by the profile kernel)
__device__ inline float w4_f32_inner(float x, float y) {
    typedef cplx<float> C;
    float s = x + y;
    if (s >= 5.5f) return w4_region2_f32(x, y);
    C t = {y, -x};
    C num, den;
    if (y >= 0.195f * x - 0.176f) {
        w4_region3(t, num, den);
        return re_div_f(num, den);
    }
    C u = cmul(t, t);
    w4_region4(t, u, num, den);
    return expf(u.r) * cosf(u.i) - re_div_f(num, den);
}

// Zaghloul & Ali (2011) ACM TOMS 38(2) Algorithm 916, series for Re w, with
// a = 1/2: the n-sums use e^{-(an)^2} = e^{-n^2/4} and e^{-(an +- x)^2} =
// e^{-x^2} e^{-n^2/4} e^{-+ n x}, so only three exponentials per point.
__device__ inline double voigt916_f64(double x, double y) {
    x = fabs(x);
    double ex2 = exp(-x * x);
    if (x == 0.0) return erfcx(y);
    if (y == 0.0) return ex2;
    double yy = y * y;
    double q = exp(-x), qi = exp(x);  // e^{-x}, e^{x}
    double qn = 1.0, qin = 1.0;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int nmax = (int)ceil(2.0 * x) + 16;
    for (int n = 1; n <= nmax; ++n) {
        double an = 0.5 * n;
        double cn = exp(-an * an);
        double f = cn / (an * an + yy);
        qn *= q;
        qin *= qi;
        s1 += f;
        s2 += f * qn;
        s3 += f * qin;
    }
    s1 *= ex2;
    s2 *= ex2;
    s3 *= ex2;
    double xy = x * y;
    double c2 = cos(2.0 * xy), sn = sin(xy);
    const double a2pi = 1.0 / 3.14159265358979323846;  // 2a/pi at a = 1/2
    double k = ex2 * erfcx(y) * c2;
    k += a2pi * sn * sn * ex2 / y;
    k += a2pi * y * (-c2 * s1 + 0.5 * (s2 + s3));
    return k;
}

// The same series in FP32 for x^2 + y^2 < 100.  The factors e^{-x^2} e^{-+nx}
// of the FP64 form under- and overflow in FP32 (x up to 10, n up to 2x + 16),
// so each term is taken whole: e^{-(an +- x)^2} by one native exponential,
// and e^{-(an)^2} from a table.  Terms of the first two sums beyond n = 10 are
// below e^{-25} of the leading ones, and the third sum peaks at n ~ 2x and is
// cut at 2x + 10.  sin^2(xy) and cos(2xy) = (c - s)(c + s) come from one
// sincos of xy (1 - cos(2xy) would cancel for small xy).  Against wofz the
// result is within 2e-6 relative over the region (4e-7 of K(0, y)).
__device__ inline float voigt916_f32(float x, float y) {
    // e^{-n^2/4}, n = 1..10
    const float g[10] = {0.7788007830714049f,   0.36787944117144233f,  0.10539922456186433f,
                         0.01831563888873418f,  0.0019304541362277093f, 1.2340980408667956e-4f,
                         4.785117392129009e-6f, 1.1253517471925912e-7f, 1.6052280551856116e-9f,
                         1.3887943864964021e-11f};
    x = fabsf(x);
    if (y == 0.0f) return __expf(-x * x);
    if (x == 0.0f) return erfcxf(y);
    const float yy = y * y;
    float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
#pragma unroll
    for (int n = 1; n <= 10; ++n) {
        const float an = 0.5f * (float)n;
        const float u = __builtin_amdgcn_rcpf(an * an + yy);
        const float p = an + x, m = an - x;
        s1 = fmaf(g[n - 1], u, s1);
        s2 = fmaf(__expf(-p * p), u, s2);
        s3 = fmaf(__expf(-m * m), u, s3);
    }
    const int nmax = (int)ceilf(2.0f * x) + 10;
    for (int n = 11; n <= nmax; ++n) {
        const float an = 0.5f * (float)n;
        const float u = __builtin_amdgcn_rcpf(an * an + yy);
        const float m = an - x;
        s3 = fmaf(__expf(-m * m), u, s3);
    }
    const float ex2 = __expf(-x * x);
    float sn, cs;
    sincosf(x * y, &sn, &cs);
    const float c2 = (cs - sn) * (cs + sn);
    const float a2pi = 0.3183098861837907f;  // 2a/pi at a = 1/2
    float k = ex2 * erfcxf(y) * c2;
    k += a2pi * sn * sn * ex2 / y;
    k += a2pi * y * fmaf(-c2 * ex2, s1, 0.5f * (s2 + s3));
    return k;
}

// Voigt-Quad in FP32 where x^2 + y^2 < kQuadR2: the Zaghloul-Ali series
// inside x^2 + y^2 < 100 (the reference's core rule), the 4-point
// Gauss-Hermite pairs of w4_region2_f32 beyond.  The reference's far rule is
// the 11-point Gauss-Hermite sum; for x^2 + y^2 >= 100 both quadratures equal
// K to within 1e-14 (11 points) and 1.5e-7 (4 points) relative, so the
// 4-point sum stands in for the 11-point one with two reciprocals instead of
// six.  Beyond kQuadR2 the profile kernel evaluates the one-term continued
// fraction (Humlicek's region-I rational, with 1/sqrt(pi)), which is within
// 1e-6 relative of K there (2.5 / (x^2 + y^2)^2).
constexpr float kQuadR2 = 1600.0f;

__device__ __forceinline__ float quad_inner(float x, float y) {
    const float r = fmaf(x, x, y * y);
    return r < 100.0f ? voigt916_f32(x, y) : w4_region2_f32(x, y);
}

// 11-point Gauss-Hermite quadrature of K (A&S 25.4.46, table 25.10) in FP64,
// the reference's far rule, for the plain FP64 path: the +-t_k pair sums to
// 2 w_k (r + t_k^2) / ((r + t_k^2)^2 - 4 x^2 t_k^2), r = x^2 + y^2.
__device__ __forceinline__ double gh11_f64(double x, double y) {
    const double t2[5] = {0.4313988071478515, 1.7597536984236963, 4.104465362828315,
                          7.746703779542555, 13.457678352057581};
    const double w[5] = {0.42935975235612495, 0.11722787516770851, 0.011911395444911507,
                         0.00034681946632334544, 1.4395603937142596e-06};
    double r = x * x + y * y;
    double acc = 0.6547592869145916 / r;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double qq = r + t2[k];
        acc += w[k] * 2.0 * qq / (qq * qq - 4.0 * x * x * t2[k]);
    }
    return acc * y / 3.14159265358979323846;
}

// erfc of the bin-averaged Gaussian as kErfcPieces cubic pieces over
// [-8, 8), kErfcPerUnit per unit: piece k, e = k / kErfcPerUnit - 8 + f /
// kErfcPerUnit with f in [0, 1), is the cubic Hermite interpolant of erfc
// and its derivative -2/sqrt(pi) exp(-e^2) at the piece's ends, kept as FP32
// coefficients (a0, a1, a2, a3) of f (gxk::init_gauss_tables computes them
// in FP64 on the device).  Interpolation error: the largest fourth
// derivative of erfc (5.8) / 384 / kErfcPerUnit^4 = 1.4e-8; the FP32
// evaluation adds ~2 ulp of erfc.
constexpr int kErfcPieces = 512;
constexpr float kErfcPerUnit = 32.0f;

__device__ __forceinline__ double4 erfc_piece(int k) {
    const double e0 = (double)k / (double)kErfcPerUnit - 8.0;
    const double e1 = (double)(k + 1) / (double)kErfcPerUnit - 8.0;
    const double H0 = erfc(e0), H1 = erfc(e1);
    const double s = -1.1283791670955126 / (double)kErfcPerUnit;  // -2/sqrt(pi) per piece
    const double D0 = s * exp(-e0 * e0), D1 = s * exp(-e1 * e1);
    return make_double4(H0, D0, 3.0 * (H1 - H0) - 2.0 * D0 - D1, 2.0 * (H0 - H1) + D0 + D1);
}

}  // namespace gx
