// kernels.hip -- gfx950 kernels of the line-by-line cross-section path.
//
// Per (T,P) the path is
//   k_line_rec      FP64 S, gammaL, 1/w, window -> 32-B FP32 record per line
//                   (amplitude as mantissa and exponent), N_eval, the largest
//                   exponent (the FP32 sums run in units of 2^E)
//   k_segments      per tile, the lines centred near it (Humlicek regions II-IV)
//   k_profile_tiles a tile of 64*PPL grid points per wave or per workgroup (its
//                   waves share the tile's nu-sorted lines); lines are staged per
//                   64-line batch into LDS rows, the profile is evaluated in FP32
//                   with per-point Kahan sums in registers; each point is written
//                   once (no atomics on the grid)
//   k_group_sum     tiles cut into line groups: group partial rows added in order
// Grid-dependent, T,P-independent steps (once per chunk and grid):
//   k_line_pos      window [ib, ie) and nearest point c per line (FP64, bit-exact
//                   to OpenMPManager.cpp:399-400)
//   k_tiles         binary-searched line range per tile -> work items
// Baseline:
//   k_fp64_points   one work-item per grid point, binary-searched line range,
//                   FP64 W4 and FP64 sum (the "plain FP64" GPU path)
// Partition function:
//   k_partition / k_partition_final  FP64 sum of g exp(-beta E) in fixed order
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "gexs_internal.h"
#include "kernels.h"
#include "profiles_device.h"

namespace {

// ----------------------------------------------------------- FP64 physics
// Operation order follows Managers/CPU/OpenMPManager.cpp:385-404.
__device__ __forceinline__ double dpw_coef(const TPParams& p) {
    return sqrt(2.0 * GX_LN2 * GX_BOLTZ * p.T * GX_NA / (p.mass)) / GX_VELLGT;
}

__device__ __forceinline__ double line_strength(const TPParams& p, double nu, double E, double A,
                                                int g) {
#pragma clang fp contract(off)
    double cmcoef = 1.0 / (8.0 * GX_PI * GX_VELLGT);
    double beta = GX_PLANCK * GX_VELLGT / (GX_BOLTZ * p.T);
    double v = nu == 0.0 ? 1e-6 : nu;
    double s = cmcoef * A * g;
    if (p.profile == GX_PROF_VOIGT || p.profile == GX_PROF_VOIGT_QUAD) s = s * GX_ISQRTPI;
    return s * exp(-beta * E) * (1.0 - exp(-beta * v)) / (v * v * p.Q);
}

__device__ __forceinline__ double inv_width(const TPParams& p, double nu) {
    if (p.profile == GX_PROF_GAUSSIAN) return GX_SQRTLN2 / p.gauss_hwhm;
    double v = nu == 0.0 ? 1e-6 : nu;
    if (p.width_conv) return GX_SQRTLN2 / (v * dpw_coef(p));
    return 1.0 / (v * dpw_coef(p));
}

__device__ __forceinline__ double gamma_l(const TPParams& p, double gamma, double nexp) {
    double gl = gamma * pow(p.t_ref / p.T, nexp);
    if (p.pressure_mode) gl = gl * (p.P * p.p_scale);
    return gl;
}

// amplitude that multiplies the FP32 profile value of one line
__device__ __forceinline__ double line_amp(const TPParams& p, double S, double iw) {
    if (p.profile == GX_PROF_VOIGT || p.profile == GX_PROF_VOIGT_QUAD) return S * iw;
    return S * (0.5 / p.dfreq);
}


__device__ __forceinline__ long long wave_sum_ll(long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Workgroup reductions (256 threads): the result is valid in thread 0.  One
// atomic per workgroup instead of one per wave keeps same-address atomics
// off the critical path of the per-line passes.
__device__ __forceinline__ long long block_sum_ll(long long v, long long* sh) {
    v = wave_sum_ll(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) v = sh[0] + sh[1] + sh[2] + sh[3];
    return v;
}

__device__ __forceinline__ int block_max_i(int v, int* sh) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) v = max(max(sh[0], sh[1]), max(sh[2], sh[3]));
    return v;
}

}  // namespace

// ------------------------------------------------------------- k_line_pos
namespace {
// window [ib, ie) and nearest grid index c of a line at x
__device__ __forceinline__ LinePos window_of(const TPParams& p, double x) {
#pragma clang fp contract(off)
    const double v = x == 0.0 ? 1e-6 : x;
    // OpenMPManager.cpp:399-400: round(((nu - cut) - start)/dnu), clamped
    double b = round((v - p.cutoff - p.start) / p.dfreq);
    double e = round((v + p.cutoff - p.start) / p.dfreq);
    const double lo = (double)p.win_lo, hi = (double)p.win_hi;
    double bb = (b < lo) ? lo : b;
    double ee = (hi < e) ? hi : e;
    // keep the integers in range; far-away lines have empty windows anyway
    bb = fmin(bb, 2.0e9);
    ee = fmax(ee, -2.0e9);
    double c = round((v - p.start) / p.dfreq);
    c = fmin(fmax(c, -1.0e9), 1.0e9);
    LinePos lp;
    lp.ib = (int)bb;
    lp.ie = (int)ee;
    lp.c = (int)c;
    lp.cf = (float)c;
    return lp;
}
}  // namespace

// blockIdx.y = job; each job has its own lines, grid and window
__global__ __launch_bounds__(256) void k_line_pos(const double* __restrict__ nu,
                                                  const JobDesc* __restrict__ jd,
                                                  const TPParams* __restrict__ tp,
                                                  LinePos* __restrict__ pos) {
    const JobDesc d = jd[blockIdx.y];
    const TPParams p = tp[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n; i += gridDim.x * blockDim.x)
        pos[d.pos_off + i] = window_of(p, nu[d.l0 + i]);
}

// Amplitudes of one pass span many decades (line strengths reach 1e-125 at
// 2000 K), beyond FP32.  Each record therefore holds its amplitude as an FP32
// mantissa m in [0.5, 1) times 2^ex (ex in VRec::ex), and the pass keeps the
// largest ex of the lines that reach the grid, E; the profile kernel stages a
// line's constants as ldexp(m, ex - E), so the largest amplitude lands in
// [0.5, 1).  Power-of-two scaling commutes with rounding, so these are the
// FP32 values of amp * 2^-E (bit for bit, away from FP32 subnormals).  E is
// kept biased by kExBias in a 64-bit counter, zero before the pass (no line).
constexpr int kExBias = 4096;

__device__ __forceinline__ int pass_exponent(unsigned long long biased) {
    return biased == 0ull ? 0 : (int)biased - kExBias;
}

// ------------------------------------------------------------- k_line_rec
// FP32 record of one line at one (T,P) (FP64 arithmetic, no contraction):
// returns the amplitude exponent ex (0 when the amplitude is 0) and sets
// *neval to the line's window length when S != 0 and the window is not empty,
// and *k15 (Voigt) to the distance beyond which its points are in region I.
__device__ __forceinline__ int line_record(const TPParams& p, bool voigt, double nu, double E,
                                           double A, int g, const double* gamma,
                                           const double* nexp, int64_t i, const LinePos& lp,
                                           VRec& r, VRecX& x, long long* neval) {
#pragma clang fp contract(off)
    const double v = nu == 0.0 ? 1e-6 : nu;
    const double S = line_strength(p, nu, E, A, g);
    const double iw = inv_width(p, nu);
    const bool reaches = lp.ie > lp.ib;
    *neval = (S != 0.0 && reaches) ? (long long)(lp.ie - lp.ib) : 0;
    // offset of the line from its nearest grid point, FP64 (f_c = start + c*dfreq)
    const double fc = p.start + (double)lp.c * p.dfreq;
    const double delta = v - fc;
    const double Af = p.dfreq * iw;
    const double Bf = delta * iw;
    // a line whose window misses the grid never adds to it: amplitude 0
    double amp = reaches ? line_amp(p, S, iw) : 0.0;
    int ex = 0;
    if (isfinite(amp) && amp > 0.0) {
        frexp(amp, &ex);
        amp = ldexp(amp, -ex);  // exact: the mantissa in [0.5, 1)
    } else {
        amp = 0.0;
    }
    r.cf = lp.cf;
    r.A = (float)Af;
    r.B = (float)Bf;
    r.Kg = (float)amp;
    if (voigt) {
        const double y = (gamma != nullptr) ? gamma_l(p, gamma[i], nexp[i]) * iw : 0.0;
        const bool quad = p.profile == GX_PROF_VOIGT_QUAD;
        // Voigt: Humlicek's region-I constant; Voigt-Quad: 1/sqrt(pi)
        r.Kl = (float)(amp * (quad ? 0.5641895835477563 : 0.5641896) * y);
        r.E2 = (float)(y * y - 0.5);
        r.H = (float)(2.0 * y * y);
        x.y = (float)y;
        // |x| >= |c - j| A - |B|; region I needs |x| + y >= 15 (Voigt) or
        // x^2 + y^2 >= kQuadR2 (Voigt-Quad), with margin
        double kk;
        if (quad) {
            const double xm2 = (double)gx::kQuadR2 * (1.0 + 1e-5) - y * y;
            kk = xm2 > 0.0 ? (sqrt(xm2) + fabs(Bf)) / Af : 0.0;
        } else {
            kk = (15.0 * (1.0 + 1e-5) - y + fabs(Bf)) / Af;
        }
        x.k15 = kk <= 0.0 ? 0 : (kk >= 1.0e8 ? 100000000 : (int)ceil(kk));
    } else {
        r.Kl = 0.0f;
        r.E2 = 0.0f;
        r.H = 0.0f;
        x.y = (float)(0.5 * p.dfreq * iw);  // H, half bin width in units of the width
        x.k15 = 0;
    }
    r.ex = ex;
    return amp > 0.0 ? ex : 0;
}

__global__ __launch_bounds__(256) void k_line_rec(const double* __restrict__ nu,
                                                  const double* __restrict__ E,
                                                  const double* __restrict__ A,
                                                  const int* __restrict__ g,
                                                  const double* __restrict__ gamma,
                                                  const double* __restrict__ nexp,
                                                  const LinePos* __restrict__ pos,
                                                  const JobDesc* __restrict__ jd,
                                                  const TPParams* __restrict__ tp,
                                                  unsigned long long* __restrict__ emax,
                                                  VRec* __restrict__ rec, VRecX* __restrict__ recx,
                                                  int* __restrict__ kmax,
                                                  unsigned long long* __restrict__ neval) {
    const int k = blockIdx.y;
    const TPParams p = tp[k];
    const JobDesc d = jd[k];
    const bool voigt = p.profile == GX_PROF_VOIGT || p.profile == GX_PROF_VOIGT_QUAD;
    long long ne = 0;
    int km = 0;
    int em = 0;  // largest biased exponent of a line reaching the grid (0: none)
    for (int ii = blockIdx.x * blockDim.x + threadIdx.x; ii < d.n; ii += gridDim.x * blockDim.x) {
        const int64_t i = d.l0 + ii;
        const LinePos lp = pos[d.pos_off + ii];
        VRec r;
        VRecX x;
        long long nl;
        line_record(p, voigt, nu[i], E[i], A[i], g[i], gamma, nexp, i, lp, r, x, &nl);
        ne += nl;
        if (r.Kg > 0.0f) em = max(em, r.ex + kExBias);
        if (voigt && lp.ie > lp.ib) km = max(km, x.k15);
        rec[d.rec_off + ii] = r;
        recx[d.recx_off + ii] = x;
    }
    __shared__ long long sh_ne[4];
    __shared__ int sh_km[4];
    __shared__ int sh_em[4];
    ne = block_sum_ll(ne, sh_ne);
    km = block_max_i(km, sh_km);
    em = block_max_i(em, sh_em);
    if (threadIdx.x == 0) {
        if (neval != nullptr && ne != 0) atomicAdd(&neval[k], (unsigned long long)ne);
        if (kmax != nullptr && km > 0) atomicMax(&kmax[k], km);
        if (em > 0) atomicMax(&emax[k], (unsigned long long)em);
    }
}

// ---------------------------------------------------------------- k_tiles
// One work item per tile: lines whose window meets [w0, w1) form the
// contiguous range [lo, hi) because ib and ie are non-decreasing in nu.
// blockIdx.y = job - job0 (jobs job0 .. job0 + gridDim.y - 1 share the span).
__global__ __launch_bounds__(256) void k_tiles(const LinePos* __restrict__ pos,
                                               const JobDesc* __restrict__ jd,
                                               const TPParams* __restrict__ tp, int span,
                                               int job0, WorkItem* __restrict__ items,
                                               uint32_t* __restrict__ cost_key,
                                               uint32_t* __restrict__ item_idx) {
    const int job = job0 + blockIdx.y;
    const JobDesc d = jd[job];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.ntiles) return;
    const int win_lo = tp[job].win_lo, win_hi = tp[job].win_hi;
    const LinePos* jp = pos + d.pos_off;
    const int n = d.n;
    const int w0 = win_lo + t * span;
    const int w1 = min(w0 + span, win_hi);
    // lo = first i with ie > w0
    int a = 0, b = n;
    while (a < b) {
        int m = (a + b) >> 1;
        if (jp[m].ie > w0) b = m; else a = m + 1;
    }
    const int lo = a;
    // hi = first i with ib >= w1
    a = lo;
    b = n;
    while (a < b) {
        int m = (a + b) >> 1;
        if (jp[m].ib >= w1) b = m; else a = m + 1;
    }
    const int hi = a;
    // with ib and ie non-decreasing in nu, the lines whose window covers the
    // whole tile are [lf, rf): lf = first with ie >= w1, rf = first with ib > w0
    a = lo;
    b = hi;
    while (a < b) {
        const int m = (a + b) >> 1;
        if (jp[m].ie >= w1) b = m; else a = m + 1;
    }
    const int lf = a;
    a = lo;
    b = hi;
    while (a < b) {
        const int m = (a + b) >> 1;
        if (jp[m].ib > w0) b = m; else a = m + 1;
    }
    const int rf = a;
    WorkItem it;
    it.lo = lo;
    it.hi = hi;
    it.w0 = w0;
    it.w1 = w1;
    it.job = job;
    it.s1 = lf < rf ? lf : hi;
    it.s4 = lf < rf ? rf : hi;
    it.s2 = it.s3 = it.s4;  // [s1, s4) all region I until k_segments runs
    const int G = max(d.groups, 1);
    it.ngroups = G;
    for (int g = 0; g < G; ++g) {
        const int slot = d.items_off + t * G + g;
        it.group = g;
        it.slot = t * G + g;  // row of the job's partial sums
        items[slot] = it;
        if (cost_key) {
            const uint32_t c = (uint32_t)max(hi - lo, 0) / (uint32_t)G;
            cost_key[slot] = 0xFFFFFFFFu - c;  // ascending key = heaviest first
            item_idx[slot] = (uint32_t)slot;
        }
    }
}

// Near-centre segment of every work item: among the lines [s1, s4) that cover
// the whole tile (set by k_tiles), those whose centre c lies within kmax of
// the tile, [s2, s3), may have points outside Humlicek region I.  kmax = max
// k15 over the job's lines (T,P-dependent, from k_line_rec).
// out may equal in (in place) or be a per-pass copy.
__global__ __launch_bounds__(256) void k_segments(const WorkItem* in, WorkItem* out, int nitems,
                                                  const JobDesc* __restrict__ jd,
                                                  const LinePos* __restrict__ pos,
                                                  const int* __restrict__ kmax) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nitems) return;
    WorkItem it = in[t];
    int s2 = it.s4, s3 = it.s4;
    if (it.s1 < it.s4) {
        const LinePos* jp = pos + jd[it.job].pos_off;
        const int K = kmax[it.job];
        const long long cl = (long long)it.w0 - K, ch = (long long)it.w1 + K;
        int a = it.s1, b = it.s4;
        while (a < b) {  // first c >= w0 - K
            const int m = (a + b) >> 1;
            if ((long long)jp[m].c >= cl) b = m; else a = m + 1;
        }
        s2 = a;
        b = it.s4;
        while (a < b) {  // first c >= w1 + K
            const int m = (a + b) >> 1;
            if ((long long)jp[m].c >= ch) b = m; else a = m + 1;
        }
        s3 = a;
    }
    it.s2 = s2;
    it.s3 = s3;
    out[t] = it;
}

__global__ __launch_bounds__(256) void k_permute_items(const WorkItem* __restrict__ in,
                                                       const uint32_t* __restrict__ idx, int n,
                                                       WorkItem* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = in[idx[t]];
}

// --------------------------------------------------------- profile tiles
namespace {

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Line records are read-only while a profile kernel runs; addressing them in
// the constant address space lets wave-uniform loads issue as scalar loads
// (s_load) into SGPRs, so per-line decisions are scalar branches (the window
// bounds of edge lines, the records of the generic checked loop).  The
// staged and batched loops read records with vector loads through
// global-space pointers (plain pointers would be flat accesses, whose waits
// cannot be counted per load).  (The host pass of this single-source file
// only parses device code.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) LinePos* ConstPos;
typedef const __attribute__((address_space(4))) VRec* ConstRec;
typedef const __attribute__((address_space(4))) VRecX* ConstRecX;
typedef const __attribute__((address_space(1))) LinePos* GlobPos;
typedef const __attribute__((address_space(1))) VRec* GlobRec;
typedef const __attribute__((address_space(1))) VRecX* GlobRecX;
#else
typedef const LinePos* ConstPos;
typedef const VRec* ConstRec;
typedef const VRecX* ConstRecX;
typedef const LinePos* GlobPos;
typedef const VRec* GlobRec;
typedef const VRecX* GlobRecX;
#endif

// Region-I value (times the line amplitude) at x: Kl (s + 1) / (s^2 + H),
// s = x^2 + E2.  In region I (|x| + y >= 15) every term is positive: either
// y^2 >= 1/2 or x^2 > 200, so s has no cancellation.
__device__ __forceinline__ float region1(float x, float Kl, float E2, float H) {
    const float s = fmaf(x, x, E2);
    const float den = fmaf(s, s, H);
    const float num = fmaf(s, Kl, Kl);
    return num * __builtin_amdgcn_rcpf(den);
}

constexpr int kChunk = 32;  // lines per uncompensated FP32 partial sum (at most kChunk + 3)

template <int PPL>
struct TileAcc {
    float sum[PPL], cmp[PPL], part[PPL], jf[PPL];
    int jj[PPL];
    int n = 0;  // lines in part since the last fold
    int E = 0;  // the pass's largest amplitude exponent: constants are staged times 2^(ex - E)
    __device__ __forceinline__ void fold() {
#pragma unroll
        for (int q = 0; q < PPL; ++q) {
            kahan_add(sum[q], cmp[q], part[q]);
            part[q] = 0.0f;
        }
        n = 0;
    }
    __device__ __forceinline__ void tick(int k) {
        n += k;
        if (n >= kChunk) fold();
    }
};

// window masks: which window edges may fall inside the tile
enum { kMaskNone = 0, kMaskHi = 1, kMaskLo = 2, kMaskBoth = 3 };

template <int MASK>
__device__ __forceinline__ bool in_window(int j, const LinePos& lp) {
    if (MASK == kMaskNone) return true;
    if (MASK == kMaskHi) return j < lp.ie;
    if (MASK == kMaskLo) return j >= lp.ib;
    return j >= lp.ib && j < lp.ie;
}

// how many of the 64 points j0 .. j0+63 of one slice lie in the window:
// 0 none (the slice is skipped), 1 some (masked), 2 all (unmasked).  j0 and
// the window are wave-uniform, so this is a scalar branch.
template <int MASK>
__device__ __forceinline__ int slice_in_window(int j0, const LinePos& lp) {
    if (MASK == kMaskNone) return 2;
    const bool none = (MASK != kMaskLo && j0 >= lp.ie) || (MASK != kMaskHi && j0 + GX_WAVE <= lp.ib);
    if (none) return 0;
    const bool all = (MASK == kMaskHi || j0 >= lp.ib) && (MASK == kMaskLo || j0 + GX_WAVE <= lp.ie);
    return all ? 2 : 1;
}

// lines [a, b) whose every point in the tile is in Humlicek region I.  Per
// batch of 64 lines, lane k loads line a + k's record with one vector load
// and writes row k of the wave's LDS block xs: A, Kl, E2, H, then
// X0q = (cf - (w0 + 64 q)) A + B, the line's x at lane 0 of slice q, for
// every slice (rows of stage_row<PPL>() floats, 16-byte aligned).  The
// batch's lines then read their row with broadcast ds_read_b128 and, per
// point, x = X0q - lane A: one FMA, no per-line vector work and no scalar
// loads in the loop (LDS reads return in order, so the loop waits only on
// the rows it uses).  |X0q - x| <= 63 A, so x keeps FP32 accuracy except
// where |x| << 63 A; in region I there |x| + y >= 15 makes y >= 15 - |x|,
// and K's relative slope in x, below 2|x| / (x^2 + y^2), keeps the error of
// K under 1e-6.
// MASK kMaskNone: the windows cover the tile; one Kahan fold per batch.
// kMaskHi / kMaskLo: the windows end / start inside the tile; slices outside
// the window are skipped, the one the edge cuts is masked per lane (window
// bounds by scalar loads, four lines at a time).
template <int PPL>
constexpr int stage_row() { return 4 + (PPL + 3) / 4 * 4; }

template <int PPL, int MASK>
__device__ __forceinline__ void staged_masked(TileAcc<PPL>& t, const float* row,
                                              const LinePos& lp, float nl, int w0) {
    const float A = row[0], Kl = row[1], E2 = row[2], H = row[3];
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        const int st = slice_in_window<MASK>(w0 + GX_WAVE * q, lp);
        if (st == 0) continue;
        const float v = region1(fmaf(nl, A, row[4 + q]), Kl, E2, H);
        t.part[q] += (st == 2 || in_window<MASK>(t.jj[q], lp)) ? v : 0.0f;
    }
    t.tick(1);
}

template <int PPL, int MASK>
__device__ __forceinline__ void voigt_staged(TileAcc<PPL>& t, GlobRec grec, ConstPos pos, int a,
                                             int b, int w0, int lane, float* xs) {
    constexpr int R = stage_row<PPL>();
    if (a >= b) return;
    if (MASK == kMaskNone && t.n > 0) t.fold();
    const float nl = -(float)lane;
    VRec vn = grec[min(a + lane, b - 1)];
    for (int base = a; base < b; base += GX_WAVE) {
        const VRec vc = vn;
        // unconditional (the last batch reloads line b - 1): every batch has
        // the same load in flight
        vn = grec[min(base + GX_WAVE + lane, b - 1)];
        float* row = xs + lane * R;
        row[0] = vc.A;
        row[1] = ldexpf(vc.Kl, vc.ex - t.E);
        row[2] = vc.E2;
        row[3] = vc.H;
#pragma unroll
        for (int q = 0; q < PPL; ++q) row[4 + q] = fmaf(vc.cf - (float)(w0 + GX_WAVE * q), vc.A, vc.B);
        const int nb = min(GX_WAVE, b - base);
        if (MASK != kMaskNone) {
            const ConstPos pb = pos + base;
            int k = 0;
            if (nb >= 8) {
                LinePos p0 = pb[0], p1 = pb[1], p2 = pb[2], p3 = pb[3];
                for (; k + 8 <= nb; k += 4) {
                    const LinePos q0 = pb[k + 4], q1 = pb[k + 5], q2 = pb[k + 6], q3 = pb[k + 7];
                    __builtin_amdgcn_sched_barrier(0);
                    staged_masked<PPL, MASK>(t, xs + (k + 0) * R, p0, nl, w0);
                    staged_masked<PPL, MASK>(t, xs + (k + 1) * R, p1, nl, w0);
                    staged_masked<PPL, MASK>(t, xs + (k + 2) * R, p2, nl, w0);
                    staged_masked<PPL, MASK>(t, xs + (k + 3) * R, p3, nl, w0);
                    p0 = q0;
                    p1 = q1;
                    p2 = q2;
                    p3 = q3;
                }
                staged_masked<PPL, MASK>(t, xs + (k + 0) * R, p0, nl, w0);
                staged_masked<PPL, MASK>(t, xs + (k + 1) * R, p1, nl, w0);
                staged_masked<PPL, MASK>(t, xs + (k + 2) * R, p2, nl, w0);
                staged_masked<PPL, MASK>(t, xs + (k + 3) * R, p3, nl, w0);
                k += 4;
            }
            for (; k < nb; ++k) staged_masked<PPL, MASK>(t, xs + k * R, pb[k], nl, w0);
            continue;
        }
#pragma unroll 4
        for (int k = 0; k < nb; ++k) {
            const float* rk = xs + k * R;
            const float A = rk[0], Kl = rk[1], E2 = rk[2], H = rk[3];
#pragma unroll
            for (int q = 0; q < PPL; ++q) t.part[q] += region1(fmaf(nl, A, rk[4 + q]), Kl, E2, H);
        }
        t.fold();
    }
}

// Points outside the profile kernel's region I, and their value.  Voigt:
// Humlicek's regions II-IV (|x| + y < 15).  Voigt-Quad: x^2 + y^2 < kQuadR2,
// the Zaghloul-Ali core and the Gauss-Hermite wings (gx::quad_inner).  Beyond,
// both profiles take the region-I rational of the line's record.
template <int PROF>
__device__ __forceinline__ bool inner_point(float x, float y) {
    if constexpr (PROF == GX_PROF_VOIGT_QUAD) return fmaf(x, x, y * y) < gx::kQuadR2;
    return x + y < 15.0f;
}
template <int PROF>
__device__ __forceinline__ float inner_value(float x, float y) {
    if constexpr (PROF == GX_PROF_VOIGT_QUAD) return gx::quad_inner(x, y);
    return gx::w4_f32_inner(x, y);
}

// one line with per-line decisions: window masks per lane (only the edges
// MASK names; slices wholly outside the window are skipped, slices wholly
// inside are unmasked), region I inline, the inner points per lane
// DEFER: inner points add nothing here; their values come from batch_w4.
template <int PROF, int PPL, int MASK, bool DEFER = false>
__device__ __forceinline__ void checked_line(TileAcc<PPL>& t, const LinePos& lp, const VRec& r,
                                             const VRecX& rx, int w0, int w1) {
    static_assert(PROF == GX_PROF_VOIGT || PROF == GX_PROF_VOIGT_QUAD,
                  "bin-Gauss lines run gauss_staged");
    const int kmin = lp.c < w0 ? (w0 - lp.c) : (lp.c >= w1 ? (lp.c - w1 + 1) : 0);
    if (kmin >= rx.k15) {
#pragma unroll
        for (int q = 0; q < PPL; ++q) {
            const int st = slice_in_window<MASK>(w0 + GX_WAVE * q, lp);
            if (st == 0) continue;
            const float v = region1(fmaf(r.cf - t.jf[q], r.A, r.B), r.Kl, r.E2, r.H);
            if (st == 2)
                t.part[q] += v;
            else
                t.part[q] += in_window<MASK>(t.jj[q], lp) ? v : 0.0f;
        }
    } else {
#pragma unroll
        for (int q = 0; q < PPL; ++q) {
            const int st = slice_in_window<MASK>(w0 + GX_WAVE * q, lp);
            if (st == 0) continue;
            const float x = fabsf(fmaf(r.cf - t.jf[q], r.A, r.B));
            float v = region1(x, r.Kl, r.E2, r.H);
            if (inner_point<PROF>(x, rx.y)) v = DEFER ? 0.0f : r.Kg * inner_value<PROF>(x, rx.y);
            if (st == 2)
                t.part[q] += v;
            else
                t.part[q] += in_window<MASK>(t.jj[q], lp) ? v : 0.0f;
        }
    }
    t.tick(1);
}

// a record with its amplitudes scaled to the pass: Kl, Kg times 2^(ex - E)
__device__ __forceinline__ VRec to_pass(VRec r, int E) {
    r.Kl = ldexpf(r.Kl, r.ex - E);
    r.Kg = ldexpf(r.Kg, r.ex - E);
    return r;
}

// lines [a, b) through checked_line, records by scalar loads, two lines at
// a time, the next two loading while these are evaluated
template <int PROF, int PPL, int MASK>
__device__ __forceinline__ void lines_checked(TileAcc<PPL>& t, ConstPos pos, ConstRec rec,
                                              ConstRecX recx, int a, int b, int w0, int w1) {
    int i = a;
    if (b - a >= 4) {
        LinePos l0 = pos[i], l1 = pos[i + 1];
        VRec r0 = rec[i], r1 = rec[i + 1];
        VRecX x0 = recx[i], x1 = recx[i + 1];
        for (; i + 4 <= b; i += 2) {
            const LinePos m0 = pos[i + 2], m1 = pos[i + 3];
            const VRec s0 = rec[i + 2], s1 = rec[i + 3];
            const VRecX y0 = recx[i + 2], y1 = recx[i + 3];
            __builtin_amdgcn_sched_barrier(0);
            checked_line<PROF, PPL, MASK>(t, l0, to_pass(r0, t.E), x0, w0, w1);
            checked_line<PROF, PPL, MASK>(t, l1, to_pass(r1, t.E), x1, w0, w1);
            l0 = m0;
            l1 = m1;
            r0 = s0;
            r1 = s1;
            x0 = y0;
            x1 = y1;
        }
        checked_line<PROF, PPL, MASK>(t, l0, to_pass(r0, t.E), x0, w0, w1);
        checked_line<PROF, PPL, MASK>(t, l1, to_pass(r1, t.E), x1, w0, w1);
        i += 2;
    }
    for (; i < b; ++i)
        checked_line<PROF, PPL, MASK>(t, pos[i], to_pass(rec[i], t.E), recx[i], w0, w1);
}

// Lines centred near the tile (kMaskNone) staged like the region-I
// loop: per batch of 64 lines lane k loads line k's LinePos, VRec and VRecX
// with vector loads and writes a 12-word row (cf, A, B, Kl | E2, H, Kg, y |
// c, k15); each line then reads its row with three broadcast ds_read_b128,
// takes c and k15 into SGPRs (readfirstlane) for its scalar decisions and
// runs checked_line on VGPR operands.  No scalar loads in the loop.
constexpr int kNearRow = 12;

// words per line row of the wave's LDS block (the widest row the profile
// stages: Voigt region-I and near-centre rows, or bin-Gauss rows)
template <int PPL>
constexpr int gauss_row();
template <int PROF, int PPL>
constexpr int wave_row() {
    if (PROF == GX_PROF_DOPPLER) return gauss_row<PPL>();
    return stage_row<PPL>() > kNearRow ? stage_row<PPL>() : kNearRow;
}

template <int PROF, int PPL>
__device__ __forceinline__ void near_staged(TileAcc<PPL>& t, GlobPos gpos, GlobRec grec,
                                            GlobRecX grecx, int a, int b, int w0, int w1,
                                            int lane, float* xs) {
    if (a >= b) return;
    int li = min(a + lane, b - 1);
    LinePos pn = gpos[li];
    VRec rn = grec[li];
    VRecX xn = grecx[li];
    for (int base = a; base < b; base += GX_WAVE) {
        const LinePos pc = pn;
        const VRec rc = rn;
        const VRecX xc = xn;
        li = min(base + GX_WAVE + lane, b - 1);
        pn = gpos[li];
        rn = grec[li];
        xn = grecx[li];
        float* row = xs + lane * kNearRow;
        row[0] = rc.cf;
        row[1] = rc.A;
        row[2] = rc.B;
        row[3] = ldexpf(rc.Kl, rc.ex - t.E);
        row[4] = rc.E2;
        row[5] = rc.H;
        row[6] = ldexpf(rc.Kg, rc.ex - t.E);
        row[7] = xc.y;
        row[8] = __int_as_float(pc.c);
        row[9] = __int_as_float(xc.k15);
        const int nb = min(GX_WAVE, b - base);
        for (int k = 0; k < nb; ++k) {
            const float* rk = xs + k * kNearRow;
            VRec r;
            r.cf = rk[0];
            r.A = rk[1];
            r.B = rk[2];
            r.Kl = rk[3];
            r.E2 = rk[4];
            r.H = rk[5];
            r.Kg = rk[6];
            VRecX rx;
            rx.y = rk[7];
            rx.k15 = __builtin_amdgcn_readfirstlane(__float_as_int(rk[9]));
            LinePos lp;
            lp.c = __builtin_amdgcn_readfirstlane(__float_as_int(rk[8]));
            lp.ib = lp.ie = 0;  // the windows cover the tile (kMaskNone)
            checked_line<PROF, PPL, kMaskNone>(t, lp, r, rx, w0, w1);
        }
    }
}

// Bin-averaged Gaussian (Doppler and the fixed-HWHM Gaussian) staged like
// the region-I loop.  The bins of neighbouring points share an edge: point j
// spans x_j + h .. x_j - h (x falls by A = 2h per point), so each lane
// evaluates erfc once, at its point's far edge x_j - h, and takes the near
// edge from the lane before (DPP wave shift by one).  Lane 0's near edge in
// slice q is the far edge of point w0 + 64q - 1, E_q: per batch of 64 lines,
// lane k writes row k (cf, A, B, Kg | h, ib, ie, - | E_q, F_q for every
// slice), computing E_q with the very expression lane 63 of slice q - 1
// evaluates, so each shared edge is one FP32 value and a line's bins
// telescope exactly.  The slices of a line are then independent (no value
// passes from one slice to the next).  With F = erfc(|e|):
//   both edges >= 0:  erf(a) - erf(b) = F(b) - F(a)
//   both edges <  0:  erf(a) - erf(b) = F(a) - F(b)
//   a >= 0 > b:       erf(a) - erf(b) = 2 - F(a) - F(b)
// (no cancellation beyond that of the bin itself).
template <int PPL>
constexpr int gauss_row() { return 8 + (2 * PPL + 3) / 4 * 4; }

__device__ __forceinline__ float wave_shr1(float v, float lane0) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lane0), __float_as_int(v),
                                                      0x138 /* wave_shr:1 */, 0xF, 0xF, false));
}

__device__ __forceinline__ float bin_from_edges(float ea, float Fa, float eb, float Fb) {
    const float d = Fb - Fa;
    const float across = (2.0f - Fa) - Fb;
    const float v = ea < 0.0f ? -d : across;  // selects, not branches
    return eb >= 0.0f ? d : v;
}

template <int PPL>
__device__ __forceinline__ void gauss_staged(TileAcc<PPL>& t, GlobPos gpos, GlobRec grec,
                                             GlobRecX grecx, int a, int b, int w0, int lane,
                                             float* xs) {
    constexpr int R = gauss_row<PPL>();
    if (a >= b) return;
    int li = min(a + lane, b - 1);
    LinePos pn = gpos[li];
    VRec rn = grec[li];
    VRecX xn = grecx[li];
    for (int base = a; base < b; base += GX_WAVE) {
        const LinePos pc = pn;
        const VRec rc = rn;
        const VRecX xc = xn;
        li = min(base + GX_WAVE + lane, b - 1);
        pn = gpos[li];
        rn = grec[li];
        xn = grecx[li];
        float* row = xs + lane * R;
        row[0] = rc.cf;
        row[1] = rc.A;
        row[2] = rc.B;
        row[3] = ldexpf(rc.Kg, rc.ex - t.E);
        row[4] = xc.y;
        row[5] = __int_as_float(pc.ib);
        row[6] = __int_as_float(pc.ie);
#pragma unroll
        for (int q = 0; q < PPL; ++q) {
            const float e = fmaf(rc.cf - (float)(w0 + GX_WAVE * q - 1), rc.A, rc.B) - xc.y;
            row[8 + 2 * q] = e;
            row[9 + 2 * q] = gx::erfc_nonneg(fabsf(e));
        }
        // Per line the window test is per lane (VALU compares into a
        // select); the only scalar decision is whether a slice meets the
        // window at all.  The partial sums are folded once per chunk of
        // kChunk lines rather than counted line by line: with a bare scalar
        // unit per CU shared by four SIMDs, per-line scalar work would bound
        // this loop.
        const int nb = min(GX_WAVE, b - base);
        if (t.n > 0) t.fold();
        for (int k0 = 0; k0 < nb; k0 += kChunk) {
            const int k1 = min(nb, k0 + kChunk);
            for (int k = k0; k < k1; ++k) {
                const float* rk = xs + k * R;
                const float cf = rk[0], A = rk[1], B = rk[2], Kg = rk[3], h = rk[4];
                const int ib = __builtin_amdgcn_readfirstlane(__float_as_int(rk[5]));
                const int ie = __builtin_amdgcn_readfirstlane(__float_as_int(rk[6]));
#pragma unroll
                for (int q = 0; q < PPL; ++q) {
                    const int j0 = w0 + GX_WAVE * q;
                    if (j0 >= ie || j0 + GX_WAVE <= ib) continue;  // slice outside the window
                    const float x = fmaf(cf - t.jf[q], A, B);
                    const float eb = x - h;
                    const float Fb = gx::erfc_nonneg(fabsf(eb));
                    const float ea = wave_shr1(eb, rk[8 + 2 * q]);
                    const float Fa = wave_shr1(Fb, rk[9 + 2 * q]);
                    const float v = bin_from_edges(ea, Fa, eb, Fb);
                    const bool in = t.jj[q] >= ib && t.jj[q] < ie;
                    t.part[q] = fmaf(Kg, in ? v : 0.0f, t.part[q]);
                }
            }
            t.fold();
        }
    }
}

__device__ __forceinline__ float bcast(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ __forceinline__ int bcast(int v, int k) { return __builtin_amdgcn_readlane(v, k); }

// W4 values of one batch of lines at one point per lane (64-point tiles),
// compacted: a line's points outside Humlicek region I (|x| + y < 15, all
// within k15 of its centre) are few and adjacent, so evaluating W4 line by
// line leaves most lanes idle.  Lane k holds line k's records (rc, pc, xc);
// its candidates are the points of the tile and of its window with
// |c - j| < k15.  A wave scan numbers the candidates, each lane takes one
// (binary search in the scan, held in the wave's 64-int LDS block incl),
// repeats the region test of checked_line bit for bit and evaluates W4;
// the values then go to the lanes owning the points line by line in batch
// order (ds_bpermute), so each point sums them in a fixed order.
template <int PROF, int MASK>
__device__ __forceinline__ void batch_w4(TileAcc<1>& t, const VRec& rc, const LinePos& pc,
                                         const VRecX& xc, int nb, int w0, int w1, int lane,
                                         int* incl) {
    int n = 0, jl = 0;
    if (lane < nb && xc.k15 > 0) {
        jl = max(pc.c - xc.k15 + 1, w0);
        int jh = min(pc.c + xc.k15, w1);
        if (MASK == kMaskLo || MASK == kMaskBoth) jl = max(jl, pc.ib);
        if (MASK == kMaskHi || MASK == kMaskBoth) jh = min(jh, pc.ie);
        n = max(jh - jl, 0);
    }
    int sc = n;
#pragma unroll
    for (int d = 1; d < GX_WAVE; d <<= 1) {
        const int u = __shfl_up(sc, d);
        if (lane >= d) sc += u;
    }
    const int total = __builtin_amdgcn_readlane(sc, GX_WAVE - 1);
    if (total == 0) return;
    incl[lane] = sc;
    const int ex = sc - n;  // the first candidate of this lane's line
    for (int p0 = 0; p0 < total; p0 += GX_WAVE) {
        const int pp = p0 + lane;
        int k = 0;  // the line of candidate pp: smallest k with incl[k] > pp
        if (pp < total) {
#pragma unroll
            for (int step = GX_WAVE / 2; step > 0; step >>= 1)
                if (incl[k + step - 1] <= pp) k += step;
        }
        const float cf = __shfl(rc.cf, k), A = __shfl(rc.A, k), B = __shfl(rc.B, k);
        const float Kg = __shfl(rc.Kg, k), y = __shfl(xc.y, k);
        const int jlk = __shfl(jl, k), exk = __shfl(ex, k);
        float v = 0.0f;
        if (pp < total) {
            const int j = jlk + (pp - exk);
            const float x = fabsf(fmaf(cf - (float)j, A, B));
            if (inner_point<PROF>(x, y)) v = Kg * inner_value<PROF>(x, y);
        }
        const int last = min(total - p0, GX_WAVE) - 1;
        const int k0 = __builtin_amdgcn_readlane(k, 0), k1 = __builtin_amdgcn_readlane(k, last);
        for (int kk = k0; kk <= k1; ++kk) {
            const int jq = __builtin_amdgcn_readlane(jl, kk);
            const int eq = __builtin_amdgcn_readlane(ex, kk);
            const int nq = __builtin_amdgcn_readlane(n, kk);
            const int d = t.jj[0] - jq;    // this lane's point among line kk's candidates
            const int L = eq + d - p0;     // the lane holding that candidate in this chunk
            const bool mine = d >= 0 && d < nq && L >= 0 && L < GX_WAVE;
            const float got = __shfl(v, L & (GX_WAVE - 1));
            t.part[0] += mine ? got : 0.0f;
            t.tick(1);
        }
    }
}

// One point per lane (64-point tiles), lines with per-line decisions: a
// line's evaluation is a handful of VALU instructions, too few to cover a
// scalar load's latency when few waves share a SIMD.  Here lane k loads the
// records of line base + k with vector loads (64 lines per batch, the next
// batch in flight while this one is evaluated) and each line is broadcast to
// the wave with readlane.  Voigt's W4 points are left to batch_w4.
template <int PROF, int MASK>
__device__ __forceinline__ void lines_batched(TileAcc<1>& t, GlobPos pos, GlobRec rec,
                                              GlobRecX recx, int a, int b, int w0, int w1,
                                              int lane, int* incl) {
    constexpr bool kDefer = true;
    if (a >= b) return;
    LinePos pc, pn;
    VRec rc, rn;
    VRecX xc, xn;
    {
        const int li = min(a + lane, b - 1);
        rc = rec[li];
        pc = pos[li];
        xc = recx[li];
    }
    for (int base = a; base < b; base += GX_WAVE) {
        const int nb = min(GX_WAVE, b - base);
        rc = to_pass(rc, t.E);  // lane k's line, amplitudes scaled to the pass
        {
            // unconditional (the last batch reloads line b - 1), so every
            // iteration has the same loads in flight and the wait before the
            // broadcasts covers only the current batch
            const int li = min(base + GX_WAVE + lane, b - 1);
            rn = rec[li];
            pn = pos[li];
            xn = recx[li];
        }
        for (int k = 0; k < nb; ++k) {
            VRec r;
            r.cf = bcast(rc.cf, k);
            r.A = bcast(rc.A, k);
            r.B = bcast(rc.B, k);
            r.Kl = bcast(rc.Kl, k);
            r.E2 = bcast(rc.E2, k);
            r.H = bcast(rc.H, k);
            r.Kg = bcast(rc.Kg, k);
            LinePos lp;
            lp.ib = bcast(pc.ib, k);
            lp.ie = bcast(pc.ie, k);
            lp.c = bcast(pc.c, k);
            VRecX rx;
            rx.y = bcast(xc.y, k);
            rx.k15 = bcast(xc.k15, k);
            checked_line<PROF, 1, MASK, kDefer>(t, lp, r, rx, w0, w1);
        }
        if (kDefer) batch_w4<PROF, MASK>(t, rc, pc, xc, nb, w0, w1, lane, incl);
        rc = rn;
        pc = pn;
        xc = xn;
    }
}

// Kahan sums of the lines [ca, cb) of the item (a sub-range of [lo, hi)) at
// the 64*PPL points of the tile; lines stay in nu order across the segments
// (init false: add to the sums already in t)
template <int PROF, int PPL>
__device__ __forceinline__ void sum_lines(TileAcc<PPL>& t, const WorkItem& it, const Job& jb,
                                          int lane, int ca, int cb, float* xs, bool init = true) {
    const ConstPos pos = (ConstPos)jb.pos;
    const ConstRec rec = (ConstRec)jb.rec;
    const ConstRecX recx = (ConstRecX)jb.recx;
    if (init) {
#pragma unroll
        for (int q = 0; q < PPL; ++q) {
            t.sum[q] = 0.0f;
            t.cmp[q] = 0.0f;
            t.part[q] = 0.0f;
            t.jj[q] = it.w0 + lane + GX_WAVE * q;
            t.jf[q] = (float)t.jj[q];
        }
        t.n = 0;
    }
    const int w0 = it.w0, w1 = it.w1;
    if constexpr (PROF == GX_PROF_DOPPLER) {
        gauss_staged<PPL>(t, (GlobPos)jb.pos, (GlobRec)jb.rec, (GlobRecX)jb.recx, ca, cb, w0, lane,
                          xs);
    } else if constexpr (PPL == 1) {
        // region-I segments staged, the others by vector loads broadcast
        // per line; the wave's LDS block holds the staged X0 or the W4
        // candidate scan of batch_w4
        int* incl = reinterpret_cast<int*>(xs);
        const GlobPos gpos = (GlobPos)jb.pos;
        const GlobRec grec = (GlobRec)jb.rec;
        const GlobRecX grecx = (GlobRecX)jb.recx;
        static_assert(PROF == GX_PROF_VOIGT || PROF == GX_PROF_VOIGT_QUAD, "Voigt profiles");
        if (it.s1 < it.s4) {
            const int s1 = min(max(it.s1, ca), cb), s2 = min(max(it.s2, ca), cb);
            const int s3 = min(max(it.s3, ca), cb), s4 = min(max(it.s4, ca), cb);
            // the region-I segments through the staged loop (6 VALU per
            // line against 14 with broadcasts), the rest batched
            const ConstPos pos = (ConstPos)jb.pos;
            if (it.s2 > it.s1)
                voigt_staged<1, kMaskHi>(t, grec, pos, ca, s1, w0, lane, xs);
            else
                lines_batched<PROF, kMaskHi>(t, gpos, grec, grecx, ca, s1, w0, w1, lane, incl);
            voigt_staged<1, kMaskNone>(t, grec, pos, s1, s2, w0, lane, xs);
            lines_batched<PROF, kMaskNone>(t, gpos, grec, grecx, s2, s3, w0, w1, lane, incl);
            voigt_staged<1, kMaskNone>(t, grec, pos, s3, s4, w0, lane, xs);
            if (it.s3 < it.s4)
                voigt_staged<1, kMaskLo>(t, grec, pos, s4, cb, w0, lane, xs);
            else
                lines_batched<PROF, kMaskLo>(t, gpos, grec, grecx, s4, cb, w0, w1, lane, incl);
        } else {
            lines_batched<PROF, kMaskBoth>(t, gpos, grec, grecx, ca, cb, w0, w1, lane, incl);
        }
    } else {
        const int s1 = min(max(it.s1, ca), cb), s2 = min(max(it.s2, ca), cb);
        const int s3 = min(max(it.s3, ca), cb), s4 = min(max(it.s4, ca), cb);
        if (it.s1 < it.s4) {
            // [lo, s1): ib <= w0 < ie < w1; [s1, s4): ib <= w0, ie >= w1;
            // [s4, hi): w0 < ib < w1 <= ie.
            // ν order makes c non-decreasing: when the first whole-tile line
            // is far below the tile (s2 > s1) so are all of [lo, s1); when
            // the last one is far above (s3 < s4) so are all of [s4, hi)
            const GlobRec grec = (GlobRec)jb.rec;
            if (it.s2 > it.s1)
                voigt_staged<PPL, kMaskHi>(t, grec, pos, ca, s1, w0, lane, xs);
            else
                lines_checked<PROF, PPL, kMaskHi>(t, pos, rec, recx, ca, s1, w0, w1);
            voigt_staged<PPL, kMaskNone>(t, grec, pos, s1, s2, w0, lane, xs);
            near_staged<PROF, PPL>(t, (GlobPos)jb.pos, grec, (GlobRecX)jb.recx, s2, s3, w0, w1, lane,
                                   xs);
            voigt_staged<PPL, kMaskNone>(t, grec, pos, s3, s4, w0, lane, xs);
            if (it.s3 < it.s4)
                voigt_staged<PPL, kMaskLo>(t, grec, pos, s4, cb, w0, lane, xs);
            else
                lines_checked<PROF, PPL, kMaskLo>(t, pos, rec, recx, s4, cb, w0, w1);
        } else {
            lines_checked<PROF, PPL, kMaskBoth>(t, pos, rec, recx, ca, cb, w0, w1);
        }
    }
    t.fold();
}

__device__ __forceinline__ void store_point(const Job& jb, int j, double v) {
    if (jb.accumulate)
        jb.out[j] += v;
    else
        jb.out[j] = v;
}

// line k * n / K of a range of n lines starting at lo
__device__ __forceinline__ int share_cut(int lo, int n, int k, int K) {
    return lo + (int)(((long long)n * k) / K);
}

// Kahan sums of share k of K of the tile's lines [lo, hi) (K = line groups x
// waves per group).  The near-centre lines [s2, s3) evaluate W4 over part of
// the tile and cost up to ~20x a region-I line, and they are contiguous in nu
// order, so one equal cut of [lo, hi) gives them all to one share.  When a
// Voigt tile has at least K of them, each of [lo, s2), [s2, s3) and [s3, hi)
// is cut into K equal contiguous shares instead and share k sums its three
// pieces in nu order: every share carries 1/K of each segment's work.
template <int PROF, int PPL>
__device__ __forceinline__ void sum_share(TileAcc<PPL>& t, const WorkItem& it, const Job& jb,
                                          int lane, int k, int K, float* xs) {
    const int lo = it.lo, hi = max(it.hi, it.lo);
    const int m2 = min(max(it.s2, lo), hi), m3 = min(max(it.s3, lo), hi);
    const bool three = PROF != GX_PROF_DOPPLER && K > 1 && it.s1 < it.s4 && m3 - m2 >= K;
    const int np = three ? 3 : 1;
    // one loop, so the kernel holds one copy of sum_lines
#pragma nounroll
    for (int p = 0; p < np; ++p) {
        const int a = !three ? lo : p == 0 ? lo : p == 1 ? m2 : m3;
        const int b = !three ? hi : p == 0 ? m2 : p == 1 ? m3 : hi;
        sum_lines<PROF, PPL>(t, it, jb, lane, share_cut(a, b - a, k, K),
                             share_cut(a, b - a, k + 1, K), xs, p == 0);
    }
}

// value k of the tile (point w0 + k): to the grid, or with line groups to
// the item's row of partial sums
template <int PPL>
__device__ __forceinline__ void store_tile(const WorkItem& it, const Job& jb, int k, double v) {
    if (it.ngroups > 1)
        jb.part[(size_t)it.slot * (GX_WAVE * PPL) + k] = v;
    else if (it.w0 + k < it.w1)
        store_point(jb, it.w0 + k, v);
}

// one wave evaluates the whole item; the sums are in units of 2^E (the
// pass exponent, a scalar), restored in FP64 at the store
template <int PROF, int PPL>
__device__ __forceinline__ void run_item(const WorkItem& it, const Job& jb,
                                         const unsigned long long* scale_bits, int lane, float* xs) {
    TileAcc<PPL> t;
    t.E = pass_exponent(scale_bits[jb.scale_slot]);
    sum_share<PROF, PPL>(t, it, jb, lane, it.group, max(it.ngroups, 1), xs);
    const double inv_scale = ldexp(1.0, t.E);
#pragma unroll
    for (int q = 0; q < PPL; ++q)
        store_tile<PPL>(it, jb, GX_WAVE * q + lane,
                        ((double)t.sum[q] - (double)t.cmp[q]) * inv_scale);
}

constexpr int kWavesPerBlock = GX_WG / GX_WAVE;  // SPLIT 1: items per workgroup

// workgroup size of a split: 4 waves (SPLIT 1 and 4) or 16 waves (SPLIT 16)
template <int SPLIT>
constexpr int block_threads() { return SPLIT > 4 ? SPLIT * GX_WAVE : GX_WG; }

// the SPLIT waves of the workgroup take the item's shares group * SPLIT + wv
// of its lines; their FP64 partial sums are added in wave order through LDS.
// red shares its LDS with the waves' staging rows (xs): it is written only
// once every wave of the workgroup has left its line loops.
template <int PROF, int PPL, int SPLIT>
__device__ __forceinline__ void run_item_split(const WorkItem& it, const Job& jb,
                                               const unsigned long long* scale_bits, int lane,
                                               int wv, double* red, float* xs) {
    TileAcc<PPL> t;
    t.E = pass_exponent(scale_bits[jb.scale_slot]);
    sum_share<PROF, PPL>(t, it, jb, lane, it.group * SPLIT + wv, max(it.ngroups, 1) * SPLIT, xs);
    // each wave's FP64 partial in absolute units (a power-of-two scaling in
    // FP64 is exact, so the sum equals the one scaled after adding)
    const double to_abs = ldexp(1.0, t.E);
    __syncthreads();  // the staging rows are free
#pragma unroll
    for (int q = 0; q < PPL; ++q)
        red[(size_t)wv * GX_WAVE * PPL + GX_WAVE * q + lane] =
            ((double)t.sum[q] - (double)t.cmp[q]) * to_abs;
    __syncthreads();
    for (int k = threadIdx.x; k < GX_WAVE * PPL; k += block_threads<SPLIT>()) {
        double v = red[k];
#pragma unroll
        for (int w = 1; w < SPLIT; ++w) v += red[(size_t)w * GX_WAVE * PPL + k];
        store_tile<PPL>(it, jb, k, v);
    }
    __syncthreads();  // red (the staging rows) is reused by the next item
}

// block b runs on XCD b % GX_XCDS; map the blocks of one XCD to one
// contiguous range of logical blocks so neighbouring tiles share its L2
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b % GX_XCDS, k = b / GX_XCDS;
    const int per = nb / GX_XCDS, rem = nb % GX_XCDS;
    return x * per + min(x, rem) + k;
}

}  // namespace

// single != 0: every item belongs to jobs[0], which the waves then load
// without waiting for their item (no dependent load).  The choice is an index
// select: a select between a loaded Job and one passed by value puts the Job's
// fields in VGPRs and spills the line loops' registers.
// mode 0: one item per wave (SPLIT 1) or per workgroup (SPLIT 4, 16), blocks in
// XCD-contiguous order; 3: as 0 in dispatch order (consecutive items on
// different XCDs); 1: persistent, items pulled from one counter; 2:
// persistent, items pulled from the queue of the workgroup's XCD (items in
// grid order, split into GX_XCDS cost-balanced contiguous ranges), then from
// the other queues.  xq = GX_XCDS counters (zero at launch), xqb = the
// GX_XCDS + 1 queue bounds.
// Waves per SIMD the register allocation aims at.  7 (72 VGPRs): at 4 points
// per lane the kernel takes 86 VGPRs unconstrained (5 waves per SIMD); held
// to 72 it spills two values outside its loops and runs 4 % faster, the
// extra waves hiding the scalar and LDS load latency of the line loops.
// 8-point tiles need up to ~150 VGPRs and 4-point tiles with 16 waves per
// workgroup ~100: held to 72 they spill inside the line loops, so they aim at
// 3 and 4 waves.
template <int PPL, int SPLIT>
constexpr int waves_target() { return PPL >= 8 ? 3 : ((PPL == 4 && SPLIT == 16) ? 4 : 7); }

template <int PROF, int PPL, int SPLIT>
__global__ __launch_bounds__(block_threads<SPLIT>()) __attribute__((amdgpu_waves_per_eu(waves_target<PPL, SPLIT>()))) void k_profile_tiles(const WorkItem* __restrict__ items,
                                                         int nitems, const Job* __restrict__ jobs,
                                                         int single,
                                                         const unsigned long long* __restrict__ scale_bits,
                                                         unsigned int* __restrict__ xq,
                                                         const unsigned int* __restrict__ xqb, int mode) {
    // per wave, 64 rows of wave_row() words: the staged rows of one batch of
    // 64 lines, or at PPL 1 the W4 candidate scan of one batch; after the
    // line loops of a split item, the waves' FP64 partial sums (red)
    constexpr int kXsWords = block_threads<SPLIT>() * wave_row<PROF, PPL>();
    constexpr int kRedWords = SPLIT > 1 ? 2 * SPLIT * GX_WAVE * PPL : 0;
    __shared__ __attribute__((aligned(16))) float smem[kXsWords > kRedWords ? kXsWords : kRedWords];
    __shared__ int grab;
    float* xs_all = smem;
    double* red = reinterpret_cast<double*>(smem);
    const int lane = threadIdx.x & (GX_WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float* xs = xs_all + wv * GX_WAVE * wave_row<PROF, PPL>();
    const int lb = mode == 3 ? (int)blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
    if (mode == 0 || mode == 3) {
        if (SPLIT == 1) {
            const int idx = lb * kWavesPerBlock + wv;
            if (idx >= nitems) return;
            const WorkItem it = items[idx];
            const Job jb = jobs[single ? 0 : it.job];
            run_item<PROF, PPL>(it, jb, scale_bits, lane, xs);
        } else {
            if (lb >= nitems) return;
            const WorkItem it = items[lb];
            const Job jb = jobs[single ? 0 : it.job];
            run_item_split<PROF, PPL, SPLIT>(it, jb, scale_bits, lane, wv, red, xs);
        }
        return;
    }
    const int nq = mode == 2 ? GX_XCDS : 1;
    const int home = mode == 2 ? (int)(blockIdx.x % GX_XCDS) : 0;
    for (int k = 0; k < nq; ++k) {
        const int q = (home + k) % nq;
        const int beg = mode == 2 ? (int)xqb[q] : 0;
        const int end = mode == 2 ? (int)xqb[q + 1] : nitems;
        for (;;) {
            int idx;
            if (SPLIT == 1) {
                unsigned int t = 0;
                if (lane == 0) t = atomicAdd(&xq[q], 1u);
                idx = beg + __builtin_amdgcn_readfirstlane((int)t);
            } else {
                if (threadIdx.x == 0) grab = (int)atomicAdd(&xq[q], 1u);
                __syncthreads();
                idx = beg + grab;
                __syncthreads();  // grab is rewritten by the next pull
            }
            if (idx >= end) break;
            const WorkItem it = items[idx];
            const Job jb = jobs[single ? 0 : it.job];
            if (SPLIT == 1)
                run_item<PROF, PPL>(it, jb, scale_bits, lane, xs);
            else
                run_item_split<PROF, PPL, SPLIT>(it, jb, scale_bits, lane, wv, red, xs);
        }
    }
}

// Line groups: point k of a tile is the sum of its groups' partial rows in
// group order (fixed, so a result does not depend on the launch).  One
// workgroup of `span` threads per item; only items of group 0 write.
__global__ __launch_bounds__(512) void k_group_sum(const WorkItem* __restrict__ items, int nitems,
                                                   const Job* __restrict__ jobs, int span) {
    const int i = blockIdx.x;
    if (i >= nitems) return;
    const WorkItem it = items[i];
    if (it.ngroups <= 1 || it.group != 0) return;
    const int k = threadIdx.x;
    if (k >= span || it.w0 + k >= it.w1) return;
    const Job jb = jobs[it.job];
    const double* row = jb.part + (size_t)it.slot * span + k;
    double v = 0.0;
    for (int g = 0; g < it.ngroups; ++g) v += row[(size_t)g * span];
    store_point(jb, it.w0 + k, v);
}

// ------------------------------------------------------ plain FP64 baseline
__global__ __launch_bounds__(256) void k_line_f64(const double* __restrict__ nu,
                                                  const double* __restrict__ E,
                                                  const double* __restrict__ A,
                                                  const int* __restrict__ g,
                                                  const double* __restrict__ gamma,
                                                  const double* __restrict__ nexp, int n,
                                                  const LinePos* __restrict__ pos,
                                                  const TPParams* __restrict__ tp,
                                                  LineF64* __restrict__ out,
                                                  unsigned long long* __restrict__ neval) {
    const TPParams p = tp[0];
    long long ne = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double S = line_strength(p, nu[i], E[i], A[i], g[i]);
        double iw = inv_width(p, nu[i]);
        LinePos lp = pos[i];
        if (S != 0.0 && lp.ie > lp.ib) ne += lp.ie - lp.ib;
        LineF64 r;
        if (p.profile == GX_PROF_VOIGT || p.profile == GX_PROF_VOIGT_QUAD) {
            r.Kg = S * iw;
            r.y = (gamma != nullptr) ? gamma_l(p, gamma[i], nexp[i]) * iw : 0.0;
        } else {
            r.Kg = S;
            r.y = 0.0;
        }
        r.inv_w = iw;
        out[i] = r;
    }
    __shared__ long long sh_ne[4];
    ne = block_sum_ll(ne, sh_ne);
    if (neval != nullptr && threadIdx.x == 0 && ne != 0) atomicAdd(neval, (unsigned long long)ne);
}

template <int PROF>
__global__ __launch_bounds__(256) void k_fp64_points(const double* __restrict__ nu,
                                                     const LinePos* __restrict__ pos,
                                                     const LineF64* __restrict__ lf, int n,
                                                     double start, double dfreq, int win_lo,
                                                     int win_hi, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int j = win_lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= win_hi) return;
    // lines with ib <= j < ie: [first ie > j, first ib > j)
    int a = 0, b = n;
    while (a < b) {
        int m = (a + b) >> 1;
        if (pos[m].ie > j) b = m; else a = m + 1;
    }
    const int lo = a;
    a = lo;
    b = n;
    while (a < b) {
        int m = (a + b) >> 1;
        if (pos[m].ib > j) b = m; else a = m + 1;
    }
    const int hi = a;
    const double f = start + (double)j * dfreq;
    const double x0 = dfreq * 0.5;
    double acc = 0.0;
    for (int i = lo; i < hi; ++i) {
        const LineF64 r = lf[i];
        double v = nu[i] == 0.0 ? 1e-6 : nu[i];
        double d = fabs(v - f);
        if (PROF == GX_PROF_VOIGT) {
            acc += r.Kg * gx::w4_f64(d * r.inv_w, r.y);
        } else if (PROF == GX_PROF_VOIGT_QUAD) {
            double x = d * r.inv_w;
            double k = (x * x + r.y * r.y < 100.0) ? gx::voigt916_f64(x, r.y) : gx::gh11_f64(x, r.y);
            acc += r.Kg * k;
        } else {
            double xp = r.inv_w * (d + x0), xm = r.inv_w * (d - x0);
            acc += r.Kg * (erf(xp) - erf(xm));
        }
    }
    if (PROF == GX_PROF_VOIGT || PROF == GX_PROF_VOIGT_QUAD)
        out[j] += acc;
    else
        out[j] += acc * (0.5 / dfreq);
}

// ------------------------------------------------------- partition function
// Q(T_k) = sum_i g_i exp(-beta_k E_i).  Block b sums states [b*chunk, (b+1)*chunk)
// with a fixed per-thread stride and a fixed LDS tree, then one block per T
// adds the partials in block order: the result does not depend on timing.
__global__ __launch_bounds__(256) void k_partition(const double* __restrict__ E,
                                                   const int* __restrict__ g, int64_t n,
                                                   int64_t chunk, const double* __restrict__ T,
                                                   double* __restrict__ partial, int nblk) {
    __shared__ double red[256];
    const int k = blockIdx.y;
    const double beta = GX_PLANCK * GX_VELLGT / (GX_BOLTZ * T[k]);
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    const int64_t b1 = b0 + chunk < n ? b0 + chunk : n;
    double s = 0.0;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += 256) s += g[i] * exp(-beta * E[i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)k * nblk + blockIdx.x] = red[0];
}

// Q[k] = sum of the partials of temperature k; with jd, also tp[j].Q = Q[k]
// for every job j whose q_slot is k, and block 0 zeroes clear[0 .. nclear)
// (the pass counters, so that one launch does both)
__global__ __launch_bounds__(256) void k_partition_final(const double* __restrict__ partial,
                                                         int nblk, double* __restrict__ Q,
                                                         const JobDesc* __restrict__ jd, int njobs,
                                                         TPParams* __restrict__ tp,
                                                         unsigned long long* __restrict__ clear,
                                                         int nclear) {
    __shared__ double red[256];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += partial[(int64_t)k * nblk + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) Q[k] = red[0];
    if (jd != nullptr)
        for (int j = threadIdx.x; j < njobs; j += 256)
            if (jd[j].q_slot == k) tp[j].Q = red[0];
    if (k == 0)
        for (int i = threadIdx.x; i < nclear; i += 256) clear[i] = 0ull;
}

// ------------------------------------------------------------- small kernels
// One job's tables written from kernel arguments and nclear 64-bit counters
// zeroed: the set-up of a pass whose tables live in device memory the host
// cannot stage into.
__global__ void k_write_tables(TPParams tp, JobDesc jd, Job jb, TPParams* __restrict__ dtp,
                               JobDesc* __restrict__ djd, Job* __restrict__ djob,
                               unsigned long long* __restrict__ clear, int nclear) {
    if (threadIdx.x == 0) {
        *dtp = tp;
        *djd = jd;
        *djob = jb;
    }
    for (int i = threadIdx.x; i < nclear; i += blockDim.x) clear[i] = 0ull;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double* __restrict__ p, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_add_f64(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] += src[i];
}

// ---------------------------------------------------------- host launchers
namespace gxk {

static inline int blocks_for(int64_t n, int per) { return (int)((n + per - 1) / per); }

hipError_t line_pos(hipStream_t s, const double* nu, const JobDesc* jd, int njobs,
                    int64_t max_n, const TPParams* tp, LinePos* pos) {
    if (njobs <= 0 || max_n <= 0) return hipSuccess;
    int gx = std::min(blocks_for(max_n, 256), 2048);
    hipLaunchKernelGGL(k_line_pos, dim3(gx, njobs), dim3(256), 0, s, nu, jd, tp, pos);
    return hipGetLastError();
}

hipError_t line_rec(hipStream_t s, const double* nu, const double* E, const double* A,
                    const int* g, const double* gamma, const double* nexp, const LinePos* pos,
                    const JobDesc* jd, int njobs, int64_t max_n, const TPParams* tp,
                    unsigned long long* emax, VRec* rec, VRecX* recx, int* kmax,
                    unsigned long long* neval, int max_blocks) {
    if (njobs <= 0 || max_n <= 0) return hipSuccess;
    int gx = std::min(blocks_for(max_n, 256), max_blocks > 0 ? max_blocks : 2048);
    hipLaunchKernelGGL(k_line_rec, dim3(gx, njobs), dim3(256), 0, s, nu, E, A, g, gamma, nexp, pos,
                       jd, tp, emax, rec, recx, kmax, neval);
    return hipGetLastError();
}

hipError_t segments(hipStream_t s, const WorkItem* in, WorkItem* out, int nitems,
                    const JobDesc* jd, const LinePos* pos, const int* kmax, int profile) {
    // only the Voigt profiles read them
    if (nitems <= 0 || (profile != GX_PROF_VOIGT && profile != GX_PROF_VOIGT_QUAD)) return hipSuccess;
    hipLaunchKernelGGL(k_segments, dim3(blocks_for(nitems, 256)), dim3(256), 0, s, in, out, nitems,
                       jd, pos, kmax);
    return hipGetLastError();
}

hipError_t tiles(hipStream_t s, const LinePos* pos, const JobDesc* jd, int njobs, int max_tiles,
                 const TPParams* tp, int span, WorkItem* items, uint32_t* cost_key,
                 uint32_t* item_idx, int job0) {
    if (njobs <= 0 || max_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tiles, dim3(blocks_for(max_tiles, 256), njobs), dim3(256), 0, s, pos, jd,
                       tp, span, job0, items, cost_key, item_idx);
    return hipGetLastError();
}

hipError_t group_sum(hipStream_t s, const WorkItem* items, int nitems, const Job* jobs, int span) {
    if (nitems <= 0) return hipSuccess;
    if (span <= 0 || span > 512) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_group_sum, dim3(nitems), dim3(span), 0, s, items, nitems, jobs, span);
    return hipGetLastError();
}

hipError_t permute_items(hipStream_t s, const WorkItem* in, const uint32_t* idx, int n,
                         WorkItem* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_permute_items, dim3(blocks_for(n, 256)), dim3(256), 0, s, in, idx, n,
                       out);
    return hipGetLastError();
}

template <int PROF, int PPL, int SPLIT>
static hipError_t launch_tiles_t(hipStream_t s, const WorkItem* items, int nitems,
                                 const Job* jobs, int single,
                                 const unsigned long long* scale_bits, unsigned int* xq,
                                 const unsigned int* xqb, int mode, int blocks) {
    if (nitems <= 0) return hipSuccess;
    const int ipb = SPLIT == 1 ? kWavesPerBlock : 1;  // items per block
    const bool slots = mode == 0 || mode == 3;  // one launch slot per item
    const int nb = slots ? blocks_for(nitems, ipb)
                         : std::max(1, std::min(blocks, blocks_for(nitems, ipb)));
    hipLaunchKernelGGL((k_profile_tiles<PROF, PPL, SPLIT>), dim3(nb), dim3(block_threads<SPLIT>()), 0, s, items,
                       nitems, jobs, single, scale_bits, xq, xqb, mode);
    return hipGetLastError();
}

template <int PROF, int SPLIT>
static hipError_t launch_ppl(hipStream_t s, int ppl, const WorkItem* items, int nitems,
                             const Job* jobs, int single, const unsigned long long* scale_bits,
                             unsigned int* xq, const unsigned int* xqb, int mode, int blocks) {
    switch (ppl) {
        case 1: return launch_tiles_t<PROF, 1, SPLIT>(s, items, nitems, jobs, single, scale_bits, xq, xqb, mode, blocks);
        case 2: return launch_tiles_t<PROF, 2, SPLIT>(s, items, nitems, jobs, single, scale_bits, xq, xqb, mode, blocks);
        case 4: return launch_tiles_t<PROF, 4, SPLIT>(s, items, nitems, jobs, single, scale_bits, xq, xqb, mode, blocks);
        case 8: return launch_tiles_t<PROF, 8, SPLIT>(s, items, nitems, jobs, single, scale_bits, xq, xqb, mode, blocks);
        default: return hipErrorInvalidValue;
    }
}

template <int PROF>
static hipError_t launch_prof(hipStream_t s, int ppl, int split, const WorkItem* items,
                              int nitems, const Job* jobs, int single,
                              const unsigned long long* scale_bits,
                              unsigned int* xq, const unsigned int* xqb, int mode, int blocks) {
    if (split == 1) return launch_ppl<PROF, 1>(s, ppl, items, nitems, jobs, single, scale_bits, xq, xqb, mode, blocks);
    if (split == 4) return launch_ppl<PROF, 4>(s, ppl, items, nitems, jobs, single, scale_bits, xq, xqb, mode, blocks);
    if (split == 16)
        return launch_ppl<PROF, 16>(s, ppl, items, nitems, jobs, single, scale_bits, xq, xqb, mode, blocks);
    return hipErrorInvalidValue;
}

hipError_t profile_tiles(hipStream_t s, int profile, int ppl, int split, const WorkItem* items,
                         int nitems, const Job* jobs, const unsigned long long* scale_bits,
                         unsigned int* xq, const unsigned int* xqb, int mode, int blocks,
                         bool single_job) {
    const int single = single_job ? 1 : 0;
    switch (profile) {
        case GX_PROF_VOIGT:
            return launch_prof<GX_PROF_VOIGT>(s, ppl, split, items, nitems, jobs, single, scale_bits, xq, xqb, mode,
                                              blocks);
        case GX_PROF_VOIGT_QUAD:
            return launch_prof<GX_PROF_VOIGT_QUAD>(s, ppl, split, items, nitems, jobs, single, scale_bits, xq, xqb,
                                                   mode, blocks);
        case GX_PROF_DOPPLER:
        case GX_PROF_GAUSSIAN:
            return launch_prof<GX_PROF_DOPPLER>(s, ppl, split, items, nitems, jobs, single, scale_bits, xq, xqb,
                                                mode, blocks);
        default:
            return hipErrorInvalidValue;
    }
}

hipError_t line_f64(hipStream_t s, const double* nu, const double* E, const double* A,
                    const int* g, const double* gamma, const double* nexp, int n,
                    const LinePos* pos, const TPParams* tp, LineF64* out,
                    unsigned long long* neval) {
    if (n <= 0) return hipSuccess;
    int gx = std::min(blocks_for(n, 256), 2048);
    hipLaunchKernelGGL(k_line_f64, dim3(gx), dim3(256), 0, s, nu, E, A, g, gamma, nexp, n, pos,
                       tp, out, neval);
    return hipGetLastError();
}

hipError_t fp64_points(hipStream_t s, int profile, const double* nu, const LinePos* pos,
                       const LineF64* lf, int n, double start, double dfreq, int win_lo,
                       int win_hi, double* out) {
    int np = win_hi - win_lo;
    if (np <= 0) return hipSuccess;
    dim3 grid(blocks_for(np, 256)), blk(256);
    switch (profile) {
        case GX_PROF_VOIGT:
            hipLaunchKernelGGL(k_fp64_points<GX_PROF_VOIGT>, grid, blk, 0, s, nu, pos, lf, n, start, dfreq, win_lo, win_hi, out);
            break;
        case GX_PROF_VOIGT_QUAD:
            hipLaunchKernelGGL(k_fp64_points<GX_PROF_VOIGT_QUAD>, grid, blk, 0, s, nu, pos, lf, n, start, dfreq, win_lo, win_hi, out);
            break;
        default:
            hipLaunchKernelGGL(k_fp64_points<GX_PROF_DOPPLER>, grid, blk, 0, s, nu, pos, lf, n, start, dfreq, win_lo, win_hi, out);
            break;
    }
    return hipGetLastError();
}

hipError_t partition(hipStream_t s, const double* E, const int* g, int64_t n, const double* dT,
                     int nT, double* partial, int nblk, double* dQ, const JobDesc* jd, int njobs,
                     TPParams* tp, unsigned long long* clear, int nclear) {
    if (nT <= 0) return hipSuccess;
    int64_t chunk = (n + nblk - 1) / nblk;
    if (chunk < 1) chunk = 1;
    hipLaunchKernelGGL(k_partition, dim3(nblk, nT), dim3(256), 0, s, E, g, n, chunk, dT, partial,
                       nblk);
    hipLaunchKernelGGL(k_partition_final, dim3(nT), dim3(256), 0, s, partial, nblk, dQ, jd, njobs, tp,
                       clear, nclear);
    return hipGetLastError();
}

hipError_t write_tables(hipStream_t s, const TPParams& tp, const JobDesc& jd, const Job& jb,
                        TPParams* dtp, JobDesc* djd, Job* djob, unsigned long long* clear,
                        int nclear) {
    hipLaunchKernelGGL(k_write_tables, dim3(1), dim3(64), 0, s, tp, jd, jb, dtp, djd, djob, clear,
                       nclear);
    return hipGetLastError();
}

hipError_t fill_u64(hipStream_t s, unsigned long long* p, int64_t n, unsigned long long v) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t fill_f64(hipStream_t s, double* p, int64_t n, double v) {
    if (n <= 0) return hipSuccess;
    int nb = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t add_f64(hipStream_t s, double* dst, const double* src, int64_t n) {
    if (n <= 0) return hipSuccess;
    int nb = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_add_f64, dim3(nb), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

}  // namespace gxk
