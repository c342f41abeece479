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
//                   once (no atomics on the grid).  In profile_tiles.h, its
//                   instantiations in tiles_voigt.hip, tiles_quad.hip and
//                   tiles_gauss.hip.
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
#include "profile_tiles.h"

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

// ------------------------------------------------------------- k_line_rec
// FP32 record of one line at one (T,P) (FP64 arithmetic, no contraction):
// returns the amplitude exponent ex (0 when the amplitude is 0) and sets
// *neval to the line's window length when S != 0 and the window is not empty;
// for the Voigt profiles x.k15 is the distance (in grid points) beyond which
// the line's points are in region I and *kfar the one beyond which the
// one-term wing form Kl / (s - 1) applies: region I is
// Kl (s + 1) / (s^2 + H) = Kl / (s - 1 + (H + 1) / (s + 1)), and the last
// term changes the value by at most 2^-23 relative once s^2 >= (H + 1) 2^23.
// For the bin-averaged Gaussian x.k15 = *kfar is the line's reach: at every
// point j with |c - j| >= reach both edges of the bin, e = (c - j) A + B - H
// and e + A, lie beyond the erfc table's ends (|e| >= 8) on one side, so the
// table gives them the same value and the line adds exactly zero there
// (0 when the line adds nothing anywhere).
__device__ __forceinline__ int line_record(const TPParams& p, bool voigt, double nu, double E,
                                           double A, int g, const double* gamma,
                                           const double* nexp, int64_t i, const LinePos& lp,
                                           VRec& r, VRecX& x, long long* neval, int* kfar) {
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
        const double sf = sqrt((2.0 * y * y + 1.0) * 8388608.0);  // s >= sf: one-term form
        const double xf2 = sf - (y * y - 0.5);
        const double kf = xf2 > 0.0 ? (sqrt(xf2) * (1.0 + 1e-5) + fabs(Bf)) / Af : 0.0;
        *kfar = max(x.k15, kf <= 0.0 ? 0 : (kf >= 1.0e8 ? 100000000 : (int)ceil(kf)));
    } else {
        r.Kl = 0.0f;
        r.E2 = 0.0f;
        r.H = 0.0f;
        x.y = (float)(0.5 * p.dfreq * iw);  // H, half bin width in units of the width
        // |c - j| >= K puts e = (c - j) A + B - H (and e + A) at or beyond
        // +8 (c - j >= K) or e + A at or below -8 (j - c >= K) when
        // K A >= 8 + |B| + A / 2; the relative margins cover the FP32
        // rounding of A, B - H and the kernel's fma (|e| <= 8 + A there)
        const double kk = ((8.0 * (1.0 + 1e-5) + fabs(Bf)) / Af) * (1.0 + 1e-5) + 0.5;
        x.k15 = amp > 0.0 ? (!(kk < 1.0e8) ? 100000000 : (int)ceil(kk)) : 0;
        *kfar = x.k15;
    }
    r.ex = ex;
    return amp > 0.0 ? ex : 0;
}

// the sums of one workgroup's lines, valid in thread 0
struct RecSums {
    long long ne;  // N_eval
    int km;        // largest x.k15
    int kf;        // largest kfar
    int em;        // largest biased exponent of a line reaching the grid (0: none)
};

// the records of job k's lines (the blocks of one grid row share them)
__device__ __forceinline__ RecSums line_rec_block(const TPParams& p, const JobDesc& d,
                                                  const double* __restrict__ nu,
                                                  const double* __restrict__ E,
                                                  const double* __restrict__ A,
                                                  const int* __restrict__ g,
                                                  const double* __restrict__ gamma,
                                                  const double* __restrict__ nexp,
                                                  const LinePos* __restrict__ pos,
                                                  VRec* __restrict__ rec,
                                                  VRecX* __restrict__ recx,
                                                  LinePos* __restrict__ pos_out = nullptr) {
    const bool voigt = p.profile == GX_PROF_VOIGT || p.profile == GX_PROF_VOIGT_QUAD;
    long long ne = 0;
    int km = 0, kf = 0;
    int em = 0;
    for (int ii = blockIdx.x * blockDim.x + threadIdx.x; ii < d.n; ii += gridDim.x * blockDim.x) {
        const int64_t i = d.l0 + ii;
        LinePos lp;
        if (pos_out != nullptr) {  // the windows too (k_line_pos's work, in the same pass)
            lp = window_of(p, nu[i]);
            pos_out[d.pos_off + ii] = lp;
        } else {
            lp = pos[d.pos_off + ii];
        }
        VRec r;
        VRecX x;
        long long nl;
        int kfl;
        line_record(p, voigt, nu[i], E[i], A[i], g[i], gamma, nexp, i, lp, r, x, &nl, &kfl);
        ne += nl;
        if (r.Kg > 0.0f) em = max(em, r.ex + kExBias);
        if (lp.ie > lp.ib) {
            if (voigt) km = max(km, x.k15);
            kf = max(kf, kfl);
        }
        rec[d.rec_off + ii] = r;
        recx[d.recx_off + ii] = x;
    }
    __shared__ long long sh_ne[4];
    __shared__ int sh_km[4];
    __shared__ int sh_kf[4];
    __shared__ int sh_em[4];
    return RecSums{block_sum_ll(ne, sh_ne), block_max_i(km, sh_km), block_max_i(kf, sh_kf),
                   block_max_i(em, sh_em)};
}

// adds one workgroup's (or launch's) sums to job k's counters (thread 0)
__device__ __forceinline__ void rec_publish(const RecSums& s, int k, unsigned long long* emax,
                                            int* kmax, unsigned long long* neval) {
    if (threadIdx.x != 0) return;
    if (neval != nullptr && s.ne != 0) atomicAdd(&neval[k], (unsigned long long)s.ne);
    if (kmax != nullptr && s.km > 0) atomicMax(&kmax[2 * k], s.km);
    if (kmax != nullptr && s.kf > 0) atomicMax(&kmax[2 * k + 1], s.kf);
    if (s.em > 0) atomicMax(&emax[k], (unsigned long long)s.em);
}

using gxk::kRecMaxBlocks;
using gxk::kRecPartBytes;

struct alignas(32) RecPart {
    long long ne;
    int km, kf, em;
};
static_assert(sizeof(RecPart) == kRecPartBytes, "RecPart layout");

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
                                                  unsigned long long* __restrict__ neval,
                                                  RecPart* __restrict__ parts,
                                                  LinePos* __restrict__ pos_out) {
    const int k = blockIdx.y;
    const RecSums t = line_rec_block(tp[k], jd[k], nu, E, A, g, gamma, nexp, pos, rec, recx, pos_out);
    if (parts == nullptr)
        rec_publish(t, k, emax, kmax, neval);
    else if (threadIdx.x == 0)
        parts[(size_t)k * kRecMaxBlocks + blockIdx.x] = RecPart{t.ne, t.km, t.kf, t.em};
}

// One pass of one job with its tables as kernel arguments: block 0 also
// stores the Job the profile kernel reads (djob) and clears clear[0 ..
// nclear), the counters of a later pass (the pass's own counters are zero
// before it starts).
__global__ __launch_bounds__(256) void k_line_rec_pass(TPParams p, JobDesc d, Job jb,
                                                       Job* __restrict__ djob,
                                                       unsigned long long* __restrict__ clear,
                                                       int nclear, const double* __restrict__ nu,
                                                       const double* __restrict__ E,
                                                       const double* __restrict__ A,
                                                       const int* __restrict__ g,
                                                       const double* __restrict__ gamma,
                                                       const double* __restrict__ nexp,
                                                       const LinePos* __restrict__ pos,
                                                       VRec* __restrict__ rec,
                                                       VRecX* __restrict__ recx,
                                                       RecPart* __restrict__ parts) {
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) *djob = jb;
        for (int i = threadIdx.x; i < nclear; i += blockDim.x) clear[i] = 0ull;
    }
    const RecSums t = line_rec_block(p, d, nu, E, A, g, gamma, nexp, pos, rec, recx);
    if (threadIdx.x == 0) parts[blockIdx.x] = RecPart{t.ne, t.km, t.kf, t.em};
}

// Adds the workgroup sums of a line-record launch (nb per job, job k's at
// parts + k kRecMaxBlocks) to job k's counters: one workgroup per job, four
// atomics.  The launch's own workgroups do not combine their sums:
// device-scope atomics on one address serialise, and an agent-scope release
// writes back the XCD's L2, so either, once per workgroup, made a
// 2048-workgroup launch several times longer than its loads and stores.
__device__ __forceinline__ void rec_reduce_job(const RecPart* __restrict__ parts, int nb, int k,
                                               unsigned long long* __restrict__ emax,
                                               int* __restrict__ kmax,
                                               unsigned long long* __restrict__ neval) {
    parts += (size_t)k * kRecMaxBlocks;
    RecSums v{0, 0, 0, 0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        const RecPart q = parts[i];
        v.ne += q.ne;
        v.km = max(v.km, q.km);
        v.kf = max(v.kf, q.kf);
        v.em = max(v.em, q.em);
    }
    __shared__ long long sh_ne[4];
    __shared__ int sh_km[4];
    __shared__ int sh_kf[4];
    __shared__ int sh_em[4];
    const RecSums t{block_sum_ll(v.ne, sh_ne), block_max_i(v.km, sh_km), block_max_i(v.kf, sh_kf),
                    block_max_i(v.em, sh_em)};
    rec_publish(t, k, emax, kmax, neval);
}

__global__ __launch_bounds__(256) void k_rec_reduce(const RecPart* __restrict__ parts, int nb,
                                                    unsigned long long* __restrict__ emax,
                                                    int* __restrict__ kmax,
                                                    unsigned long long* __restrict__ neval) {
    rec_reduce_job(parts, nb, blockIdx.x, emax, kmax, neval);
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
                                               uint32_t* __restrict__ item_idx, gxk::RecReduce red) {
    const int job = job0 + blockIdx.y;
    if (red.parts != nullptr && blockIdx.x == gridDim.x - 1) {  // the job's line-record sums
        rec_reduce_job((const RecPart*)red.parts, red.nb, job, red.emax, red.kmax, red.neval);
        return;
    }
    const JobDesc d = jd[job];
    const int lane = threadIdx.x & (GX_WAVE - 1);
    const int t = blockIdx.x * (blockDim.x / GX_WAVE) + (threadIdx.x >> 6);  // one wave per tile
    if (t >= d.ntiles) return;
    const int win_lo = tp[job].win_lo, win_hi = tp[job].win_hi;
    const LinePos* jp = pos + d.pos_off;
    const int n = d.n;
    const int w0 = win_lo + t * span;
    const int w1 = min(w0 + span, win_hi);
    // lo = first i with ie > w0, hi = first with ib >= w1; the lines whose
    // window covers the whole tile are [lf, rf) within [lo, hi): lf = first
    // with ie >= w1, rf = first with ib > w0.  ib and ie are non-decreasing
    // in nu, and every line before lo has ib <= max(ie, 0) <= w0 < w1, so each
    // bound is also the first over all [0, n), clamped to hi for lf and rf.
    // The wave searches 64-way: each round the lanes test 64 evenly spaced
    // lines of a bound's range [a, b] (the bound is b when no line in [a, b)
    // qualifies) and the ballot keeps the stretch before the first that
    // does, so 5e5 lines take 4-5 rounds; the four bounds share the rounds
    // (one thread per tile, bisecting, took 14-30 us per launch).
    int a[4] = {0, 0, 0, 0}, b[4] = {n, n, n, n};
    while (a[0] < b[0] || a[1] < b[1] || a[2] < b[2] || a[3] < b[3]) {
        int st[4], i[4];
        LinePos q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            st[k] = max((b[k] - a[k] + GX_WAVE - 1) / GX_WAVE, 1);
            i[k] = a[k] + lane * st[k];
            q[k] = jp[min(i[k], max(b[k] - 1, 0))];  // n >= 1 here
        }
        const bool up[4] = {q[0].ie > w0, q[1].ib >= w1, q[2].ie >= w1, q[3].ib > w0};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (a[k] >= b[k]) continue;  // wave-uniform
            const unsigned long long m = __ballot(i[k] < b[k] && up[k]);
            if (m == 0ull) {
                a[k] += ((b[k] - 1 - a[k]) / st[k]) * st[k] + 1;  // past the last sample
            } else {
                const int f = __builtin_ctzll(m);
                b[k] = a[k] + f * st[k];
                if (f > 0) a[k] += (f - 1) * st[k] + 1;
            }
        }
    }
    const int lo = a[0];
    const int hi = a[1];
    const int lf = min(a[2], hi);
    const int rf = min(a[3], hi);
    WorkItem it;
    it.lo = lo;
    it.hi = hi;
    it.w0 = w0;
    it.w1 = w1;
    it.job = job;
    it.s1 = lf < rf ? lf : hi;
    it.s4 = lf < rf ? rf : hi;
    it.s2 = it.s3 = it.s4;  // the profile kernel locates s2, s3, f2, f3 per pass
    const bool gauss = tp[job].profile == GX_PROF_DOPPLER || tp[job].profile == GX_PROF_GAUSSIAN;
    it.f2 = gauss ? lo : it.s1;  // bin-Gauss: every line within reach until then
    it.f3 = gauss ? hi : it.s4;
    it.e2 = it.s1;
    it.e3 = it.s4;
    const int G = max(d.groups, 1);
    it.ngroups = G;
    if (lane != 0) return;
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

// The segments of every work item as the profile kernel locates them
// (locate_segments, one wave per item), stored into the item: for
// inspection only, the profile kernel locates them itself.  pos: the lines of
// every job (pos_off 0).
__global__ __launch_bounds__(256) void k_segments(WorkItem* items, int nitems,
                                                  const LinePos* __restrict__ pos,
                                                  const int* __restrict__ kmax, bool gauss,
                                                  int half) {
    const int t = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & (GX_WAVE - 1);
    if (t >= nitems) return;
    WorkItem it = items[t];
    if (gauss)
        locate_segments<true>(it, pos, kmax, half, lane);
    else
        locate_segments<false>(it, pos, kmax, half, lane);
    if (lane == 0) items[t] = it;
}

__global__ __launch_bounds__(256) void k_permute_items(const WorkItem* __restrict__ in,
                                                       const uint32_t* __restrict__ idx, int n,
                                                       WorkItem* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = in[idx[t]];
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

int rec_blocks(int64_t max_n, int max_blocks) {
    const int cap = max_blocks > 0 ? std::min(max_blocks, kRecMaxBlocks) : kRecMaxBlocks;
    return std::min(blocks_for(std::max<int64_t>(max_n, 1), 256), cap);
}

hipError_t line_rec(hipStream_t s, const double* nu, const double* E, const double* A,
                    const int* g, const double* gamma, const double* nexp, const LinePos* pos,
                    const JobDesc* jd, int njobs, int64_t max_n, const TPParams* tp,
                    unsigned long long* emax, VRec* rec, VRecX* recx, int* kmax,
                    unsigned long long* neval, int max_blocks, void* parts, LinePos* pos_out,
                    bool reduce) {
    if (njobs <= 0 || max_n <= 0) return hipSuccess;
    const int gx = rec_blocks(max_n, max_blocks);
    hipLaunchKernelGGL(k_line_rec, dim3(gx, njobs), dim3(256), 0, s, nu, E, A, g, gamma, nexp, pos,
                       jd, tp, emax, rec, recx, kmax, neval, (RecPart*)parts, pos_out);
    if (parts != nullptr && reduce)
        hipLaunchKernelGGL(k_rec_reduce, dim3(njobs), dim3(256), 0, s, (const RecPart*)parts, gx, emax,
                           kmax, neval);
    return hipGetLastError();
}

hipError_t line_rec_pass(hipStream_t s, const TPParams& tp, const JobDesc& jd, const Job& jb,
                         Job* djob, unsigned long long* clear, int nclear, const double* nu,
                         const double* E, const double* A, const int* g, const double* gamma,
                         const double* nexp, const LinePos* pos, unsigned long long* emax,
                         VRec* rec, VRecX* recx, int* kmax, unsigned long long* neval,
                         void* red, int max_blocks) {
    if (red == nullptr) return hipErrorInvalidValue;
    const int cap = max_blocks > 0 ? std::min(max_blocks, kRecMaxBlocks) : kRecMaxBlocks;
    const int gx = std::max(1, std::min(blocks_for(std::max<int64_t>(jd.n, 1), 256), cap));
    hipLaunchKernelGGL(k_line_rec_pass, dim3(gx), dim3(256), 0, s, tp, jd, jb, djob, clear, nclear,
                       nu, E, A, g, gamma, nexp, pos, rec, recx, (RecPart*)red);
    hipLaunchKernelGGL(k_rec_reduce, dim3(1), dim3(256), 0, s, (const RecPart*)red, gx, emax, kmax,
                       neval);
    return hipGetLastError();
}

hipError_t segments(hipStream_t s, WorkItem* items, int nitems, const LinePos* pos,
                    const int* kmax, int profile, int span) {
    if (nitems <= 0) return hipSuccess;
    const bool gauss = profile != GX_PROF_VOIGT && profile != GX_PROF_VOIGT_QUAD;
    hipLaunchKernelGGL(k_segments, dim3(blocks_for(nitems, 4)), dim3(256), 0, s, items, nitems, pos,
                       kmax, gauss, span / 2);
    return hipGetLastError();
}

hipError_t tiles(hipStream_t s, const LinePos* pos, const JobDesc* jd, int njobs, int max_tiles,
                 const TPParams* tp, int span, WorkItem* items, uint32_t* cost_key,
                 uint32_t* item_idx, int job0, const RecReduce* red) {
    if (njobs <= 0 || (max_tiles <= 0 && red == nullptr)) return hipSuccess;
    const RecReduce r = red ? *red : RecReduce{nullptr, 0, nullptr, nullptr, nullptr};
    const int gx = blocks_for(std::max(max_tiles, 0), 256 / GX_WAVE) + (red ? 1 : 0);  // a wave per tile
    hipLaunchKernelGGL(k_tiles, dim3(gx, njobs), dim3(256), 0, s, pos, jd, tp, span, job0, items,
                       cost_key, item_idx, r);
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

hipError_t profile_tiles(hipStream_t s, int profile, int ppl, int split, const WorkItem* items,
                         int nitems, const Job* jobs, const unsigned long long* scale_bits,
                         const int* kmax, unsigned int* xq, const unsigned int* xqb, int mode,
                         int blocks, bool single_job) {
    const int single = single_job ? 1 : 0;
    switch (profile) {
        case GX_PROF_VOIGT:
            return profile_tiles_voigt(s, ppl, split, items, nitems, jobs, single, scale_bits, kmax, xq,
                                       xqb, mode, blocks);
        case GX_PROF_VOIGT_QUAD:
            return profile_tiles_quad(s, ppl, split, items, nitems, jobs, single, scale_bits, kmax, xq,
                                      xqb, mode, blocks);
        case GX_PROF_DOPPLER:
        case GX_PROF_GAUSSIAN:
            return profile_tiles_gauss(s, ppl, split, items, nitems, jobs, single, scale_bits, kmax, xq,
                                       xqb, mode, blocks);
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
