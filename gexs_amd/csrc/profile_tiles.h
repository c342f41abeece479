// profile_tiles.h -- the profile kernel k_profile_tiles<PROFILE, PPL, SPLIT>
// (a tile of 64*PPL grid points per wave or per workgroup of SPLIT waves; the
// tile's nu-sorted lines staged per 64-line batch into LDS rows; FP32 sums
// with per-point Kahan compensation in registers; each point written once)
// and its launch templates.  Each profile's instantiations are compiled in a
// translation unit of their own (tiles_voigt.hip, tiles_quad.hip,
// tiles_gauss.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "gexs_internal.h"
#include "kernels.h"
#include "profiles_device.h"

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

// For each n, the first index i of [a, b) with jp[i].c >= v[n] (b when there
// is none); c is non-decreasing in i.  The whole wave searches N bounds at
// once, 64 ways per round: lane k samples the last line of the k-th of 64
// equal chunks (the last lane line b - 1), and the first chunk whose sample
// reaches v holds the answer (1 round for 64 lines, 2 for 4096, 3 for 2^18).
template <int N>
__device__ __forceinline__ void wave_first_c(const LinePos* jp, int a0, int b0,
                                             const long long (&v)[N], int (&out)[N], int lane) {
    int a[N], b[N];
    bool open = false;
#pragma unroll
    for (int n = 0; n < N; ++n) {
        a[n] = a0;
        b[n] = b0;
        out[n] = a0 < b0 ? -1 : b0;
        open = open || a0 < b0;
    }
    while (open) {
        int c[N];
#pragma unroll
        for (int n = 0; n < N; ++n) {  // the N samples in flight together
            const int s = (b[n] - a[n] + GX_WAVE - 1) / GX_WAVE;
            c[n] = out[n] < 0 ? jp[min(a[n] + (lane + 1) * s - 1, b[n] - 1)].c : 0;
        }
        open = false;
#pragma unroll
        for (int n = 0; n < N; ++n) {
            if (out[n] >= 0) continue;
            const int s = (b[n] - a[n] + GX_WAVE - 1) / GX_WAVE;
            const unsigned long long m = __ballot((long long)c[n] >= v[n]);
            if (m == 0ull) {  // b - 1 too is below v
                out[n] = b[n];
                continue;
            }
            const int k = __builtin_ctzll(m);
            if (s == 1) {
                out[n] = a[n] + k;
                continue;
            }
            b[n] = min(a[n] + (k + 1) * s - 1, b[n] - 1) + 1;  // its sample reaches v
            a[n] = a[n] + k * s;                               // chunk k - 1's does not
            open = true;
        }
    }
}

// The pass-dependent segments of a work item (kmax[2 job] = K, the job's
// largest k15, and kmax[2 job + 1] = KF, its largest kfar or reach, from this
// pass's line records):
//   Voigt profiles: among the lines [s1, s4) that cover the whole tile, f2,
//   s2, s3 and f3 are the first with c >= w0 - KF, w0 - K, w1 + K, w1 + KF:
//   [s2, s3) may have points outside region I, [s1, f2) and [f3, s4) take
//   the one-term wing form at every point.  e2 and e3 are the first with
//   c >= w0 - max(KF, 3T + 2) and c >= max(w1 + KF, w0 + 5T + 2), T the
//   tile's half width: [s1, e2) and [e3, s4) are far wings whose centre lies
//   at least 4T + 2.5 points from the tile's centre w0 + T - 1/2 (the
//   far-field expansion's lines).
//   bin-Gauss: [f2, f3) are the lines of [lo, hi) with c in [w0 - KF,
//   w1 + KF), the only ones that add to the tile.
// Every lane of the wave gets the same values.
template <bool GAUSS>
__device__ __forceinline__ void locate_segments(WorkItem& it, const LinePos* jp, const int* kmax,
                                                int half, int lane) {
    const long long KF = kmax[2 * it.job + 1];
    if constexpr (GAUSS) {
        const long long v[2] = {(long long)it.w0 - KF, (long long)it.w1 + KF};
        int r[2];
        wave_first_c<2>(jp, it.lo, max(it.hi, it.lo), v, r, lane);
        it.f2 = r[0];
        it.f3 = r[1];
    } else {
        const long long K = kmax[2 * it.job];
        const long long T = half;
        const long long v[6] = {(long long)it.w0 - KF, (long long)it.w0 - K, (long long)it.w1 + K,
                                (long long)it.w1 + KF, (long long)it.w0 - max(KF, 3 * T + 2),
                                max((long long)it.w1 + KF, (long long)it.w0 + 5 * T + 2)};
        int r[6];
        wave_first_c<6>(jp, it.s1, max(it.s4, it.s1), v, r, lane);
        it.f2 = r[0];
        it.s2 = r[1];
        it.s3 = r[2];
        it.f3 = r[3];
        it.e2 = r[4];
        it.e3 = r[5];
    }
}

// --------------------------------------------------------- profile tiles
namespace {

// The bin-averaged Gaussian's erfc: the table of cubic pieces in device
// memory (filled once per device by gxk::init_gauss_tables) and its copy in
// the LDS of each workgroup of the Gaussian profile kernel.
static __device__ float4 gx_erfc_pieces[gx::kErfcPieces];
static __shared__ float4 gx_erfc_lds[gx::kErfcPieces];

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

// The one-term wing value Kl / (s - 1), s - 1 = x^2 + (E2 - 1), of a line
// whose whole tile lies beyond its kfar: region I without the term that is
// below 2^-23 there (4 VALU against 6).
__device__ __forceinline__ float far1(float x, float Kl, float E2m1) {
    return Kl * __builtin_amdgcn_rcpf(fmaf(x, x, E2m1));
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

template <int PPL, int MASK, bool FAR>
__device__ __forceinline__ void staged_masked(TileAcc<PPL>& t, const float* row,
                                              const LinePos& lp, float nl, int w0) {
    const float A = row[0], Kl = row[1], E2 = row[2], H = row[3];
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        const int st = slice_in_window<MASK>(w0 + GX_WAVE * q, lp);
        if (st == 0) continue;
        const float x = fmaf(nl, A, row[4 + q]);
        const float v = FAR ? far1(x, Kl, E2) : region1(x, Kl, E2, H);
        t.part[q] += (st == 2 || in_window<MASK>(t.jj[q], lp)) ? v : 0.0f;
    }
    t.tick(1);
}

// FAR: the lines' whole tile lies beyond their kfar; the row holds E2 - 1 in
// place of E2 and the one-term form is evaluated.
template <int PPL, int MASK, bool FAR = false>
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
        row[2] = FAR ? vc.E2 - 1.0f : vc.E2;
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
                    staged_masked<PPL, MASK, FAR>(t, xs + (k + 0) * R, p0, nl, w0);
                    staged_masked<PPL, MASK, FAR>(t, xs + (k + 1) * R, p1, nl, w0);
                    staged_masked<PPL, MASK, FAR>(t, xs + (k + 2) * R, p2, nl, w0);
                    staged_masked<PPL, MASK, FAR>(t, xs + (k + 3) * R, p3, nl, w0);
                    p0 = q0;
                    p1 = q1;
                    p2 = q2;
                    p3 = q3;
                }
                staged_masked<PPL, MASK, FAR>(t, xs + (k + 0) * R, p0, nl, w0);
                staged_masked<PPL, MASK, FAR>(t, xs + (k + 1) * R, p1, nl, w0);
                staged_masked<PPL, MASK, FAR>(t, xs + (k + 2) * R, p2, nl, w0);
                staged_masked<PPL, MASK, FAR>(t, xs + (k + 3) * R, p3, nl, w0);
                k += 4;
            }
            for (; k < nb; ++k) staged_masked<PPL, MASK, FAR>(t, xs + k * R, pb[k], nl, w0);
            continue;
        }
#pragma unroll 4
        for (int k = 0; k < nb; ++k) {
            const float* rk = xs + k * R;
            const float A = rk[0], Kl = rk[1], E2 = rk[2], H = rk[3];
#pragma unroll
            for (int q = 0; q < PPL; ++q) {
                const float x = fmaf(nl, A, rk[4 + q]);
                t.part[q] += FAR ? far1(x, Kl, E2) : region1(x, Kl, E2, H);
            }
        }
        t.fold();
    }
}

// Far-field expansion of the one-term wing form over a tile of half width T
// (T = 32 PPL).  A far line's value at point j, Kl / (x_j^2 + D), with
// x_j = Xc - A T tau (Xc its x at the tile's centre t_c = w0 + T - 1/2,
// tau = (j - t_c) / T in [-1, 1], D = E2 - 1), is the power series
// sum_k h_k tau^k with, for den = Xc^2 + D,
//   h_0 = Kl / den,  h_1 = a h_0,  h_k = a h_(k-1) + b h_(k-2),
//   a = 2 Xc A T / den,  b = -(A T)^2 / den
// (multiply the series by den - 2 Xc A T tau + (A T tau)^2 = x^2 + D and
// compare terms).  The expansion's lines lie 4T + 2.5 points or more from
// t_c, so A T / |Xc| < 1/4 (|B| <= A / 2); they are far lines, whose x^2
// exceeds 2800 at every point, so |D| <= 1.5 moves the series' poles by
// under 2.3 %; the terms beyond kFarTerms then add under 17 (0.256)^16 /
// 0.744^2 = 1.1e-8 of h_0.  Each lane sums the coefficients of one line per
// batch of 64 (67 VALU per batch: about one per line per wave, against 4 per
// line per 64-point slice for the one-term form); far_add adds the lanes'
// sums across the wave and evaluates the polynomial at each point.
constexpr int kFarTerms = 16;

template <int PPL>
__device__ __forceinline__ void far_expand(float (&C)[kFarTerms], GlobRec grec, int a, int b,
                                           float tcf, int E, int lane) {
    constexpr float T = (float)(GX_WAVE * PPL / 2);
    if (a >= b) return;
    // the record's first 24 bytes (cf, A, B, Kl | E2, ex): 16 + 8 byte loads
    struct Far {
        float cf, A, B, Kl, E2;
        int ex;
    };
    auto load = [&](int i) {
        const VRec* r = &grec[i];
        return Far{r->cf, r->A, r->B, r->Kl, r->E2, r->ex};
    };
    Far vn = load(min(a + lane, b - 1));
    for (int base = a; base < b; base += GX_WAVE) {
        const Far v = vn;
        vn = load(min(base + GX_WAVE + lane, b - 1));
        const float Kl = base + lane < b ? ldexpf(v.Kl, v.ex - E) : 0.0f;
        const float Xc = fmaf(v.cf - tcf, v.A, v.B);
        const float inv = __builtin_amdgcn_rcpf(fmaf(Xc, Xc, v.E2 - 1.0f));
        const float u = v.A * T;
        const float al = (2.0f * u) * (Xc * inv);
        const float be = -(u * u) * inv;
        float h0 = Kl * inv;
        float h1 = al * h0;
        C[0] += h0;
        C[1] += h1;
#pragma unroll
        for (int k = 2; k < kFarTerms; ++k) {
            const float h2 = fmaf(al, h1, be * h0);
            C[k] += h2;
            h0 = h1;
            h1 = h2;
        }
    }
}

// the wave's coefficient sums (an xor butterfly: every lane adds the same
// pairs, so all hold the same totals) evaluated at the lane's points and
// added to their sums
template <int PPL>
__device__ __forceinline__ void far_add(TileAcc<PPL>& t, float (&C)[kFarTerms], float tcf) {
    constexpr float invT = 1.0f / (float)(GX_WAVE * PPL / 2);
#pragma unroll
    for (int k = 0; k < kFarTerms; ++k) {
#pragma unroll
        for (int m = GX_WAVE / 2; m >= 1; m >>= 1) C[k] += __shfl_xor(C[k], m);
    }
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        const float tau = (t.jf[q] - tcf) * invT;  // exact
        float p = C[kFarTerms - 1];
#pragma unroll
        for (int k = kFarTerms - 2; k >= 0; --k) p = fmaf(p, tau, C[k]);
        kahan_add(t.sum[q], t.cmp[q], p);
    }
}

// the far-field expansion of the lines [a, b) and [c, d) added to the tile
template <int PPL>
__device__ __forceinline__ void far_lines(TileAcc<PPL>& t, GlobRec grec, int a, int b, int c, int d,
                                          int w0, int lane) {
    if (a >= b && c >= d) return;
    const float tcf = (float)w0 + (float)(GX_WAVE * PPL / 2) - 0.5f;  // exact below 2^23
    float C[kFarTerms];
#pragma unroll
    for (int k = 0; k < kFarTerms; ++k) C[k] = 0.0f;
    far_expand<PPL>(C, grec, a, b, tcf, t.E, lane);
    far_expand<PPL>(C, grec, c, d, tcf, t.E, lane);
    far_add<PPL>(t, C, tcf);
}

// Window-edge lines by the far-field expansion.  HI: the lines [a, b) whose
// windows end inside the tile (ie in (w0, w1], non-decreasing in nu order):
// point j takes the lines with ie > j, a suffix of [a, b).  LO: the lines
// whose windows start inside it (ib in [w0, w1), non-decreasing): point j
// takes those with ib <= j, a prefix.  The window bounds go to the wave's LDS
// block; each point finds its first line (suffix) or its end (prefix) by
// binary search there.  The lines' coefficients (kEdgeTerms terms, lane k one
// line per batch) are scanned across the lanes, batch after batch with the
// sum so far carried, in the order suffix or prefix sums need, and each point
// takes the scanned sum of its line when that line is in the batch.  The
// lines must lie 14T + 1 points or more from the tile's centre (ratio
// A T / |Xc| <= 0.072, so 8 terms leave under 9 (0.074)^8 / 0.926^2 = 1e-8);
// the caller checks nothing else: false (nothing added) when the nearest
// line is closer or the lines outnumber the LDS block, and the caller sums
// them directly.
constexpr int kEdgeTerms = 8;

template <int PPL, bool HI>
__device__ __forceinline__ bool edge_expand(TileAcc<PPL>& t, GlobPos gpos, GlobRec grec, int a,
                                            int b, int w0, int lane, float* xs, int cap) {
    constexpr int T = GX_WAVE * PPL / 2;
    if (a >= b) return true;
    const int n = b - a;
    if (n > cap) return false;
    const float tcf = (float)w0 + (float)T - 0.5f;
    const int cnear = __builtin_amdgcn_readfirstlane(gpos[HI ? b - 1 : a].c);
    const float dist = HI ? tcf - (float)cnear : (float)cnear - tcf;
    if (dist < (float)(14 * T + 1)) return false;
    int* E = reinterpret_cast<int*>(xs);
    for (int i = lane; i < n; i += GX_WAVE) E[i] = HI ? gpos[a + i].ie : gpos[a + i].ib;
    // m[q]: the number of lines with bound <= j, i.e. the first with bound
    // > j (E is non-decreasing); a fixed number of halvings (n's bit length)
    int m[PPL];
    int steps = 0;
    while ((1 << steps) <= n) ++steps;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        const int j = t.jj[q];
        int lo = 0, len = n;  // the answer lies in [lo, lo + len]
        for (int st = 0; st < steps; ++st) {
            const int half = len >> 1;
            const bool go = len > 0 && E[min(lo + half, n - 1)] <= j;
            lo = go ? lo + half + 1 : lo;
            len = len > 0 ? (go ? len - half - 1 : half) : 0;
        }
        m[q] = lo;
    }
    float carry[kEdgeTerms];
#pragma unroll
    for (int k = 0; k < kEdgeTerms; ++k) carry[k] = 0.0f;
    for (int done = 0; done < n; done += GX_WAVE) {
        const int nb = min(GX_WAVE, n - done);
        // lane k's line: HI from the top down, LO from the bottom up
        const int li = HI ? n - 1 - done - lane : done + lane;
        const bool act = lane < nb;
        const VRec* r = &grec[a + min(max(li, 0), n - 1)];
        const float cf = r->cf, A = r->A, B = r->B, Klm = r->Kl, E2 = r->E2;
        const int ex = r->ex;
        const float Kl = act ? ldexpf(Klm, ex - t.E) : 0.0f;
        const float Xc = fmaf(cf - tcf, A, B);
        const float inv = __builtin_amdgcn_rcpf(fmaf(Xc, Xc, E2 - 1.0f));
        const float u = A * (float)T;
        const float al = (2.0f * u) * (Xc * inv);
        const float be = -(u * u) * inv;
        float c[kEdgeTerms];
        c[0] = Kl * inv;
        c[1] = al * c[0];
#pragma unroll
        for (int k = 2; k < kEdgeTerms; ++k) c[k] = fmaf(al, c[k - 1], be * c[k - 2]);
        // inclusive prefix sums over the lanes, plus the lines before
#pragma unroll
        for (int k = 0; k < kEdgeTerms; ++k) {
#pragma unroll
            for (int d = 1; d < GX_WAVE; d <<= 1) {
                const float o = __shfl_up(c[k], d);
                c[k] += lane >= d ? o : 0.0f;
            }
            c[k] += carry[k];
        }
        // each point: its sum when its line is in this batch
#pragma unroll
        for (int q = 0; q < PPL; ++q) {
            // HI: suffix from line m (lane n - 1 - done - m); LO: prefix to line m - 1
            const int want = HI ? n - 1 - done - m[q] : m[q] - 1 - done;
            const bool here = want >= 0 && want < nb;
            float p = __shfl(c[kEdgeTerms - 1], want & (GX_WAVE - 1));
            const float tau = (t.jf[q] - tcf) * (1.0f / (float)T);
#pragma unroll
            for (int k = kEdgeTerms - 2; k >= 0; --k) p = fmaf(p, tau, __shfl(c[k], want & (GX_WAVE - 1)));
            if (here) kahan_add(t.sum[q], t.cmp[q], p);
        }
#pragma unroll
        for (int k = 0; k < kEdgeTerms; ++k)
            carry[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c[k]), nb - 1));
    }
    return true;
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
            const int j0 = w0 + GX_WAVE * q;
            const int st = slice_in_window<MASK>(j0, lp);
            if (st == 0) continue;
            const float x = fabsf(fmaf(r.cf - t.jf[q], r.A, r.B));
            float v = region1(x, r.Kl, r.E2, r.H);
            // the slice's nearest point to the centre is k15 or more away:
            // region I at every lane (a scalar decision per slice)
            const int kq = lp.c < j0 ? (j0 - lp.c) : (lp.c >= j0 + GX_WAVE ? (lp.c - j0 - GX_WAVE + 1) : 0);
            if (kq < rx.k15) {
                if (inner_point<PROF>(x, rx.y)) v = DEFER ? 0.0f : r.Kg * inner_value<PROF>(x, rx.y);
            }
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
// the region-I loop.  Point j's bin spans e from x_j + h down to x_j - h (x
// falls by A = 2h per point), and its value is
//   erf(x_j + h) - erf(x_j - h) = H(x_j - h) - H(x_j + h),  H = erfc,
// so the bins of neighbouring points share an edge: each lane evaluates H
// once, at its far edge e_j = x_j - h, and takes the near edge from the lane
// before (DPP wave shift by one).  Lane 0's near edge in slice q is the far
// edge of point w0 + 64q - 1: per batch of 64 lines, lane k writes row k
// (cf, A, B - h, Kg | ib, ie, H_0, H_1 | H_2 ...), H_q from the very
// expression lane 63 of slice q - 1 (or of the tile before) evaluates, so
// each shared edge is one FP32 value and a line's bins telescope exactly: a
// line's band strength is H(last edge) - H(first edge) up to the roundings of
// the bins.  H comes from the table of cubic pieces in LDS (erfc_lds): ~8
// VALU against ~15 for a rational-exponential erfc.  Below the centre H is
// near 2, where FP32 spacing is 2.4e-7: a bin there carries at most that
// much absolute error, in units where the line's bins sum to 2.
template <int PPL>
constexpr int gauss_row() { return (6 + PPL + 3) / 4 * 4; }

__device__ __forceinline__ float wave_shr1(float v, float lane0) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lane0), __float_as_int(v),
                                                      0x138 /* wave_shr:1 */, 0xF, 0xF, false));
}

// erfc(e) from the workgroup's LDS copy of the table, given v = 32 e (the
// callers stage A and B - h times 32, a power of two, so v is exactly 32
// times the edge they would compute unscaled): piece k covers e in
// [k / 32 - 8, (k + 1) / 32 - 8) as a0 + f (a1 + f (a2 + f a3)), f the
// position within the piece (exact); e outside [-8, 8) takes the end pieces
// (erfc(8) = 1.1e-29).
__device__ __forceinline__ float erfc_lds(float v32) {
    constexpr float half = 0.5f * (float)gx::kErfcPieces;
    const float v = __builtin_amdgcn_fmed3f(v32, -half, half - 1.5e-5f);
    int k;  // floor(v) as an integer in one instruction (v is finite after the clamp)
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(k) : "v"(v));
    const float f = __builtin_amdgcn_fractf(v);
    const float4 c = gx_erfc_lds[k + gx::kErfcPieces / 2];
    return fmaf(fmaf(fmaf(c.w, f, c.z), f, c.y), f, c.x);
}

// lines [a, b) of a tile at 64 * PPL points.  MASK names the window edges
// that may fall inside the tile (the segments of the work item):
// kMaskNone, every window covers the tile; kMaskHi / kMaskLo, the windows
// end / start inside it; kMaskBoth, either.  A masked point adds nothing
// outside its line's window.  At one point per lane every line of a tile
// meets its one slice; with more slices a masked line skips the slices
// outside its window (a scalar decision per slice).  The partial sums are
// folded once per chunk of kChunk lines.
template <int PPL, int MASK>
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
        // A and the far edge at c, B - h, times the table's 32 per unit
        const float A32 = rc.A * gx::kErfcPerUnit;
        const float Bh32 = (rc.B - xc.y) * gx::kErfcPerUnit;
        row[0] = rc.cf;
        row[1] = A32;
        row[2] = Bh32;
        row[3] = ldexpf(rc.Kg, rc.ex - t.E);
        row[4] = __int_as_float(pc.ib);
        row[5] = __int_as_float(pc.ie);
#pragma unroll
        for (int q = 0; q < PPL; ++q)
            row[6 + q] = erfc_lds(fmaf(rc.cf - (float)(w0 + GX_WAVE * q - 1), A32, Bh32));
        const int nb = min(GX_WAVE, b - base);
        if (t.n > 0) t.fold();
        // one line of the staged batch (row k) into the partial sums
        auto line = [&](const float* rk) {
            const float cf = rk[0], A32 = rk[1], Bh32 = rk[2], Kg = rk[3];
            LinePos lp;
            if (MASK != kMaskNone) {
                lp.ib = __float_as_int(rk[4]);
                lp.ie = __float_as_int(rk[5]);
            }
#pragma unroll
            for (int q = 0; q < PPL; ++q) {
                if (PPL > 1 && MASK != kMaskNone) {
                    LinePos ls;
                    ls.ib = __builtin_amdgcn_readfirstlane(lp.ib);
                    ls.ie = __builtin_amdgcn_readfirstlane(lp.ie);
                    if (slice_in_window<MASK>(w0 + GX_WAVE * q, ls) == 0) continue;
                }
                const float Hb = erfc_lds(fmaf(cf - t.jf[q], A32, Bh32));
                const float Ha = wave_shr1(Hb, rk[6 + q]);
                t.part[q] = fmaf(Kg, in_window<MASK>(t.jj[q], lp) ? Hb - Ha : 0.0f, t.part[q]);
            }
        };
        for (int k0 = 0; k0 < nb; k0 += kChunk) {
            const int k1 = min(nb, k0 + kChunk);
            for (int k = k0; k < k1; ++k) line(xs + k * R);
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
        const GlobPos gpos = (GlobPos)jb.pos;
        const GlobRec grec = (GlobRec)jb.rec;
        const GlobRecX grecx = (GlobRecX)jb.recx;
        if (it.s1 < it.s4) {
            // [lo, s1): windows end inside the tile; [s1, s4): cover it;
            // [s4, hi): start inside it
            const int s1 = min(max(it.s1, ca), cb), s4 = min(max(it.s4, ca), cb);
            gauss_staged<PPL, kMaskHi>(t, gpos, grec, grecx, ca, s1, w0, lane, xs);
            gauss_staged<PPL, kMaskNone>(t, gpos, grec, grecx, s1, s4, w0, lane, xs);
            gauss_staged<PPL, kMaskLo>(t, gpos, grec, grecx, s4, cb, w0, lane, xs);
        } else {
            gauss_staged<PPL, kMaskBoth>(t, gpos, grec, grecx, ca, cb, w0, lane, xs);
        }
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
            // the region-I and far segments through the staged loop (6 or 4
            // VALU per line against 14 with broadcasts), the rest batched
            const int f2 = min(max(it.f2, ca), cb), f3 = min(max(it.f3, ca), cb);
            const int e2 = min(max(it.e2, ca), cb), e3 = min(max(it.e3, ca), cb);
            const ConstPos pos = (ConstPos)jb.pos;
            constexpr int cap = GX_WAVE * wave_row<PROF, 1>();
            if (it.f2 > it.s1) {
                if (!(it.e2 > it.s1 && edge_expand<1, true>(t, gpos, grec, ca, s1, w0, lane, xs, cap)))
                    voigt_staged<1, kMaskHi, true>(t, grec, pos, ca, s1, w0, lane, xs);
            } else if (it.s2 > it.s1)
                voigt_staged<1, kMaskHi>(t, grec, pos, ca, s1, w0, lane, xs);
            else
                lines_batched<PROF, kMaskHi>(t, gpos, grec, grecx, ca, s1, w0, w1, lane, incl);
            voigt_staged<1, kMaskNone, true>(t, grec, pos, e2, f2, w0, lane, xs);
            voigt_staged<1, kMaskNone>(t, grec, pos, f2, s2, w0, lane, xs);
            lines_batched<PROF, kMaskNone>(t, gpos, grec, grecx, s2, s3, w0, w1, lane, incl);
            voigt_staged<1, kMaskNone>(t, grec, pos, s3, f3, w0, lane, xs);
            voigt_staged<1, kMaskNone, true>(t, grec, pos, f3, e3, w0, lane, xs);
            if (it.f3 < it.s4) {
                if (!(it.e3 < it.s4 && edge_expand<1, false>(t, gpos, grec, s4, cb, w0, lane, xs, cap)))
                    voigt_staged<1, kMaskLo, true>(t, grec, pos, s4, cb, w0, lane, xs);
            } else if (it.s3 < it.s4)
                voigt_staged<1, kMaskLo>(t, grec, pos, s4, cb, w0, lane, xs);
            else
                lines_batched<PROF, kMaskLo>(t, gpos, grec, grecx, s4, cb, w0, w1, lane, incl);
            t.fold();
            far_lines<1>(t, grec, s1, e2, e3, s4, w0, lane);
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
            // is far below the tile (s2 > s1, or beyond kfar: f2 > s1) so
            // are all of [lo, s1); when the last one is far above (s3 < s4,
            // f3 < s4) so are all of [s4, hi)
            const int f2 = min(max(it.f2, ca), cb), f3 = min(max(it.f3, ca), cb);
            const int e2 = min(max(it.e2, ca), cb), e3 = min(max(it.e3, ca), cb);
            const GlobRec grec = (GlobRec)jb.rec;
            constexpr int cap = GX_WAVE * wave_row<PROF, PPL>();
            const GlobPos gpos = (GlobPos)jb.pos;
            if (it.f2 > it.s1) {
                if (!(it.e2 > it.s1 && edge_expand<PPL, true>(t, gpos, grec, ca, s1, w0, lane, xs, cap)))
                    voigt_staged<PPL, kMaskHi, true>(t, grec, pos, ca, s1, w0, lane, xs);
            } else if (it.s2 > it.s1)
                voigt_staged<PPL, kMaskHi>(t, grec, pos, ca, s1, w0, lane, xs);
            else
                lines_checked<PROF, PPL, kMaskHi>(t, pos, rec, recx, ca, s1, w0, w1);
            voigt_staged<PPL, kMaskNone, true>(t, grec, pos, e2, f2, w0, lane, xs);
            voigt_staged<PPL, kMaskNone>(t, grec, pos, f2, s2, w0, lane, xs);
            near_staged<PROF, PPL>(t, (GlobPos)jb.pos, grec, (GlobRecX)jb.recx, s2, s3, w0, w1, lane,
                                   xs);
            voigt_staged<PPL, kMaskNone>(t, grec, pos, s3, f3, w0, lane, xs);
            voigt_staged<PPL, kMaskNone, true>(t, grec, pos, f3, e3, w0, lane, xs);
            if (it.f3 < it.s4) {
                if (!(it.e3 < it.s4 && edge_expand<PPL, false>(t, gpos, grec, s4, cb, w0, lane, xs, cap)))
                    voigt_staged<PPL, kMaskLo, true>(t, grec, pos, s4, cb, w0, lane, xs);
            } else if (it.s3 < it.s4)
                voigt_staged<PPL, kMaskLo>(t, grec, pos, s4, cb, w0, lane, xs);
            else
                lines_checked<PROF, PPL, kMaskLo>(t, pos, rec, recx, s4, cb, w0, w1);
            t.fold();
            far_lines<PPL>(t, grec, s1, e2, e3, s4, w0, lane);
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
    // bin-Gauss: only the lines within reach of the tile, [f2, f3)
    const int lo = PROF == GX_PROF_DOPPLER ? it.f2 : it.lo;
    const int hi = max(PROF == GX_PROF_DOPPLER ? it.f3 : it.hi, lo);
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

// The item's segments: one wave per item searches them itself; the waves of
// a split item take them from the first wave's search (segs, in LDS).  The
// item's own __syncthreads before its partial sums keep segs unchanged until
// every wave has read it.
template <int PROF, int PPL, int SPLIT>
__device__ __forceinline__ void item_segments(WorkItem& it, const LinePos* jp, const int* kmax,
                                              int lane, int wv, int* segs) {
    constexpr bool gauss = PROF == GX_PROF_DOPPLER;
    if constexpr (SPLIT == 1) {
        locate_segments<gauss>(it, jp, kmax, GX_WAVE * PPL / 2, lane);
    } else {
        if (wv == 0) {
            locate_segments<gauss>(it, jp, kmax, GX_WAVE * PPL / 2, lane);
            if (lane == 0) {
                segs[0] = it.f2;
                segs[1] = it.s2;
                segs[2] = it.s3;
                segs[3] = it.f3;
                segs[4] = it.e2;
                segs[5] = it.e3;
            }
        }
        __syncthreads();
        it.f2 = __builtin_amdgcn_readfirstlane(segs[0]);
        it.s2 = __builtin_amdgcn_readfirstlane(segs[1]);
        it.s3 = __builtin_amdgcn_readfirstlane(segs[2]);
        it.f3 = __builtin_amdgcn_readfirstlane(segs[3]);
        it.e2 = __builtin_amdgcn_readfirstlane(segs[4]);
        it.e3 = __builtin_amdgcn_readfirstlane(segs[5]);
    }
}

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
                                                         const int* __restrict__ kmax,
                                                         unsigned int* __restrict__ xq,
                                                         const unsigned int* __restrict__ xqb, int mode) {
    // per wave, 64 rows of wave_row() words: the staged rows of one batch of
    // 64 lines, or at PPL 1 the W4 candidate scan of one batch; after the
    // line loops of a split item, the waves' FP64 partial sums (red)
    constexpr int kXsWords = block_threads<SPLIT>() * wave_row<PROF, PPL>();
    constexpr int kRedWords = SPLIT > 1 ? 2 * SPLIT * GX_WAVE * PPL : 0;
    __shared__ __attribute__((aligned(16))) float smem[kXsWords > kRedWords ? kXsWords : kRedWords];
    __shared__ int grab;
    __shared__ int segs[6];
    float* xs_all = smem;
    double* red = reinterpret_cast<double*>(smem);
    const int lane = threadIdx.x & (GX_WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float* xs = xs_all + wv * GX_WAVE * wave_row<PROF, PPL>();
    const int lb = mode == 3 ? (int)blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
    if constexpr (PROF == GX_PROF_DOPPLER) {  // the erfc pieces into LDS (8 KiB)
        for (int i = threadIdx.x; i < gx::kErfcPieces; i += block_threads<SPLIT>())
            gx_erfc_lds[i] = gx_erfc_pieces[i];
        __syncthreads();
    }
    if (mode == 0 || mode == 3) {
        if (SPLIT == 1) {
            const int idx = lb * kWavesPerBlock + wv;
            if (idx >= nitems) return;
            WorkItem it = items[idx];
            const Job jb = jobs[single ? 0 : it.job];
            item_segments<PROF, PPL, 1>(it, jb.pos, kmax, lane, wv, segs);
            run_item<PROF, PPL>(it, jb, scale_bits, lane, xs);
        } else {
            if (lb >= nitems) return;
            WorkItem it = items[lb];
            const Job jb = jobs[single ? 0 : it.job];
            item_segments<PROF, PPL, SPLIT>(it, jb.pos, kmax, lane, wv, segs);
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
            WorkItem it = items[idx];
            const Job jb = jobs[single ? 0 : it.job];
            item_segments<PROF, PPL, SPLIT>(it, jb.pos, kmax, lane, wv, segs);
            if (SPLIT == 1)
                run_item<PROF, PPL>(it, jb, scale_bits, lane, xs);
            else
                run_item_split<PROF, PPL, SPLIT>(it, jb, scale_bits, lane, wv, red, xs);
        }
    }
}

namespace gxk {
namespace tile_launch {

inline int blocks_for(int64_t n, int per) { return (int)((n + per - 1) / per); }

template <int PROF, int PPL, int SPLIT>
inline hipError_t launch_tiles_t(hipStream_t s, const WorkItem* items, int nitems,
                                 const Job* jobs, int single,
                                 const unsigned long long* scale_bits, const int* kmax,
                                 unsigned int* xq, const unsigned int* xqb, int mode, int blocks) {
    if (nitems <= 0) return hipSuccess;
    const int ipb = SPLIT == 1 ? kWavesPerBlock : 1;  // items per block
    const bool slots = mode == 0 || mode == 3;  // one launch slot per item
    const int nb = slots ? blocks_for(nitems, ipb)
                         : std::max(1, std::min(blocks, blocks_for(nitems, ipb)));
    hipLaunchKernelGGL((k_profile_tiles<PROF, PPL, SPLIT>), dim3(nb), dim3(block_threads<SPLIT>()), 0, s, items,
                       nitems, jobs, single, scale_bits, kmax, xq, xqb, mode);
    return hipGetLastError();
}

template <int PROF, int SPLIT>
inline hipError_t launch_ppl(hipStream_t s, int ppl, const WorkItem* items, int nitems,
                             const Job* jobs, int single, const unsigned long long* scale_bits,
                             const int* kmax, unsigned int* xq, const unsigned int* xqb, int mode,
                             int blocks) {
    switch (ppl) {
        case 1: return launch_tiles_t<PROF, 1, SPLIT>(s, items, nitems, jobs, single, scale_bits, kmax, xq, xqb, mode, blocks);
        case 2: return launch_tiles_t<PROF, 2, SPLIT>(s, items, nitems, jobs, single, scale_bits, kmax, xq, xqb, mode, blocks);
        case 4: return launch_tiles_t<PROF, 4, SPLIT>(s, items, nitems, jobs, single, scale_bits, kmax, xq, xqb, mode, blocks);
        case 8: return launch_tiles_t<PROF, 8, SPLIT>(s, items, nitems, jobs, single, scale_bits, kmax, xq, xqb, mode, blocks);
        default: return hipErrorInvalidValue;
    }
}

template <int PROF>
inline hipError_t launch_prof(hipStream_t s, int ppl, int split, const WorkItem* items,
                              int nitems, const Job* jobs, int single,
                              const unsigned long long* scale_bits, const int* kmax,
                              unsigned int* xq, const unsigned int* xqb, int mode, int blocks) {
    if (split == 1) return launch_ppl<PROF, 1>(s, ppl, items, nitems, jobs, single, scale_bits, kmax, xq, xqb, mode, blocks);
    if (split == 4) return launch_ppl<PROF, 4>(s, ppl, items, nitems, jobs, single, scale_bits, kmax, xq, xqb, mode, blocks);
    if (split == 16)
        return launch_ppl<PROF, 16>(s, ppl, items, nitems, jobs, single, scale_bits, kmax, xq, xqb, mode, blocks);
    return hipErrorInvalidValue;
}

}  // namespace tile_launch
}  // namespace gxk
