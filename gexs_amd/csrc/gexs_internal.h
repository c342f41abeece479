// gexs_internal.h -- layouts shared by the gfx950 kernels and the host side.
//
// Data layout in HBM (one context):
//   lines (sorted by nu, SoA, FP64 + int32 g)        40 B / line, resident
//   LinePos  {ib, ie, c, cf}   per line and grid        16 B  (grid-dependent)
//   VRec     8 words           per line and (T,P)       32 B  (FP32 profile constants)
//   WorkItem {lo, hi, w0, w1, job}  per wave-tile       32 B
//   result grids FP64                                    8 B / point
#pragma once
#include <stdint.h>

// physical constants, common/defines.h:4-27 of the reference (their exact
// values are part of the parity contract)
#define GX_LN2 0.69314718056
#define GX_SQRTLN2 0.832554611
#define GX_ISQRTPI 0.564189584
#define GX_NA 6.0221412927e23
#define GX_CM_COEF 1.327209365e-12
#define GX_PI 3.14159265359
#define GX_PLANCK 6.6260693e-27
#define GX_VELLGT 2.99792458e+10
#define GX_BOLTZ 1.380658e-16

// profile ids (common/Input.h:10-17)
#define GX_PROF_GAUSSIAN 0
#define GX_PROF_DOPPLER 1
#define GX_PROF_VOIGT 2
#define GX_PROF_VOIGT_QUAD 4


// points of a wave-tile = 64 * points-per-lane
#define GX_WAVE 64
#define GX_WG 256
// XCDs of an MI355X; workgroups are dispatched to them round-robin
#define GX_XCDS 8

// per line, per grid: window [ib, ie) clamped to the job window, nearest
// grid point c = round((nu - start)/dfreq) (as int and exact float)
struct __attribute__((aligned(16))) LinePos {
    int32_t ib, ie, c;
    float cf;
};

// per line, per (T,P): FP32 constants of the profile sum, read by the
// branch-free loop of the profile kernel.
// Voigt:       x_j = fma(cf - j, A, B) (= (nu - f_j) / w), y = gammaL / w,
//              region I:  K*Kg = Kl (s + 1) / (s^2 + H),  s = x^2 + E2,
//              E2 = y^2 - 1/2, H = 2y^2, Kl = Kg*0.5641896*y, Kg = S / w as
//              the mantissa of S / w = Kg 2^ex (Kg in [0.5, 1), Kl likewise
//              scaled by 2^-ex); the profile kernel applies 2^(ex - E), E the
//              pass's largest ex.
//              (Re of 0.5641896 t / (1/2 + t^2), t = y - ix: with b = 1/2 + y^2
//              its denominator x^4 + (2y^2 - 1) x^2 + b^2 equals s^2 + 2y^2 and
//              its numerator Kl (x^2 + b) equals Kl (s + 1).)
// Voigt-Quad:  cf, A, B, Kg as Voigt.
// bin-Gauss:   cf, A, B as Voigt with 1/w = gG, Kg 2^ex = S * 0.5/dfreq.
// (cf, A, B, Kl, E2, ex first: the far-field expansion reads those 24 bytes)
struct __attribute__((aligned(32))) VRec {
    float cf, A, B, Kl, E2;
    int32_t ex;
    float H, Kg;
};

// per line, per (T,P): what the checked loop needs beyond VRec.
// Voigt: y and k15 = smallest |c - j| with |x| + y >= 15 for certain.
// bin-Gauss: y = H, half the bin width in units of the width, and k15 the
// line's reach: it adds exactly zero at points j with |c - j| >= k15.
struct __attribute__((aligned(8))) VRecX {
    float y;
    int32_t k15;
};

// One wave-tile of one job: grid points [w0, w1) and the lines [lo, hi)
// whose windows meet them, split into segments
//   [lo, s1) window ends inside the tile, [s1, f2) far wings over the whole
//   tile (Kl / (s - 1), no checks), [f2, s2) region I over the whole tile (no
//   checks), [s2, s3) line centres near the tile (checked), [s3, f3) region
//   I, [f3, s4) far wings, [s4, hi) window starts inside the tile.
//   Of the far wings, [s1, e2) and [e3, s4) lie four half-tiles or more from
//   the tile's centre: their sum over the tile is a polynomial in the point
//   offset (the far-field expansion), summed per line instead of per point.
// s1 = s4 = hi when no line covers the whole tile (every line checked).
// bin-Gauss: [f2, f3) are the lines of [lo, hi) centred within reach of the
// tile (the job's largest reach); s2 = s3 = s4 and the kernel sums
// [f2, f3) only, masked by s1 and s4.
// With ngroups > 1 the tile's lines are cut into ngroups contiguous groups,
// one work item each (group = 0 .. ngroups - 1); such an item writes its FP64
// partial sums to row `slot` (tile * ngroups + group) of its job's Job::part
// and k_group_sum adds the rows of a tile in group order.
struct __attribute__((aligned(16))) WorkItem {
    int32_t lo, hi;
    int32_t w0, w1;
    int32_t job;
    int32_t s1, s2, s3, s4;
    int32_t group, ngroups, slot;
    int32_t f2, f3;
    int32_t e2, e3;
};
static_assert(sizeof(WorkItem) == 64, "WorkItem is one 64-byte record");

// per-(T,P) physics of one job, FP64
struct TPParams {
    double T, P, Q;
    double mass, t_ref, p_scale;
    double cutoff;
    double start, dfreq;
    double gauss_hwhm;
    int32_t profile, width_conv, pressure_mode, win_lo;
    int32_t win_hi, pad0, pad1, pad2;
};

// one cross-section
struct __attribute__((aligned(16))) Job {
    const LinePos* pos;  // job's lines, [0, n)
    const VRec* rec;
    const VRecX* recx;
    double* out;         // grid (index = grid point)
    double* part;        // rows of 64 * PPL FP64 partial sums, one per work item (line groups)
    int32_t n;
    int32_t profile;
    int32_t scale_slot;  // index into the pass-exponent array (biased largest amplitude exponent)
    int32_t accumulate;  // 1: out += tile sums (chunks accumulate); 0: out = tile sums
};

// where the data of one cross-section (job) of a pass live
struct JobDesc {
    int64_t l0;         // first line of the job among the current (nu-sorted) lines
    int64_t pos_off;    // first LinePos of the job
    int64_t rec_off;    // first VRec of the job
    int64_t recx_off;   // first VRecX of the job
    int32_t n;          // number of lines
    int32_t items_off;  // first work item
    int32_t ntiles;     // number of work items (wave-tiles)
    int32_t q_slot;     // >= 0: Q(T) is taken from this slot of a device partition pass
    int32_t groups;     // line groups per tile (<= 1: one); items = ntiles * groups
};

// plain-FP64 per (line, (T,P)) values of the baseline path
struct LineF64 {
    double Kg;    // S / w   (bin-Gauss: S)
    double y;     // gammaL / w   (bin-Gauss: unused)
    double inv_w; // 1 / w
};
