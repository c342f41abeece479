/*
 * gexs_hip.h -- C ABI of the MI355X (gfx950) line-by-line cross-section library.
 *
 * This is the boundary a HipManager (the BaseManager drop-in beside
 * Managers/Accel/GPU/GpuManager) calls.  Plain pointers and sizes only.
 * Citations are to files of the GEXS reference tree.
 *
 * Contract
 *   - Every function returns 0 on success and a negative GEXS_E* code on
 *     failure; gexs_last_error() gives the message.  Nothing calls exit().
 *   - One gexs_ctx per device and per caller thread (externally
 *     synchronised).  A ctx owns its HIP stream, device memory and graphs.
 *   - Host buffers passed in are copied before the call returns (or, for the
 *     asynchronous gexs_execute, are not touched at all); none is retained.
 *   - Grid: f[j] = start + j*dfreq, j in [0, npoints)  (Profiles/BaseProfile.cpp:10,19).
 *   - Result grids are FP64, like BaseManager::TransferResults (BaseManager.h:31-32).
 */
#ifndef GEXS_HIP_H
#define GEXS_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* profile ids: the values of ProfileType, common/Input.h:10-17 */
enum {
    GEXS_GAUSSIAN = 0,
    GEXS_DOPPLER = 1,
    GEXS_VOIGT = 2,
    GEXS_VOIGT_QUAD = 4
};

/* arithmetic of the profile sum */
enum {
    GEXS_FP32C = 0, /* FP32 evaluation, per-point compensated (Kahan) sums, FP64 fold */
    GEXS_FP64 = 1   /* plain FP64: one work-item per grid point, FP64 W4 and sums */
};

enum {
    GEXS_OK = 0,
    GEXS_EINVAL = -1,   /* bad argument */
    GEXS_ENOMEM = -2,   /* device allocation failed or over the ctx budget */
    GEXS_EHIP = -3,     /* HIP runtime error */
    GEXS_ESTATE = -4,   /* call out of order (e.g. execute before transfer) */
    GEXS_EBUSY = -5,    /* serving queue full */
    GEXS_ENOTFOUND = -6 /* unknown request id */
};

typedef struct gexs_ctx gexs_ctx;

/* Arguments of BaseManager::InitializeConstants (BaseManager.h:25) plus the
 * conventions the reference leaves implicit.  T, P and Q here are defaults;
 * gexs_execute passes its own. */
typedef struct gexs_constants {
    double half_width;    /* line-selection half width (VoigtProfile.h:15 = 100) */
    double temperature;   /* K */
    double partition;     /* Q(T) */
    double dfreq;         /* grid step, cm^-1 */
    double mean_mass;     /* amu */
    double pressure;      /* in the unit p_scale converts to gamma's unit */
    double ref_temp;      /* T_ref of gamma_ref (BaseProfile.cpp:49,53) */
    double ref_press;     /* pressure unit scale (BaseProfile.cpp:50,54) */
    double cutoff;        /* window half-width cm^-1; <= 0 selects min(25 P, 100)
                             (OpenMPManager.cpp:386) */
    double gauss_hwhm;    /* GAUSSIAN fixed HWHM (HALF-WIDTH keyword, Input.cpp:168-169) */
    int32_t width_conv;   /* 0: 1/w = 1/(nu*dpwcoef) as OpenMPManager.cpp:404;
                             1: sqrt(ln2)/(nu*dpwcoef) as cross_kernal_v2.cu:441 */
    int32_t pressure_mode;/* 0: gammaL = gamma*(Tref/T)^n (OpenMPManager.cpp:398);
                             1: times P*ref_press (cross_kernal_v2.cu:142) */
    int32_t precision;    /* GEXS_FP32C or GEXS_FP64 */
    int32_t reserved;
} gexs_constants;

/* ---------------------------------------------------------------- manager
 * One-to-one with the BaseManager virtuals (Managers/BaseManager.h:22-35). */

/* GpuManager::GpuManager (GPUManager.cpp:5-18); profile = GEXS_* profile id */
int gexs_create(int device, int profile, gexs_ctx** out);
/* GpuManager::Cleanup + delete */
int gexs_destroy(gexs_ctx* ctx);
/* InitializeVectors(int Npoints, int num_cs) (BaseManager.h:23): grid plus
 * num_cs result grids; *line_capacity receives GetNtrans() (BaseManager.h:35),
 * capped at 2^31-1 and at mem_budget (0 = 90% of free device memory). */
int gexs_init_vectors(gexs_ctx* ctx, int npoints, int num_cs, int64_t mem_budget,
                      int64_t* line_capacity);
/* InitializeConstants (BaseManager.h:25) */
int gexs_init_constants(gexs_ctx* ctx, const gexs_constants* k);
/* TransferFreq(double* freq, double* intens, int N) (BaseManager.h:27):
 * takes the grid (start = freq[0]) and zeroes every result grid. */
int gexs_transfer_freq(gexs_ctx* ctx, const double* freq, int n);
/* TransferVectors (BaseManager.h:26): one chunk of lines, host FP64 SoA.
 * gamma / nexp may be NULL for profiles without pressure broadening.
 * Waits for earlier work, uploads, then sorts the chunk by nu on the device.
 * Resets the execute counter (GPUManager.cpp:149). */
int gexs_transfer_lines(gexs_ctx* ctx, int64_t n, const double* E, const double* nu,
                        const double* A, const int32_t* g, const double* gamma,
                        const double* nexp);
/* Same, from device pointers already resident in HBM (no host copy). */
int gexs_transfer_lines_device(gexs_ctx* ctx, int64_t n, const double* d_E, const double* d_nu,
                               const double* d_A, const int32_t* d_g, const double* d_gamma,
                               const double* d_nexp);
/* ExecuteCrossSection(N, N_ener, start_idx, T, P, Q) (BaseManager.h:30):
 * asynchronous; accumulates the chunk's lines over grid points
 * [start_idx, start_idx+npoints) into result grid number k, where k counts
 * the executes since the last transfer_lines. */
int gexs_execute(gexs_ctx* ctx, int npoints, int64_t nlines, int start_idx, double T,
                 double P, double Q);
/* TransferResults(freq, out, N, I) (BaseManager.h:32): waits, then copies
 * result grid I to out (accumulate != 0 adds into out). */
int gexs_transfer_results(gexs_ctx* ctx, double* out, int n, int cs_index, int accumulate);
/* ReadyForWork (BaseManager.h:34): 1 when no work is pending, else 0 */
int gexs_ready(gexs_ctx* ctx);
/* wait for all work of the ctx */
int gexs_sync(gexs_ctx* ctx);
const char* gexs_last_error(const gexs_ctx* ctx);

/* Counters of the last gexs_execute or gexs_run_batch: N_eval (sum over
 * lines with S != 0 of the window length), lines, tiles and device times in
 * ms (HIP events on the ctx stream).  kernel_ms: an execute's profile kernel;
 * a batch's device work from the table upload through the profile kernel
 * (one graph launch when graphs are on).  total_ms adds what ran before it. */
typedef struct gexs_stats {
    int64_t n_eval;
    int64_t n_lines;
    int64_t n_tiles;
    double kernel_ms;
    double total_ms;
} gexs_stats;
int gexs_last_stats(gexs_ctx* ctx, gexs_stats* out);

/* Sums of the kernel_ms and total_ms of gexs_last_stats over every execute /
 * run_batch since the last reset, and their count
 * (at most 2048 are kept).  Waits for the ctx stream.  reset != 0 clears. */
int gexs_timing(gexs_ctx* ctx, int reset, double* kernel_ms, double* total_ms, int* count);

/* The lines of the last execute / run_batch in device order (ascending nu)
 * with the window [ib, ie) each received (OpenMPManager.cpp:399-400).  Any
 * pointer may be NULL; n must equal the pass's line count (gexs_pass_lines). */
int gexs_line_windows(gexs_ctx* ctx, int64_t n, double* nu_sorted, int32_t* ib, int32_t* ie);
/* The line selection of the last pass: lines [first, first + count) of the
 * nu-sorted current lines (all lines for gexs_execute; for gexs_run_batch the
 * lines with start - half_width <= nu <= end + half_width,
 * Profiles/BaseProfile.cpp:152-157). */
int gexs_pass_lines(gexs_ctx* ctx, int64_t* first, int64_t* count);
/* The work items (tiles) the last FP32C profile kernel read, 14 int32 each:
 * {lo, hi, w0, w1, s1, s2, s3, s4, job, group, f2, f3, e2, e3} -- the tile's
 * line range [lo, hi) among the pass's lines, its grid points [w0, w1) and
 * the segment bounds the kernel located: [lo, s1) windows ending inside the
 * tile, [s1, s4) windows covering it, of which [s2, s3) are centred near it
 * (Humlicek regions II-IV possible), [s1, f2) and [f3, s4) so far from it
 * that the one-term wing form applies, and of those [s1, e2) and [e3, s4)
 * summed by the far-field expansion; [s4, hi) windows starting inside it.
 * Bin-averaged Gaussian: [f2, f3) are the lines within reach of the tile.
 * *n receives the item count (0 after an FP64 pass); out needs cap >= *n
 * rows (cap 0: count only).  Waits for the ctx streams.  No counterpart in
 * the reference: it lets a caller see which line shares a launch shape gave
 * each wave. */
int gexs_pass_items(gexs_ctx* ctx, int64_t cap, int32_t* out, int64_t* n);

/* Tuning of the FP32C profile kernel: points per lane (1, 2, 4 or 8; a
 * wave covers 64*ppl grid points), the dispatch order of the tiles (0 = grid
 * order, 1 = heaviest first, 2 = per-XCD queues of neighbouring tiles) and
 * the waves that share one tile's lines (1, 4 or 16). */
int gexs_set_tuning(gexs_ctx* ctx, int points_per_lane, int order, int waves_per_tile);

/* ------------------------------------------------------ resident line list
 * ExoMol-style list held in HBM: a state table and transitions indexing it
 * (Readers/Exomol/exomol_objects.h:3-14, ExomolStateReader.cpp:158-162). */
typedef struct gexs_states {
    int64_t n;
    const double* E;     /* cm^-1 */
    const int32_t* g;    /* total degeneracy */
} gexs_states;

typedef struct gexs_transitions {
    int64_t n;
    const uint32_t* upper;  /* 0-based state index */
    const uint32_t* lower;
    const float* A;         /* Einstein A, s^-1 */
    const float* gamma;     /* Lorentz gamma_ref (NULL: Doppler-only list) */
    const float* nexp;      /* temperature exponent */
    /* FP64 lists: each non-NULL array here is used in place of its FP32
     * twin above (which may then be NULL), so an FP64 list keeps its values
     * (the device holds A, gamma and n in FP64 either way). */
    const double* A64;
    const double* gamma64;
    const double* nexp64;
} gexs_transitions;

/* Upload, assemble lines (nu = E_u - E_l, E'' = E_l, g' = g_u), drop lines
 * with nu == 0 (never selected, Profiles/BaseProfile.cpp:156) and sort the
 * rest by nu on the device.  Replaces any previous list. */
int gexs_load_list(gexs_ctx* ctx, const gexs_states* st, const gexs_transitions* tr);
/* Number of current lines (after load_list: the transitions with nu != 0). */
int gexs_line_count(gexs_ctx* ctx, int64_t* n);
/* Q(T) on the device from the resident states, FP64, fixed reduction order
 * (exomol_functions.cpp:50-73). */
int gexs_partition(gexs_ctx* ctx, const double* T, int nT, double* Q_out);

typedef struct gexs_grid {
    double start;
    double end;        /* dfreq = (end - start) / npoints */
    int32_t npoints;
    int32_t reserved;
} gexs_grid;

typedef struct gexs_tp {
    double T;
    double P;
    double Q;          /* <= 0: computed on the device from the resident states */
} gexs_tp;

/* Cross-sections of the resident list at ntp (T,P) points over one grid, in
 * one launch sequence (partition function, records, one profile kernel over
 * all (T,P) x tiles).  out (host or device pointer) receives ntp*npoints
 * doubles, row-major by (T,P); with out == NULL they stay in the ctx's
 * device buffer (gexs_batch_device_out).  Synchronous.  Lines are selected
 * by start - half_width <= nu <= end + half_width. */
int gexs_run_batch(gexs_ctx* ctx, const gexs_tp* tp, int ntp, const gexs_grid* grid,
                   double* out);
/* Device address of the last run_batch's results (ntp*npoints doubles) and
 * the buffer's capacity in doubles; valid until the next run_batch or
 * gexs_destroy. */
int gexs_batch_device_out(gexs_ctx* ctx, const double** out, int64_t* count);
/* ----------------------------------------------------- streamed host list
 * Cross-sections of a list that stays in host memory, streamed through the
 * device in chunks, as the reference driver feeds a manager
 * (Profiles/BaseProfile.cpp:120-200: transitions read in file order, a chunk
 * cut when it holds the manager's capacity, :174-181, which
 * Managers/Accel/GPU/GPUManager.cpp:98-101 derives from free device memory;
 * then TransferVectors + ExecuteCrossSection per (T,P), accumulated into the
 * grids).  Here:
 *   - chunks of `chunk` transitions in list order, the largest that fits
 *     mem_budget bytes of device memory together with the states, the grids
 *     and the pass buffers (stats->device_bytes);
 *   - a chunk is copied (host -> pinned staging on a host thread, then DMA on
 *     a copy stream) while the chunk before it is sorted and the one before
 *     that computed; nu = E_u - E_l, the selection start - half_width <= nu
 *     <= end + half_width (nu != 0) and the sort by nu run on the device;
 *   - each chunk's pass adds into the ntp grids in the ctx's device buffer
 *     (gexs_batch_device_out); out (host or device pointer, may be NULL)
 *     receives them at the end, ntp*npoints doubles row-major by (T,P).
 * Q <= 0 in tp: computed on the device from st.  The ctx's precision,
 * constants and tuning apply (persistent launch and tile orders as in
 * gexs_run_batch).  Synchronous.  Indices outside the state table stop the
 * call with GEXS_EINVAL (the grids then hold the chunks before it); the
 * resident list of gexs_load_list, if any, is left as it was. */
typedef struct gexs_stream_stats {
    int64_t n_transitions;  /* transitions streamed */
    int64_t n_lines;        /* of them selected */
    int64_t n_eval;         /* profile evaluations, summed over the (T,P) */
    int64_t chunk;          /* transitions per chunk */
    int64_t device_bytes;   /* device memory the call's buffers take (<= mem_budget) */
    int32_t n_chunks;
    int32_t reserved;
    double wall_ms;         /* the call, host clock */
    double device_ms;       /* first copy to last pass, HIP events */
    /* per stage, summed over the chunks (the stages overlap, so the sums may
     * exceed device_ms): host copies into the pinned staging (host clock),
     * their DMA to the device, keys + sort + gather, passes (HIP events) */
    double stage_ms, copy_ms, sort_ms, pass_ms;
} gexs_stream_stats;
int gexs_stream_xsec(gexs_ctx* ctx, const gexs_states* st, const gexs_transitions* tr,
                     const gexs_tp* tp, int ntp, const gexs_grid* grid, int64_t mem_budget,
                     double* out, gexs_stream_stats* stats);

/* Launch shape of batched passes: persistent_blocks > 0 runs the profile
 * kernel as that many workgroups pulling tiles from a counter (0: one wave
 * per tile); use_graph != 0 captures a batch's device work into a HIP graph
 * on first use and replays it while the plan (sizes, grid, tuning) is
 * unchanged, rewriting only the (T,P) table. */
int gexs_set_launch(gexs_ctx* ctx, int persistent_blocks, int use_graph);

/* Autotuner of the FP32C profile kernel: runs the batch once per candidate
 * (points per lane 1, 2, 4, 8 = tile widths of 64..512 points; 1, 4 or 16
 * waves per tile; grid, heaviest-first or per-XCD tile order; one launch slot
 * per tile or 4 / 8 persistent workgroups per CU), times the batch's device work
 * with HIP events (the pre-pass in it is the same for every candidate), and
 * keeps the fastest on the ctx.  report (may be NULL)
 * receives up to report_cap candidates; *n_report their count. */
typedef struct gexs_tuning {
    int32_t points_per_lane;
    int32_t order;
    int32_t persistent_blocks;
    int32_t waves_per_tile;
    double kernel_ms;
} gexs_tuning;
int gexs_autotune(gexs_ctx* ctx, const gexs_tp* tp, int ntp, const gexs_grid* grid,
                  gexs_tuning* best, gexs_tuning* report, int report_cap, int* n_report);

/* ------------------------------------------------------ no-allocation pass
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gexs_run_batch for one (T,P) over lines the caller holds
 * in device memory, writing a device grid the caller owns, with every
 * intermediate (pass tables, counters, per-line windows and FP32 records,
 * per-tile work items) in one file-scope device array of the library,
 * about 64 MiB, which each GPU the library runs on holds until the process
 * ends.  The call makes no hipMalloc, hipFree or other allocation.
 *   - lines: the pass's lines (the caller's selection), ascending in nu,
 *     at most GEXS_FIXED_MAX_LINES; gamma / nexp may be NULL for Doppler and
 *     Gaussian profiles.
 *   - grid: at most GEXS_FIXED_MAX_TILES tiles of 64 * points_per_lane
 *     points; d_out receives its npoints doubles (every point is written).
 *   - tp: T > 0, P >= 0 and Q > 0 (Q is not computed here).
 *   - FP32C precision, the ctx's constants and tuning; tile orders that need
 *     a host-built queue or a tile sort (heaviest first, per-XCD queues) run
 *     in dispatch order, which gives the same bits.
 *   - Asynchronous on the ctx stream when n_eval is NULL; otherwise waits and
 *     stores the pass's N_eval in *n_eval.
 * Every call on one GPU uses that GPU's array: calls that overlap on one GPU
 * (from any ctx or thread) race and must be serialised by the caller.
 * Sizes beyond the maxima are refused with GEXS_EINVAL before any launch. */
#define GEXS_FIXED_MAX_LINES (1 << 20)
#define GEXS_FIXED_MAX_TILES (1 << 17)
typedef struct gexs_lines_dev {
    int64_t n;
    const double* nu;     /* cm^-1, ascending */
    const double* E;      /* lower-state energy E'', cm^-1 */
    const double* A;      /* Einstein A, s^-1 */
    const int32_t* g;     /* upper-state degeneracy g' */
    const double* gamma;  /* Lorentz gamma_ref */
    const double* nexp;   /* temperature exponent */
} gexs_lines_dev;
int gexs_run_fixed(gexs_ctx* ctx, const gexs_lines_dev* lines, const gexs_tp* tp,
                   const gexs_grid* grid, double* d_out, int64_t* n_eval);

/* ---------------------------------------------------------------- serving
 * Continuous batching over the resident list: requests are admitted by HBM
 * footprint as they arrive and run while others are on the device.
 * gexs_serve_start: max_inflight concurrent batches (one stream each), a
 * device arena of mem_budget bytes from which each admitted request takes
 * 56 B per selected line + 8 B per grid point until it completes.  The grid
 * of a request has round((end - start)/dfreq) points with step
 * (end - start)/npoints.  While serving, only submit / poll / wait /
 * release / serve_stop may be called on the ctx (the list is read-only). */
typedef struct gexs_request {
    double T, P;
    double start, end;   /* range, cm^-1 */
    double dfreq;        /* resolution */
} gexs_request;

typedef struct gexs_result {
    int32_t status;      /* 0 done, 1 pending, < 0 error */
    int32_t npoints;
    const double* sigma; /* valid until gexs_release(id) */
    double latency_ms;   /* submit -> completion */
    double device_ms;
    /* the launch shape the server chose for the request (from its own size):
     * tile width in points, waves per tile, workgroups (line groups) per tile */
    int32_t tile_points, waves_per_tile, line_groups, reserved;
    int64_t n_eval;      /* line-point evaluations of the request (N_eval) */
} gexs_result;

int gexs_serve_start(gexs_ctx* ctx, int max_inflight, int64_t mem_budget);
int gexs_submit(gexs_ctx* ctx, const gexs_request* req, uint64_t* id);
int gexs_poll(gexs_ctx* ctx, uint64_t id, gexs_result* out);
int gexs_wait(gexs_ctx* ctx, uint64_t id, gexs_result* out);
int gexs_release(gexs_ctx* ctx, uint64_t id);
int gexs_serve_stop(gexs_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
