// capi.hip -- the C ABI (include/gexs_hip.h) over the gfx950 kernels.
//
// BaseManager semantics (Managers/BaseManager.h:22-35, as GpuManager implements
// them in Managers/Accel/GPU/GPUManager.cpp):
//   transfer_lines  uploads one chunk, sorts it by nu on the device, resets the
//                   execute counter (GPUManager.cpp:149)
//   execute         asynchronous; result grid = number of executes since the
//                   last transfer (one grid per (T,P), GPUManager.cpp:186-201)
//   transfer_results waits and copies one grid out (GPUManager.cpp:280-287)
// Unlike GpuManager the line data are never modified in place (the reference
// rescales gamma in place per (T,P) on concurrent streams,
// cross_kernal_v2.cu:136-146); per-(T,P) values go to separate records.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <cstdio>
#include <cstring>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gexs_hip.h"
#include "gexs_internal.h"
#include "kernels.h"
#include "serve.h"

namespace {

struct DevArr {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevArr() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t b) {
        if (b <= bytes && p) return hipSuccess;
        release();
        size_t want = std::max<size_t>(b, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        bytes = want;
        return hipSuccess;
    }
    template <typename T>
    T* as() const { return (T*)p; }
};

// Pinned host staging for small uploads (parameters, job tables): a ring of
// slots, each reused only after the copy that last read it has completed.
struct Staging {
    struct Slot {
        void* p = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    std::vector<Slot> ring;
    size_t next = 0;
    ~Staging() {
        for (auto& s : ring) {
            if (s.p) (void)hipHostFree(s.p);
            if (s.done) (void)hipEventDestroy(s.done);
        }
    }
    // copies src to dst on stream through a pinned slot
    hipError_t upload(void* dst, const void* src, size_t bytes, hipStream_t stream) {
        if (bytes == 0) return hipSuccess;
        if (ring.empty()) ring.resize(64);
        Slot& sl = ring[next];
        next = (next + 1) % ring.size();
        hipError_t e;
        if (sl.pending && (e = hipEventSynchronize(sl.done)) != hipSuccess) return e;
        if (sl.bytes < bytes) {
            if (sl.p) (void)hipHostFree(sl.p);
            sl.p = nullptr;
            size_t want = std::max<size_t>(bytes, 4096);
            if ((e = hipHostMalloc(&sl.p, want, hipHostMallocDefault)) != hipSuccess) return e;
            sl.bytes = want;
        }
        if (!sl.done && (e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess)
            return e;
        std::memcpy(sl.p, src, bytes);
        if ((e = hipMemcpyAsync(dst, sl.p, bytes, hipMemcpyHostToDevice, stream)) != hipSuccess) return e;
        if ((e = hipEventRecord(sl.done, stream)) != hipSuccess) return e;
        sl.pending = true;
        return hipSuccess;
    }
};

struct PosKey {
    double cutoff = NAN, start = NAN, dfreq = NAN;
    int win_lo = -1, win_hi = -1;
    int64_t l0 = -1, n = -1;  // line sub-range [l0, l0 + n) of the current lines
    uint64_t gen = 0;
    bool operator==(const PosKey& o) const {
        return cutoff == o.cutoff && start == o.start && dfreq == o.dfreq && win_lo == o.win_lo &&
               win_hi == o.win_hi && l0 == o.l0 && n == o.n && gen == o.gen;
    }
};

void swap_arr(DevArr& a, DevArr& b) {
    std::swap(a.p, b.p);
    std::swap(a.bytes, b.bytes);
}

// Buffers of gexs_stream_xsec, kept (and grown) between calls.  Two sets of
// chunk buffers: while the pass of chunk k reads sorted set k mod 2, chunk
// k + 1 is copied into raw set (k + 1) mod 2, sorted and gathered into the
// other sorted set, and chunk k + 2 is staged on the host.
struct StreamBufs {
    hipStream_t cp = nullptr;  // host -> device copies
    hipEvent_t h2d_done[2] = {nullptr, nullptr};   // raw set written, staging set free
    hipEvent_t gathered[2] = {nullptr, nullptr};   // sorted set written, raw set free
    hipEvent_t pass_done[2] = {nullptr, nullptr};  // sorted set free
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::vector<hipEvent_t> tev;  // per chunk: copy, sort, pass (begin, end each)
    void* h[2] = {nullptr, nullptr};  // pinned staging, one chunk's columns each
    size_t h_bytes = 0;
    unsigned long long* h_small = nullptr;  // pinned read-back slot
    DevArr raw[2];                          // device copy of one chunk's columns
    DevArr nu[2], E[2], A[2], gamma[2], nexp[2], g[2];  // a chunk's selected lines, by nu
    DevArr keys, vals, dropped, st_E, st_g, q_T, q_part, q_Q;
    gxk::SortBuffers sb;
    hipError_t init() {
        hipError_t e = hipStreamCreateWithFlags(&cp, hipStreamNonBlocking);
        for (int i = 0; i < 2 && e == hipSuccess; ++i) {
            e = hipEventCreateWithFlags(&h2d_done[i], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&gathered[i], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&pass_done[i], hipEventDisableTiming);
        }
        if (e == hipSuccess) e = hipEventCreate(&t0);
        if (e == hipSuccess) e = hipEventCreate(&t1);
        if (e == hipSuccess) e = hipHostMalloc((void**)&h_small, 64, hipHostMallocDefault);
        return e;
    }
    ~StreamBufs() {
        if (cp) (void)hipStreamSynchronize(cp);
        for (int i = 0; i < 2; ++i) {
            if (h[i]) (void)hipHostFree(h[i]);
            if (h2d_done[i]) (void)hipEventDestroy(h2d_done[i]);
            if (gathered[i]) (void)hipEventDestroy(gathered[i]);
            if (pass_done[i]) (void)hipEventDestroy(pass_done[i]);
        }
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        for (hipEvent_t e : tev) (void)hipEventDestroy(e);
        if (h_small) (void)hipHostFree(h_small);
        gxk::sort_buffers_free(sb);
        if (cp) (void)hipStreamDestroy(cp);
    }
};

// ------------------------------------------------------ no-allocation pass
// gexs_run_fixed keeps every intermediate of its pass in one file-scope
// device array (each GPU the library runs on holds its own copy until the
// process ends): the pass tables and counters, then per line its window
// (LinePos), FP32 record (VRec, VRecX), then per tile its work item.
constexpr int64_t kFixedMaxLines = GEXS_FIXED_MAX_LINES;
constexpr int64_t kFixedMaxTiles = GEXS_FIXED_MAX_TILES;
constexpr size_t kFixedTabOff = 0;     // TPParams | JobDesc | Job (256-B aligned each)
constexpr size_t kFixedCntOff = 1024;  // xq (8 x u32) | scale | neval | kmax: 7 x u64
constexpr int kFixedCntWords = 7;
constexpr size_t kFixedPosOff = 2048;
constexpr size_t kFixedRecOff = kFixedPosOff + (size_t)kFixedMaxLines * sizeof(LinePos);
constexpr size_t kFixedRecxOff = kFixedRecOff + (size_t)kFixedMaxLines * sizeof(VRec);
constexpr size_t kFixedItemsOff = kFixedRecxOff + (size_t)kFixedMaxLines * sizeof(VRecX);
constexpr size_t kFixedBytes = kFixedItemsOff + (size_t)kFixedMaxTiles * sizeof(WorkItem);
static_assert(3 * 256 <= kFixedCntOff, "pass tables overlap the counters");

}  // namespace

static __device__ __attribute__((aligned(256))) unsigned char scratch[kFixedBytes];

namespace {
// device address of `scratch` per device (hipGetSymbolAddress, taken once per
// device with it current; the array's host-side name is not a device address)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// called by gexs_create with `device` current
void scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch_ptr[device] = p;
}
}  // namespace

// buffer sets of the execute path: pass k uses set k mod kPassSets and clears
// the counters of set k + 1 (serial mode) or k + 2 (alternating mode)
constexpr int kPassSets = 4;
// profile kernels at least this long (ms, the last measured execute) run
// back to back on `stream`, their line records on `pre` (joined by events,
// ~30 us of host time per pass); shorter passes alternate between the two
// streams with no event per pass
constexpr float kSerialKernelMs = 0.15f;

struct gexs_ctx {
    int device = 0;
    int profile = GEXS_VOIGT;
    hipStream_t stream = nullptr;
    // one (start, kernel begin, kernel end) event triple per timed execute
    // since the last gexs_timing(reset); the last one feeds gexs_last_stats
    std::vector<hipEvent_t> evpool;
    int tcount = 0;
    hipEvent_t ev_start = nullptr, ev_k0 = nullptr, ev_k1 = nullptr;
    std::string err;

    gexs_constants k{};
    bool have_k = false;
    int npoints = 0;
    int num_cs = 0;
    double start = 0.0;
    bool have_grid = false;
    std::vector<double*> results;
    int64_t capacity = 0;
    int tasks = 0;

    // current lines, sorted by nu (FP64 SoA + int32 g)
    int64_t n = 0;
    bool has_gamma = false;
    uint64_t gen = 0;
    DevArr nu, E, A, gamma, nexp, g;
    // unsorted upload staging and sort scratch
    DevArr r_nu, r_E, r_A, r_gamma, r_nexp, r_g, keys, perm, scratch8;
    gxk::SortBuffers sb;

    // resident states (for the device partition function)
    int64_t n_states = 0;
    DevArr st_E, st_g, q_T, q_part, q_Q;

    // derived per grid / per (T,P)
    DevArr pos, rec, recx, lf64, items, items_tmp, cost_key, item_idx;
    DevArr tab;      // per-pass tables [TPParams | JobDesc | Job] x njobs (one upload)
    DevArr small;    // per-pass counters [xq | scale | neval | kmax] (one fill)
    DevArr rec_parts;  // workgroup sums of a multi-job line-record launch (njobs x gxk::kRecRedBytes)
    DevArr neval;    // N_eval of the plain FP64 pass
    DevArr counter;  // bounds of the per-XCD tile queues
    DevArr item_tab; // window and job tables of the tiles kernel
    const void* last_neval = nullptr;  // N_eval of the last pass (device)
    const WorkItem* last_items = nullptr;  // work items the last FP32C profile kernel read
    const int* last_kmax = nullptr;        // and the kmax it located their segments with

    // execute path: a pass's line records run on `pre` or on `stream` (see
    // run_fp32c), so that they overlap the profile kernel of the pass before
    hipStream_t pre = nullptr;
    struct PassBuf {
        DevArr tab, small, rec, recx;
        DevArr red;  // the line records' workgroup sums (gxk::kRecRedBytes)
        bool clean = false;  // counters zero (cleared by the pass two before)
        hipEvent_t pre_done = nullptr, done = nullptr;  // line records / profile kernel
        bool done_recorded = false;
    };
    PassBuf pb[kPassSets];
    int64_t npass = 0;  // passes of the execute path so far
    // work on `pre` that `stream` has not waited for, and work on `stream`
    // (anything but the execute passes) that `pre` has not waited for
    bool pre_pending = false, main_dirty = true;
    hipEvent_t ev_pre = nullptr, ev_main = nullptr;
    // profile-kernel time of the last execute whose events are complete, the
    // timing events of the last execute, and the mode of the last pass
    float pass_kernel_ms = 1.0f;
    hipEvent_t prev_k0 = nullptr, prev_k1 = nullptr;
    int pass_mode = -1;  // 1: serial, 0: alternating, -1: none yet
    int cus = 256;  // compute units of the device
    PosKey pos_key;
    PosKey items_key;
    int items_span = 0, items_order = -1, items_njobs = 0, n_items = 0;

    int ppl = 4;
    int split = 1;  // waves per tile
    int order = 1;
    int persistent_blocks = 0;
    gexs_stats last{};
    bool last_neval_pending = false;

    // batched pass: pinned tables, cached selection and captured graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::string graph_key;
    bool use_graph = false;
    void* h_tab = nullptr;
    size_t h_tab_cap = 0;
    double sel_lo = NAN, sel_hi = NAN;
    uint64_t sel_gen = 0;
    int64_t sel_l0 = 0, sel_l1 = 0;
    DevArr batch_out;

    std::unique_ptr<gxs::Server> server;
    std::unique_ptr<StreamBufs> strm;  // gexs_stream_xsec
    Staging staging;
    DevArr pos_jd, pos_tp;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int hipfail(hipError_t e, const char* where) {
        err = std::string(where) + ": " + hipGetErrorString(e);
        return GEXS_EHIP;
    }
};

#define HIPCHK(ctx, call)                                   \
    do {                                                    \
        hipError_t e__ = (call);                            \
        if (e__ != hipSuccess) return (ctx)->hipfail(e__, #call); \
    } while (0)

namespace {

double effective_cutoff(const gexs_constants& k, double P) {
    if (k.cutoff > 0.0) return k.cutoff;
    return std::min(25.0 * P, 100.0);  // OpenMPManager.cpp:386
}

TPParams make_tp(const gexs_ctx* c, double T, double P, double Q, double start, double dfreq,
                 int win_lo, int win_hi) {
    TPParams p{};
    p.T = T;
    p.P = P;
    p.Q = Q;
    p.mass = c->k.mean_mass;
    p.t_ref = c->k.ref_temp;
    p.p_scale = c->k.ref_press;
    p.cutoff = effective_cutoff(c->k, P);
    p.start = start;
    p.dfreq = dfreq;
    p.gauss_hwhm = c->k.gauss_hwhm;
    p.profile = c->profile;
    p.width_conv = c->k.width_conv;
    p.pressure_mode = c->k.pressure_mode;
    p.win_lo = win_lo;
    p.win_hi = win_hi;
    return p;
}

constexpr int kMaxTimed = 2048;

// selects the event triple of the next timed execute
int next_events(gexs_ctx* c) {
    int t = c->tcount < kMaxTimed ? c->tcount : kMaxTimed - 1;
    while ((int)c->evpool.size() < 3 * (t + 1)) {
        hipEvent_t e;
        hipError_t err = hipEventCreate(&e);
        if (err != hipSuccess) return c->hipfail(err, "hipEventCreate");
        c->evpool.push_back(e);
    }
    c->ev_start = c->evpool[3 * t];
    c->ev_k0 = c->evpool[3 * t + 1];
    c->ev_k1 = c->evpool[3 * t + 2];
    if (c->tcount < kMaxTimed) c->tcount++;
    return GEXS_OK;
}

int set_device_only(gexs_ctx* c) {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return c->hipfail(e, "hipSetDevice");
    return GEXS_OK;
}

// the execute passes on `pre` join `stream`: what is enqueued on `stream`
// from here on runs after them
int join_pre(gexs_ctx* c) {
    if (!c->pre_pending) return GEXS_OK;
    HIPCHK(c, hipEventRecord(c->ev_pre, c->pre));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_pre, 0));
    c->pre_pending = false;
    return GEXS_OK;
}

// every entry point but the execute passes: the device current, the
// execute passes on `pre` joined into `stream`, and `pre` to wait for
// what the caller enqueues on `stream` before its next execute pass
int set_device(gexs_ctx* c) {
    int rc;
    if ((rc = set_device_only(c)) != GEXS_OK) return rc;
    if ((rc = join_pre(c)) != GEXS_OK) return rc;
    c->main_dirty = true;
    return GEXS_OK;
}

// sort the staged (r_*) lines by nu into the current arrays
// drop_zero: lines at nu == 0 are removed (the driver never selects them,
// Profiles/BaseProfile.cpp:156); the manager path keeps every line it is given.
int sort_and_commit(gexs_ctx* c, int64_t n, bool has_gamma, bool drop_zero) {
    hipStream_t s = c->stream;
    HIPCHK(c, c->keys.ensure((size_t)n * 8));
    HIPCHK(c, c->perm.ensure((size_t)n * 4));
    HIPCHK(c, c->scratch8.ensure(64));
    HIPCHK(c, gxk::fill_u64(s, c->scratch8.as<unsigned long long>(), 1, 0ull));
    HIPCHK(c, gxk::make_nu_keys(s, c->r_nu.as<double>(), n, c->keys.as<uint64_t>(), c->perm.as<uint32_t>(),
                                drop_zero ? 1 : 0, c->scratch8.as<unsigned long long>()));
    HIPCHK(c, gxk::radix_sort_u64(s, c->keys.as<uint64_t>(), c->perm.as<uint32_t>(), n, c->sb, 0, 64));
    if (drop_zero) {
        unsigned long long zeros = 0;
        HIPCHK(c, hipMemcpyAsync(&zeros, c->scratch8.p, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        n -= (int64_t)zeros;
    }
    HIPCHK(c, c->nu.ensure((size_t)n * 8));
    HIPCHK(c, c->E.ensure((size_t)n * 8));
    HIPCHK(c, c->A.ensure((size_t)n * 8));
    HIPCHK(c, c->g.ensure((size_t)n * 4));
    const uint32_t* pm = c->perm.as<uint32_t>();
    HIPCHK(c, gxk::gather_f64(s, c->r_nu.as<double>(), pm, n, c->nu.as<double>()));
    HIPCHK(c, gxk::gather_f64(s, c->r_E.as<double>(), pm, n, c->E.as<double>()));
    HIPCHK(c, gxk::gather_f64(s, c->r_A.as<double>(), pm, n, c->A.as<double>()));
    HIPCHK(c, gxk::gather_i32(s, c->r_g.as<int32_t>(), pm, n, c->g.as<int32_t>()));
    if (has_gamma) {
        HIPCHK(c, c->gamma.ensure((size_t)n * 8));
        HIPCHK(c, c->nexp.ensure((size_t)n * 8));
        HIPCHK(c, gxk::gather_f64(s, c->r_gamma.as<double>(), pm, n, c->gamma.as<double>()));
        HIPCHK(c, gxk::gather_f64(s, c->r_nexp.as<double>(), pm, n, c->nexp.as<double>()));
    }
    c->n = n;
    c->has_gamma = has_gamma;
    c->gen++;
    c->tasks = 0;
    return GEXS_OK;
}

// window + nearest point of every current line for one grid window:
// lines [l0, l0 + n) of the current lines take part in the pass
int ensure_pos(gexs_ctx* c, const TPParams& tp, int64_t l0, int64_t n) {
    PosKey key;
    key.cutoff = tp.cutoff;
    key.start = tp.start;
    key.dfreq = tp.dfreq;
    key.win_lo = tp.win_lo;
    key.win_hi = tp.win_hi;
    key.l0 = l0;
    key.n = n;
    key.gen = c->gen;
    if (key == c->pos_key) return GEXS_OK;
    int rc;
    if ((rc = join_pre(c)) != GEXS_OK) return rc;  // passes in flight read the old positions
    c->main_dirty = true;
    HIPCHK(c, c->pos.ensure((size_t)std::max<int64_t>(n, 1) * sizeof(LinePos)));
    JobDesc d{};
    d.l0 = l0;
    d.n = (int32_t)n;
    d.q_slot = -1;
    HIPCHK(c, c->pos_jd.ensure(sizeof(JobDesc)));
    HIPCHK(c, c->pos_tp.ensure(sizeof(TPParams)));
    HIPCHK(c, c->staging.upload(c->pos_jd.p, &d, sizeof d, c->stream));
    HIPCHK(c, c->staging.upload(c->pos_tp.p, &tp, sizeof tp, c->stream));
    HIPCHK(c, gxk::line_pos(c->stream, c->nu.as<double>(), c->pos_jd.as<JobDesc>(), 1, n,
                            c->pos_tp.as<TPParams>(), c->pos.as<LinePos>()));
    c->pos_key = key;
    c->items_key = PosKey();  // tiles depend on positions
    return GEXS_OK;
}

// Per-pass tables of njobs jobs: TPParams, JobDesc and Job arrays at 256-B
// aligned offsets of c->tab (uploaded with one copy), and the counters the
// pass clears with one fill: [xq (GX_XCDS u32) | scale | neval | kmax].
struct PassTables {
    TPParams* tp;
    JobDesc* jd;
    Job* jobs;
    size_t jd_off, jobs_off, bytes;
    unsigned int* xq;
    unsigned long long* scale;
    unsigned long long* neval;
    int* kmax;
    int64_t zero_u64;
};

size_t al256(size_t b) { return (b + 255) / 256 * 256; }

int pass_tables_in(gexs_ctx* c, DevArr& tab, DevArr& small, int njobs, PassTables* t) {
    const size_t tp_b = al256((size_t)njobs * sizeof(TPParams));
    const size_t jd_b = al256((size_t)njobs * sizeof(JobDesc));
    const size_t jobs_b = al256((size_t)njobs * sizeof(Job));
    HIPCHK(c, tab.ensure(tp_b + jd_b + jobs_b));
    HIPCHK(c, small.ensure((4 + 3 * (size_t)njobs) * 8));
    char* b = (char*)tab.p;
    t->jd_off = tp_b;
    t->jobs_off = tp_b + jd_b;
    t->bytes = tp_b + jd_b + jobs_b;
    t->tp = (TPParams*)b;
    t->jd = (JobDesc*)(b + t->jd_off);
    t->jobs = (Job*)(b + t->jobs_off);
    char* s = (char*)small.p;
    t->xq = (unsigned int*)s;
    t->scale = (unsigned long long*)(s + 32);
    t->neval = t->scale + njobs;
    t->kmax = (int*)(t->neval + njobs);
    t->zero_u64 = 4 + 3 * (int64_t)njobs;
    return GEXS_OK;
}

int pass_tables(gexs_ctx* c, int njobs, PassTables* t) { return pass_tables_in(c, c->tab, c->small, njobs, t); }

int tiles_per_job(const gexs_ctx* c) {
    const int span = GX_WAVE * c->ppl;
    const int w = c->pos_key.win_hi - c->pos_key.win_lo;
    return w > 0 ? (w + span - 1) / span : 0;
}

// profile-kernel launch shape: order 2 = per-XCD queues (persistent);
// persistent_blocks > 0 with order 0/1 = one shared queue; otherwise one
// launch slot per tile, XCD-contiguous for order 0 and in dispatch order
// (consecutive tiles on different XCDs) for heaviest-first and order 3
int launch_mode(const gexs_ctx* c) {
    if (c->order == 2) return 2;
    if (c->persistent_blocks > 0) return 1;
    return c->order == 0 ? 0 : 3;
}
int launch_blocks(const gexs_ctx* c) {
    if (c->persistent_blocks > 0) return c->persistent_blocks;
    return 8 * c->cus;
}

// work items for njobs jobs over one grid window, from the jobs' window and
// job tables (host copies, uploaded when the items are rebuilt); sorted
// heaviest first when order == 1; with order == 2 in grid order and split
// into GX_XCDS contiguous ranges of equal line count for the per-XCD queues.
// Rebuilds run on c->stream after the execute passes in flight.
int ensure_items(gexs_ctx* c, int njobs, const TPParams* htp, const JobDesc* hjd) {
    const int span = GX_WAVE * c->ppl;
    const PosKey key = c->pos_key;
    if (key == c->items_key && c->items_span == span && c->items_order == c->order &&
        c->items_njobs == njobs)
        return GEXS_OK;
    int rc;
    if ((rc = join_pre(c)) != GEXS_OK) return rc;  // passes in flight read the old items
    c->main_dirty = true;
    const int per = tiles_per_job(c);
    const int total = per * njobs;
    HIPCHK(c, c->items.ensure((size_t)std::max(total, 1) * sizeof(WorkItem)));
    HIPCHK(c, c->items_tmp.ensure((size_t)std::max(total, 1) * sizeof(WorkItem)));
    HIPCHK(c, c->cost_key.ensure((size_t)std::max(total, 1) * 4));
    HIPCHK(c, c->item_idx.ensure((size_t)std::max(total, 1) * 4));
    const bool sorted = c->order == 1 && total > 1;
    WorkItem* dst = sorted ? c->items_tmp.as<WorkItem>() : c->items.as<WorkItem>();
    const size_t tp_b = al256((size_t)njobs * sizeof(TPParams));
    HIPCHK(c, c->item_tab.ensure(tp_b + (size_t)njobs * sizeof(JobDesc)));
    HIPCHK(c, c->staging.upload(c->item_tab.p, htp, (size_t)njobs * sizeof(TPParams), c->stream));
    HIPCHK(c, c->staging.upload((char*)c->item_tab.p + tp_b, hjd, (size_t)njobs * sizeof(JobDesc), c->stream));
    HIPCHK(c, gxk::tiles(c->stream, c->pos.as<LinePos>(), (const JobDesc*)((char*)c->item_tab.p + tp_b), njobs,
                         per, c->item_tab.as<TPParams>(), span, dst,
                         sorted ? c->cost_key.as<uint32_t>() : nullptr,
                         sorted ? c->item_idx.as<uint32_t>() : nullptr));
    if (sorted) {
        HIPCHK(c, gxk::radix_sort_u32(c->stream, c->cost_key.as<uint32_t>(), c->item_idx.as<uint32_t>(),
                                      total, c->sb, 0, 32));
        HIPCHK(c, gxk::permute_items(c->stream, c->items_tmp.as<WorkItem>(), c->item_idx.as<uint32_t>(),
                                     total, c->items.as<WorkItem>()));
    }
    HIPCHK(c, c->counter.ensure(256));
    if (c->order == 2 && total > 0) {
        // cost-balanced split of the grid-ordered items over the XCD queues
        std::vector<WorkItem> h((size_t)total);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->items.p, (size_t)total * sizeof(WorkItem), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double tot = 0.0;
        for (const auto& w : h) tot += (double)std::max(w.hi - w.lo, 0) + 1.0;
        uint32_t bounds[GX_XCDS + 1];
        bounds[0] = 0;
        double acc = 0.0;
        int q = 1;
        for (int i = 0; i < total && q < GX_XCDS; ++i) {
            acc += (double)std::max(h[(size_t)i].hi - h[(size_t)i].lo, 0) + 1.0;
            while (q < GX_XCDS && acc >= tot * q / GX_XCDS) bounds[q++] = (uint32_t)(i + 1);
        }
        while (q <= GX_XCDS) bounds[q++] = (uint32_t)total;
        bounds[GX_XCDS] = (uint32_t)total;
        HIPCHK(c, c->staging.upload(c->counter.p, bounds, sizeof bounds, c->stream));
    }
    c->items_key = key;
    c->items_span = span;
    c->items_order = c->order;
    c->items_njobs = njobs;
    c->n_items = total;
    return GEXS_OK;
}

int select_range(gexs_ctx* c, double lo, double hi, int64_t* l0, int64_t* l1) {
    if (c->n == 0) {
        *l0 = *l1 = 0;
        return GEXS_OK;
    }
    HIPCHK(c, c->scratch8.ensure(64));
    HIPCHK(c, gxk::nu_range(c->stream, c->nu.as<double>(), c->n, lo, hi, c->scratch8.as<long long>()));
    long long r[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(r, c->scratch8.p, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *l0 = r[0];
    *l1 = r[1];
    return GEXS_OK;
}

// The FP32C pass of one execute over the lines and window of c->pos_key into
// out: the line records (k_line_rec_pass with the pass's tables as kernel
// arguments, its line records into buffer set k mod kPassSets, the counters
// of a later pass's set cleared; then k_rec_reduce adds its workgroups' sums
// to the pass's counters) and the profile kernel.  The line records of one pass
// overlap the profile kernel of the pass before, in one of two modes:
//   serial (profile kernels of kSerialKernelMs or more): line records on
//   `pre`, clearing set k + 1's counters once the profile kernel of pass
//   k - 3 (that set's last reader, after pass k - 4, this set's) is done
//   (they start as pass k - 1's kernel does); profile kernels back to back
//   on `stream`, each after its pass's line records;
//   alternating (shorter kernels, whose passes the host's event calls would
//   limit): pass k runs on `stream` (k even) or `pre` (k odd) and clears set
//   k + 2's counters, so every set is used by the passes of one stream, in
//   order, with no event per pass.
// A change of mode first joins the two streams both ways.  Every pass
// computes exactly what it would alone.
int run_fp32c(gexs_ctx* c, const TPParams& tp, double* out, bool timed) {
    const int n = (int)c->pos_key.n;
    const int64_t l0 = c->pos_key.l0;
    int rc;
    const int per = tiles_per_job(c);
    JobDesc d{};
    d.l0 = l0;
    d.pos_off = 0;
    d.rec_off = 0;
    d.recx_off = 0;
    d.n = n;
    d.items_off = 0;
    d.ntiles = per;
    d.q_slot = -1;
    if ((rc = ensure_items(c, 1, &tp, &d)) != GEXS_OK) return rc;
    // the profile-kernel time of the last execute, once its events are complete
    if (c->prev_k1 && hipEventQuery(c->prev_k1) == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->prev_k0, c->prev_k1) == hipSuccess) c->pass_kernel_ms = ms;
    }
    const int mode = c->pass_kernel_ms >= kSerialKernelMs ? 1 : 0;
    const bool serial = mode == 1;
    const int64_t k = c->npass;
    gexs_ctx::PassBuf& B = c->pb[k % kPassSets];
    gexs_ctx::PassBuf& N = c->pb[(k + (serial ? 1 : 2)) % kPassSets];  // counters cleared here
    PassTables t, tn;
    const void* small_was = B.small.p;
    if ((rc = pass_tables_in(c, B.tab, B.small, 1, &t)) != GEXS_OK) return rc;
    if (B.small.p != small_was) B.clean = false;  // new memory
    if ((rc = pass_tables_in(c, N.tab, N.small, 1, &tn)) != GEXS_OK) return rc;
    HIPCHK(c, B.rec.ensure((size_t)std::max(n, 1) * sizeof(VRec)));
    HIPCHK(c, B.recx.ensure((size_t)std::max(n, 1) * sizeof(VRecX)));
    HIPCHK(c, B.red.ensure(gxk::kRecRedBytes));
    HIPCHK(c, c->counter.ensure(256));
    Job jb{};
    jb.pos = c->pos.as<LinePos>();
    jb.rec = B.rec.as<VRec>();
    jb.recx = B.recx.as<VRecX>();
    jb.out = out;
    jb.n = n;
    jb.profile = c->profile;
    jb.scale_slot = 0;
    jb.accumulate = 1;
    if (mode != c->pass_mode) {  // join the streams both ways
        HIPCHK(c, hipEventRecord(c->ev_pre, c->pre));
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_pre, 0));
        HIPCHK(c, hipEventRecord(c->ev_main, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->pre, c->ev_main, 0));
        c->pre_pending = false;
        c->main_dirty = false;
        c->pass_mode = mode;
    }
    hipStream_t P = (serial || (k & 1)) ? c->pre : c->stream;  // line records
    hipStream_t S = serial ? c->stream : P;                    // profile kernel
    if (P == c->pre && c->main_dirty) {  // lines, grids or tiles changed on `stream`
        HIPCHK(c, hipEventRecord(c->ev_main, c->stream));
        HIPCHK(c, hipStreamWaitEvent(P, c->ev_main, 0));
        c->main_dirty = false;
    }
    if (serial) {
        // pass k - 3's profile kernel must be done; waiting for the marker
        // just before pass k - 1's (on `stream` after pass k - 2's) also
        // puts these line records behind that kernel's dispatch, so their
        // blocks take the CU slots its tail leaves free instead of slots its
        // first workgroups need (a wait only if the marker is still ahead)
        hipEvent_t w = c->prev_k0 ? c->prev_k0 : (N.done_recorded ? N.done : nullptr);
        if (w) {
            const hipError_t q = hipEventQuery(w);
            if (q == hipErrorNotReady)
                HIPCHK(c, hipStreamWaitEvent(P, w, 0));
            else if (q != hipSuccess)
                HIPCHK(c, q);
        }
    }
    if (!B.clean) HIPCHK(c, gxk::fill_u64(P, (unsigned long long*)B.small.p, t.zero_u64, 0ull));
    HIPCHK(c, hipEventRecord(c->ev_start, P));
    const double* gam = c->has_gamma ? c->gamma.as<double>() : nullptr;
    const double* nex = c->has_gamma ? c->nexp.as<double>() : nullptr;
    // workgroups: four per CU beside a long kernel, one per CU beside a short
    // one (sweeps of 1-8 per CU: 1e6-line Voigt steps 0.352-0.363 ms, 2e5-line
    // Doppler steps 26-31 us; the best shapes are these, by 1-5 %)
    HIPCHK(c, gxk::line_rec_pass(P, tp, d, jb, t.jobs, (unsigned long long*)N.small.p, (int)tn.zero_u64,
                                 c->nu.as<double>(), c->E.as<double>(), c->A.as<double>(),
                                 c->g.as<int32_t>(), gam, nex, c->pos.as<LinePos>(), t.scale,
                                 B.rec.as<VRec>(), B.recx.as<VRecX>(), t.kmax, t.neval, B.red.p,
                                 serial ? 4 * c->cus : c->cus));
    N.clean = true;
    if (serial) {
        HIPCHK(c, hipEventRecord(B.pre_done, P));
        HIPCHK(c, hipStreamWaitEvent(S, B.pre_done, 0));
    }
    if (timed) HIPCHK(c, hipEventRecord(c->ev_k0, S));
    HIPCHK(c, gxk::profile_tiles(S, c->profile, c->ppl, c->split, c->items.as<WorkItem>(), c->n_items,
                                 t.jobs, t.scale, t.kmax, t.xq, c->counter.as<unsigned int>(),
                                 launch_mode(c), launch_blocks(c), true));
    if (timed) HIPCHK(c, hipEventRecord(c->ev_k1, S));
    if (serial) HIPCHK(c, hipEventRecord(B.done, S));
    B.done_recorded = serial;
    c->prev_k0 = timed ? c->ev_k0 : nullptr;
    c->prev_k1 = timed ? c->ev_k1 : nullptr;
    B.clean = false;
    c->npass = k + 1;
    if (S == c->pre) c->pre_pending = true;  // serial: `stream` waited for `pre`
    c->last.n_tiles = c->n_items;
    c->last_neval = t.neval;
    c->last_items = c->items.as<WorkItem>();
    c->last_kmax = t.kmax;
    return GEXS_OK;
}

// reset_neval = false adds this pass's N_eval to the count of the passes
// before it (a batch of several (T,P))
int run_fp64(gexs_ctx* c, const TPParams& tp, double* out, bool timed, bool reset_neval = true) {
    const int n = (int)c->pos_key.n;
    const int64_t l0 = c->pos_key.l0;
    hipStream_t s = c->stream;
    int rc;
    PassTables t;
    if ((rc = pass_tables(c, 1, &t)) != GEXS_OK) return rc;
    HIPCHK(c, c->staging.upload(t.tp, &tp, sizeof tp, s));
    HIPCHK(c, c->lf64.ensure((size_t)std::max(n, 1) * sizeof(LineF64)));
    HIPCHK(c, c->neval.ensure(8));
    if (reset_neval) HIPCHK(c, gxk::fill_u64(s, c->neval.as<unsigned long long>(), 1, 0ull));
    const double* gam = c->has_gamma ? c->gamma.as<double>() + l0 : nullptr;
    const double* nex = c->has_gamma ? c->nexp.as<double>() + l0 : nullptr;
    HIPCHK(c, gxk::line_f64(s, c->nu.as<double>() + l0, c->E.as<double>() + l0, c->A.as<double>() + l0,
                            c->g.as<int32_t>() + l0, gam, nex, n, c->pos.as<LinePos>(), t.tp,
                            c->lf64.as<LineF64>(), c->neval.as<unsigned long long>()));
    if (timed) HIPCHK(c, hipEventRecord(c->ev_k0, s));
    HIPCHK(c, gxk::fp64_points(s, c->profile, c->nu.as<double>() + l0, c->pos.as<LinePos>(), c->lf64.as<LineF64>(), n,
                               tp.start, tp.dfreq, tp.win_lo, tp.win_hi, out));
    if (timed) HIPCHK(c, hipEventRecord(c->ev_k1, s));
    c->last.n_tiles = 0;
    c->last_neval = c->neval.p;
    return GEXS_OK;
}

bool valid_profile(int p) {
    return p == GEXS_GAUSSIAN || p == GEXS_DOPPLER || p == GEXS_VOIGT || p == GEXS_VOIGT_QUAD;
}

}  // namespace

// =====================================================================  C ABI
extern "C" {

int gexs_create(int device, int profile, gexs_ctx** out) {
    if (!out) return GEXS_EINVAL;
    *out = nullptr;
    if (!valid_profile(profile)) return GEXS_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GEXS_EHIP;
    if (device < 0 || device >= ndev) return GEXS_EINVAL;
    gexs_ctx* c = new (std::nothrow) gexs_ctx();
    if (!c) return GEXS_ENOMEM;
    c->device = device;
    c->profile = profile;
    bool ok = set_device(c) == GEXS_OK &&
              hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&c->pre, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&c->ev_pre, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->ev_main, hipEventDisableTiming) == hipSuccess &&

              next_events(c) == GEXS_OK;
    for (auto& b : c->pb)
        ok = ok && hipEventCreateWithFlags(&b.pre_done, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&b.done, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        gexs_destroy(c);
        return GEXS_EHIP;
    }
    c->tcount = 0;
    scratch_init(device);  // the device is current (set_device above)
    // the Gaussian profiles' erfc table of this device (rewritten with the
    // same values by every context created on it)
    if (gxk::init_gauss_tables(c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        gexs_destroy(c);
        return GEXS_EHIP;
    }
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            c->cus = prop.multiProcessorCount;
    }
    // defaults of the reference (Input.cpp:2-15, BaseProfile.cpp:49-54, defines.h)
    c->k.half_width = 100.0;
    c->k.temperature = 296.0;
    c->k.partition = 1.0;
    c->k.dfreq = 0.01;
    c->k.mean_mass = 18.0;
    c->k.pressure = 1.0;
    c->k.ref_temp = 296.0;
    c->k.ref_press = 1.0;
    c->k.cutoff = 0.0;
    c->k.gauss_hwhm = 0.1;
    c->k.width_conv = 0;
    c->k.pressure_mode = 1;
    c->k.precision = GEXS_FP32C;
    *out = c;
    return GEXS_OK;
}

int gexs_destroy(gexs_ctx* c) {
    if (!c) return GEXS_EINVAL;
    (void)hipSetDevice(c->device);
    c->server.reset();
    (void)hipStreamSynchronize(c->stream);
    if (c->pre) (void)hipStreamSynchronize(c->pre);
    c->strm.reset();
    for (double* p : c->results) (void)hipFree(p);
    c->results.clear();
    gxk::sort_buffers_free(c->sb);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->h_tab) (void)hipHostFree(c->h_tab);
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    if (c->pre) (void)hipStreamSynchronize(c->pre);
    if (c->ev_pre) (void)hipEventDestroy(c->ev_pre);
    if (c->ev_main) (void)hipEventDestroy(c->ev_main);
    for (auto& b : c->pb) {
        if (b.pre_done) (void)hipEventDestroy(b.pre_done);
        if (b.done) (void)hipEventDestroy(b.done);
    }
    if (c->pre) (void)hipStreamDestroy(c->pre);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GEXS_OK;
}

int gexs_init_vectors(gexs_ctx* c, int npoints, int num_cs, int64_t mem_budget,
                      int64_t* line_capacity) {
    if (!c || npoints <= 0 || num_cs <= 0) return GEXS_EINVAL;
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (double* p : c->results) (void)hipFree(p);
    c->results.assign((size_t)num_cs, nullptr);
    for (int i = 0; i < num_cs; ++i) {
        hipError_t e = hipMalloc(&c->results[i], (size_t)npoints * 8);
        if (e != hipSuccess) return c->fail(GEXS_ENOMEM, "result grids do not fit");
        HIPCHK(c, gxk::fill_f64(c->stream, c->results[i], npoints, 0.0));
    }
    c->npoints = npoints;
    c->num_cs = num_cs;
    // line capacity: resident bytes per line = sorted SoA (40) + staging (40)
    // + sort keys (12) + positions (16) + one FP32 record (32) + FP64 record (24)
    size_t free_b = 0, total_b = 0;
    HIPCHK(c, hipMemGetInfo(&free_b, &total_b));
    int64_t budget = mem_budget > 0 ? mem_budget : (int64_t)(0.9 * (double)free_b);
    int64_t per_line = 40 + 40 + 12 + 12 + 16 + 32 + 24;
    int64_t cap = budget / per_line;
    cap = std::min<int64_t>(cap, (int64_t)0x7FFFFFFF);  // GetNtrans() is held in an int (BaseProfile.h:38)
    if (cap <= 0) return c->fail(GEXS_ENOMEM, "no room for lines");
    c->capacity = cap;
    if (line_capacity) *line_capacity = cap;
    return GEXS_OK;
}

int gexs_init_constants(gexs_ctx* c, const gexs_constants* k) {
    if (!c || !k) return GEXS_EINVAL;
    if (!(k->dfreq > 0.0) || !(k->mean_mass > 0.0) || !(k->ref_temp > 0.0))
        return c->fail(GEXS_EINVAL, "dfreq, mean_mass and ref_temp must be positive");
    if (k->precision != GEXS_FP32C && k->precision != GEXS_FP64)
        return c->fail(GEXS_EINVAL, "unknown precision");
    if ((c->profile == GEXS_GAUSSIAN) && !(k->gauss_hwhm > 0.0))
        return c->fail(GEXS_EINVAL, "GAUSSIAN needs gauss_hwhm > 0");
    c->k = *k;
    c->have_k = true;
    return GEXS_OK;
}

int gexs_transfer_freq(gexs_ctx* c, const double* freq, int n) {
    if (!c || !freq || n <= 0) return GEXS_EINVAL;
    if (n != c->npoints) return c->fail(GEXS_ESTATE, "transfer_freq: n differs from init_vectors npoints");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    c->start = freq[0];
    c->have_grid = true;
    for (double* p : c->results) HIPCHK(c, gxk::fill_f64(c->stream, p, n, 0.0));
    return GEXS_OK;
}

static int transfer_common(gexs_ctx* c, int64_t n, const double* E, const double* nu,
                           const double* A, const int32_t* g, const double* gamma,
                           const double* nexp, hipMemcpyKind kind) {
    if (!c || n < 0 || (n > 0 && (!E || !nu || !A || !g))) return GEXS_EINVAL;
    if (c->server) return c->fail(GEXS_ESTATE, "lines cannot change while serving");
    if (n > 0x7FFFFFFF) return c->fail(GEXS_EINVAL, "more than 2^31-1 lines in one chunk");
    if (c->capacity > 0 && n > c->capacity) return c->fail(GEXS_ENOMEM, "chunk exceeds line capacity");
    const bool needs_gamma = (c->profile == GEXS_VOIGT || c->profile == GEXS_VOIGT_QUAD);
    const bool has_gamma = gamma != nullptr && nexp != nullptr;
    if (needs_gamma && !has_gamma)
        return c->fail(GEXS_EINVAL, "Voigt profiles need gamma and n per line");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    hipStream_t s = c->stream;
    HIPCHK(c, hipStreamSynchronize(s));  // TransferVectors waits for earlier work
    size_t nb = (size_t)std::max<int64_t>(n, 1);
    HIPCHK(c, c->r_nu.ensure(nb * 8));
    HIPCHK(c, c->r_E.ensure(nb * 8));
    HIPCHK(c, c->r_A.ensure(nb * 8));
    HIPCHK(c, c->r_g.ensure(nb * 4));
    if (n > 0) {
        HIPCHK(c, hipMemcpyAsync(c->r_nu.p, nu, (size_t)n * 8, kind, s));
        HIPCHK(c, hipMemcpyAsync(c->r_E.p, E, (size_t)n * 8, kind, s));
        HIPCHK(c, hipMemcpyAsync(c->r_A.p, A, (size_t)n * 8, kind, s));
        HIPCHK(c, hipMemcpyAsync(c->r_g.p, g, (size_t)n * 4, kind, s));
    }
    if (has_gamma) {
        HIPCHK(c, c->r_gamma.ensure(nb * 8));
        HIPCHK(c, c->r_nexp.ensure(nb * 8));
        if (n > 0) {
            HIPCHK(c, hipMemcpyAsync(c->r_gamma.p, gamma, (size_t)n * 8, kind, s));
            HIPCHK(c, hipMemcpyAsync(c->r_nexp.p, nexp, (size_t)n * 8, kind, s));
        }
    }
    if ((rc = sort_and_commit(c, n, has_gamma, false)) != GEXS_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(s));  // caller's buffers are free on return
    return GEXS_OK;
}

int gexs_transfer_lines(gexs_ctx* c, int64_t n, const double* E, const double* nu, const double* A,
                        const int32_t* g, const double* gamma, const double* nexp) {
    return transfer_common(c, n, E, nu, A, g, gamma, nexp, hipMemcpyHostToDevice);
}

int gexs_transfer_lines_device(gexs_ctx* c, int64_t n, const double* E, const double* nu,
                               const double* A, const int32_t* g, const double* gamma,
                               const double* nexp) {
    return transfer_common(c, n, E, nu, A, g, gamma, nexp, hipMemcpyDeviceToDevice);
}

int gexs_execute(gexs_ctx* c, int npoints, int64_t nlines, int start_idx, double T, double P,
                 double Q) {
    if (!c) return GEXS_EINVAL;
    if (!c->have_k || !c->have_grid) return c->fail(GEXS_ESTATE, "execute before init_constants/transfer_freq");
    if (c->tasks >= (int)c->results.size())
        return c->fail(GEXS_ESTATE, "more executes than result grids since the last transfer");
    if (nlines != c->n) return c->fail(GEXS_EINVAL, "nlines must equal the transferred chunk size");
    if (start_idx < 0 || npoints < 0 || (int64_t)start_idx + npoints > c->npoints)
        return c->fail(GEXS_EINVAL, "window outside the grid");
    if (!(T > 0.0) || !(Q > 0.0)) return c->fail(GEXS_EINVAL, "T and Q must be positive");
    int rc;
    // the FP32C passes do not join the streams (run_fp32c); the FP64 pass runs
    // on `stream` after them
    if ((rc = (c->k.precision == GEXS_FP64 ? set_device(c) : set_device_only(c))) != GEXS_OK) return rc;
    double* out = c->results[(size_t)c->tasks];
    c->tasks++;
    c->last = gexs_stats{};
    c->last.n_lines = c->n;
    if (c->n == 0 || npoints == 0) return GEXS_OK;
    const double dfreq = c->k.dfreq;
    const int win_lo = start_idx, win_hi = start_idx + npoints;
    TPParams tp = make_tp(c, T, P, Q, c->start, dfreq, win_lo, win_hi);
    if ((rc = ensure_pos(c, tp, 0, c->n)) != GEXS_OK) return rc;
    if ((rc = next_events(c)) != GEXS_OK) return rc;
    if (c->k.precision == GEXS_FP64) {
        HIPCHK(c, hipEventRecord(c->ev_start, c->stream));
        rc = run_fp64(c, tp, out, true);
    } else {
        rc = run_fp32c(c, tp, out, true);
    }
    c->last_neval_pending = true;
    return rc;
}

int gexs_transfer_results(gexs_ctx* c, double* out, int n, int cs_index, int accumulate) {
    if (!c || !out || n <= 0) return GEXS_EINVAL;
    if (cs_index < 0 || cs_index >= (int)c->results.size()) return c->fail(GEXS_EINVAL, "bad cs_index");
    if (n > c->npoints) return c->fail(GEXS_EINVAL, "n exceeds the grid");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!accumulate) {
        HIPCHK(c, hipMemcpy(out, c->results[(size_t)cs_index], (size_t)n * 8, hipMemcpyDeviceToHost));
        return GEXS_OK;
    }
    std::vector<double> tmp((size_t)n);
    HIPCHK(c, hipMemcpy(tmp.data(), c->results[(size_t)cs_index], (size_t)n * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) out[i] += tmp[(size_t)i];
    return GEXS_OK;
}

int gexs_ready(gexs_ctx* c) {
    if (!c) return GEXS_EINVAL;
    return (hipStreamQuery(c->stream) == hipSuccess && hipStreamQuery(c->pre) == hipSuccess) ? 1 : 0;
}

int gexs_sync(gexs_ctx* c) {
    if (!c) return GEXS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipStreamSynchronize(c->pre));
    return GEXS_OK;
}

const char* gexs_last_error(const gexs_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gexs_last_stats(gexs_ctx* c, gexs_stats* out) {
    if (!c || !out) return GEXS_EINVAL;
    if (c->last_neval_pending) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipStreamSynchronize(c->pre));  // the last execute pass may have run there
        unsigned long long ne = 0;
        if (c->last_neval) HIPCHK(c, hipMemcpy(&ne, c->last_neval, 8, hipMemcpyDeviceToHost));
        c->last.n_eval = (int64_t)ne;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_k0, c->ev_k1) == hipSuccess) c->last.kernel_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev_start, c->ev_k1) == hipSuccess) c->last.total_ms = ms;
        c->last_neval_pending = false;
    }
    *out = c->last;
    return GEXS_OK;
}

int gexs_timing(gexs_ctx* c, int reset, double* kernel_ms, double* total_ms, int* count) {
    if (!c) return GEXS_EINVAL;
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double ks = 0.0, ts = 0.0;
    const int n = c->tcount;
    if (!reset) {
        for (int t = 0; t < n; ++t) {
            float a = 0.f, b = 0.f;
            HIPCHK(c, hipEventElapsedTime(&a, c->evpool[3 * t + 1], c->evpool[3 * t + 2]));
            HIPCHK(c, hipEventElapsedTime(&b, c->evpool[3 * t], c->evpool[3 * t + 2]));
            ks += a;
            ts += b;
        }
    } else {
        c->tcount = 0;
    }
    if (kernel_ms) *kernel_ms = ks;
    if (total_ms) *total_ms = ts;
    if (count) *count = reset ? 0 : n;
    return GEXS_OK;
}

int gexs_line_windows(gexs_ctx* c, int64_t n, double* nu_sorted, int32_t* ib, int32_t* ie) {
    if (!c) return GEXS_EINVAL;
    if (!(c->pos_key.gen == c->gen && c->pos_key.n >= 0))
        return c->fail(GEXS_ESTATE, "no pass over the current lines yet (execute first)");
    if (n != c->pos_key.n) return c->fail(GEXS_EINVAL, "n differs from the number of lines of the last pass");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n == 0) return GEXS_OK;
    if (nu_sorted)
        HIPCHK(c, hipMemcpy(nu_sorted, c->nu.as<double>() + c->pos_key.l0, (size_t)n * 8, hipMemcpyDeviceToHost));
    if (ib || ie) {
        std::vector<LinePos> lp((size_t)n);
        HIPCHK(c, hipMemcpy(lp.data(), c->pos.p, (size_t)n * sizeof(LinePos), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) {
            if (ib) ib[i] = lp[(size_t)i].ib;
            if (ie) ie[i] = lp[(size_t)i].ie;
        }
    }
    return GEXS_OK;
}

int gexs_pass_lines(gexs_ctx* c, int64_t* first, int64_t* count) {
    if (!c || !first || !count) return GEXS_EINVAL;
    if (!(c->pos_key.gen == c->gen && c->pos_key.n >= 0))
        return c->fail(GEXS_ESTATE, "no pass over the current lines yet");
    *first = c->pos_key.l0;
    *count = c->pos_key.n;
    return GEXS_OK;
}

int gexs_pass_items(gexs_ctx* c, int64_t cap, int32_t* out, int64_t* n) {
    if (!c || !n || cap < 0 || (cap > 0 && !out)) return GEXS_EINVAL;
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t m = c->last_items ? c->last.n_tiles : 0;
    *n = m;
    if (m == 0 || cap == 0) return GEXS_OK;
    if (cap < m) return c->fail(GEXS_EINVAL, "pass_items: cap is smaller than the item count");
    // the segments the profile kernel located (its waves do not store them):
    // the same search over the same kmax, into the items' segment fields
    if (c->last_kmax) {
        HIPCHK(c, gxk::segments(c->stream, const_cast<WorkItem*>(c->last_items), (int)m,
                                c->pos.as<LinePos>(), c->last_kmax, c->profile, c->items_span));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    std::vector<WorkItem> h((size_t)m);
    HIPCHK(c, hipMemcpy(h.data(), c->last_items, (size_t)m * sizeof(WorkItem), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < m; ++i) {
        const WorkItem& w = h[(size_t)i];
        const int32_t v[14] = {w.lo, w.hi,    w.w0, w.w1, w.s1, w.s2, w.s3,
                               w.s4, w.job, w.group, w.f2, w.f3, w.e2, w.e3};
        std::memcpy(out + 14 * i, v, sizeof v);
    }
    return GEXS_OK;
}

int gexs_set_tuning(gexs_ctx* c, int points_per_lane, int order, int waves_per_tile) {
    if (!c) return GEXS_EINVAL;
    if (points_per_lane != 1 && points_per_lane != 2 && points_per_lane != 4 && points_per_lane != 8)
        return c->fail(GEXS_EINVAL, "points_per_lane must be 1, 2, 4 or 8");
    if (order < 0 || order > 3) return c->fail(GEXS_EINVAL, "order must be 0, 1, 2 or 3");
    if (waves_per_tile != 1 && waves_per_tile != 4 && waves_per_tile != 16)
        return c->fail(GEXS_EINVAL, "waves_per_tile must be 1, 4 or 16");
    c->ppl = points_per_lane;
    c->order = order;
    c->split = waves_per_tile;
    return GEXS_OK;
}

// ------------------------------------------------------------ resident list
int gexs_load_list(gexs_ctx* c, const gexs_states* st, const gexs_transitions* tr) {
    if (!c || !st || !tr || st->n <= 0 || tr->n < 0 || !st->E || !st->g) return GEXS_EINVAL;
    if (tr->n > 0 && (!tr->upper || !tr->lower || (!tr->A && !tr->A64))) return GEXS_EINVAL;
    if (c->server) return c->fail(GEXS_ESTATE, "the list cannot change while serving");
    if (tr->n > 0x7FFFFFFF) return c->fail(GEXS_EINVAL, "more than 2^31-1 transitions");
    if (st->n > 0xFFFFFFFFll) return c->fail(GEXS_EINVAL, "more than 2^32 states");
    const bool has_gamma = (tr->gamma != nullptr || tr->gamma64 != nullptr) &&
                           (tr->nexp != nullptr || tr->nexp64 != nullptr);
    if ((c->profile == GEXS_VOIGT || c->profile == GEXS_VOIGT_QUAD) && !has_gamma)
        return c->fail(GEXS_EINVAL, "Voigt profiles need gamma and n per transition");
    // indices are checked on the host before any device gather uses them
    for (int64_t i = 0; i < tr->n; ++i)
        if ((int64_t)tr->upper[i] >= st->n || (int64_t)tr->lower[i] >= st->n)
            return c->fail(GEXS_EINVAL, "transition index outside the state table");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    hipStream_t s = c->stream;
    HIPCHK(c, hipStreamSynchronize(s));
    const int64_t n = tr->n;
    const size_t nb = (size_t)std::max<int64_t>(n, 1);
    HIPCHK(c, c->st_E.ensure((size_t)st->n * 8));
    HIPCHK(c, c->st_g.ensure((size_t)st->n * 4));
    HIPCHK(c, hipMemcpyAsync(c->st_E.p, st->E, (size_t)st->n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->st_g.p, st->g, (size_t)st->n * 4, hipMemcpyHostToDevice, s));
    c->n_states = st->n;
    // stage transitions: indices in the key/perm buffers, FP32 A/gamma/n widened
    DevArr up, lo, f32;
    HIPCHK(c, up.ensure(nb * 4));
    HIPCHK(c, lo.ensure(nb * 4));
    HIPCHK(c, f32.ensure(nb * 4));
    HIPCHK(c, c->r_nu.ensure(nb * 8));
    HIPCHK(c, c->r_E.ensure(nb * 8));
    HIPCHK(c, c->r_A.ensure(nb * 8));
    HIPCHK(c, c->r_g.ensure(nb * 4));
    if (n > 0) {
        HIPCHK(c, hipMemcpyAsync(up.p, tr->upper, (size_t)n * 4, hipMemcpyHostToDevice, s));
        HIPCHK(c, hipMemcpyAsync(lo.p, tr->lower, (size_t)n * 4, hipMemcpyHostToDevice, s));
        HIPCHK(c, gxk::assemble_lines(s, c->st_E.as<double>(), c->st_g.as<int32_t>(), up.as<uint32_t>(),
                                      lo.as<uint32_t>(), n, c->r_nu.as<double>(), c->r_E.as<double>(),
                                      c->r_g.as<int32_t>()));
        // an FP64 array is copied as it is; an FP32 one is widened on the
        // device through the f32 staging buffer (one at a time: the copy
        // into f32 must wait for the widening that read it before)
        auto put = [&](const double* h64, const float* h32, DevArr& dst) -> int {
            if (h64) {
                HIPCHK(c, hipMemcpyAsync(dst.p, h64, (size_t)n * 8, hipMemcpyHostToDevice, s));
                return GEXS_OK;
            }
            HIPCHK(c, hipStreamSynchronize(s));
            HIPCHK(c, hipMemcpyAsync(f32.p, h32, (size_t)n * 4, hipMemcpyHostToDevice, s));
            HIPCHK(c, gxk::widen_f32(s, f32.as<float>(), n, dst.as<double>()));
            return GEXS_OK;
        };
        if ((rc = put(tr->A64, tr->A, c->r_A)) != GEXS_OK) return rc;
        if (has_gamma) {
            HIPCHK(c, c->r_gamma.ensure(nb * 8));
            HIPCHK(c, c->r_nexp.ensure(nb * 8));
            if ((rc = put(tr->gamma64, tr->gamma, c->r_gamma)) != GEXS_OK) return rc;
            if ((rc = put(tr->nexp64, tr->nexp, c->r_nexp)) != GEXS_OK) return rc;
        }
    }
    if ((rc = sort_and_commit(c, n, has_gamma, true)) != GEXS_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(s));
    return GEXS_OK;
}

int gexs_line_count(gexs_ctx* c, int64_t* n) {
    if (!c || !n) return GEXS_EINVAL;
    *n = c->n;
    return GEXS_OK;
}

int gexs_partition(gexs_ctx* c, const double* T, int nT, double* Q_out) {
    if (!c || !T || !Q_out || nT <= 0) return GEXS_EINVAL;
    if (c->n_states <= 0) return c->fail(GEXS_ESTATE, "no resident state table (gexs_load_list)");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    const int nblk = 512;
    HIPCHK(c, c->q_T.ensure((size_t)nT * 8));
    HIPCHK(c, c->q_part.ensure((size_t)nT * nblk * 8));
    HIPCHK(c, c->q_Q.ensure((size_t)nT * 8));
    HIPCHK(c, hipMemcpyAsync(c->q_T.p, T, (size_t)nT * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, gxk::partition(c->stream, c->st_E.as<double>(), c->st_g.as<int32_t>(), c->n_states,
                             c->q_T.as<double>(), nT, c->q_part.as<double>(), nblk, c->q_Q.as<double>()));
    HIPCHK(c, hipMemcpyAsync(Q_out, c->q_Q.p, (size_t)nT * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GEXS_OK;
}

static int pinned(gexs_ctx* c, void** p, size_t* cap, size_t want) {
    if (*p && *cap >= want) return GEXS_OK;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    want = std::max<size_t>(want, 4096);
    HIPCHK(c, hipHostMalloc(p, want, hipHostMallocDefault));
    *cap = want;
    return GEXS_OK;
}

// Device work of one FP32C batch.  Only kernels and copies from the ctx's
// pinned tables, so it can be captured into a graph: the tables are read from
// c->h_tab ([TPParams | JobDesc | Job] in the layout of c->tab, then the T of
// the device partition sums) when the copies execute.
static int enqueue_batch(gexs_ctx* c, int njobs, int nq, const PassTables& t) {
    hipStream_t s = c->stream;
    const int n = (int)c->pos_key.n;
    HIPCHK(c, hipMemcpyAsync(c->tab.p, c->h_tab, t.bytes, hipMemcpyHostToDevice, s));
    if (nq > 0) {
        const int nblk = 256;
        HIPCHK(c, hipMemcpyAsync(c->q_T.p, (char*)c->h_tab + t.bytes, (size_t)nq * 8,
                                 hipMemcpyHostToDevice, s));
        // Q(T) into the jobs' TPParams and the pass counters zeroed, one launch
        HIPCHK(c, gxk::partition(s, c->st_E.as<double>(), c->st_g.as<int32_t>(), c->n_states, c->q_T.as<double>(),
                                 nq, c->q_part.as<double>(), nblk, c->q_Q.as<double>(), t.jd, njobs, t.tp,
                                 (unsigned long long*)c->small.p, (int)t.zero_u64));
    } else {
        HIPCHK(c, gxk::fill_u64(s, (unsigned long long*)c->small.p, t.zero_u64, 0ull));
    }
    const double* gam = c->has_gamma ? c->gamma.as<double>() : nullptr;
    const double* nex = c->has_gamma ? c->nexp.as<double>() : nullptr;
    HIPCHK(c, c->rec_parts.ensure((size_t)njobs * gxk::kRecRedBytes));
    HIPCHK(c, gxk::line_rec(s, c->nu.as<double>(), c->E.as<double>(), c->A.as<double>(), c->g.as<int32_t>(),
                            gam, nex, c->pos.as<LinePos>(), t.jd, njobs, n, t.tp, t.scale,
                            c->rec.as<VRec>(), c->recx.as<VRecX>(), t.kmax, t.neval, 0, c->rec_parts.p));
    HIPCHK(c, gxk::profile_tiles(s, c->profile, c->ppl, c->split, c->items.as<WorkItem>(), c->n_items,
                                 t.jobs, t.scale, t.kmax, t.xq, c->counter.as<unsigned int>(),
                                 launch_mode(c), launch_blocks(c)));
    return GEXS_OK;
}

int gexs_run_batch(gexs_ctx* c, const gexs_tp* tp, int ntp, const gexs_grid* grid, double* out) {
    if (!c || !tp || ntp <= 0 || !grid || grid->npoints <= 0 || !(grid->end > grid->start))
        return GEXS_EINVAL;
    if (!c->have_k) return c->fail(GEXS_ESTATE, "run_batch before init_constants");
    if (c->server) return c->fail(GEXS_ESTATE, "run_batch while serving");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    const double dfreq = (grid->end - grid->start) / (double)grid->npoints;  // BaseProfile.cpp:10
    int nq = 0;
    for (int i = 0; i < ntp; ++i) {
        if (!(tp[i].T > 0.0) || !(tp[i].P >= 0.0)) return c->fail(GEXS_EINVAL, "T must be positive, P >= 0");
        if (!(tp[i].Q > 0.0)) nq++;
    }
    if (nq > 0 && c->n_states <= 0)
        return c->fail(GEXS_ESTATE, "Q(T) on the device needs a resident state table (gexs_load_list)");
    // one cutoff for the whole batch (positions and tiles are shared)
    const double cutoff = effective_cutoff(c->k, tp[0].P);
    for (int i = 1; i < ntp; ++i)
        if (effective_cutoff(c->k, tp[i].P) != cutoff)
            return c->fail(GEXS_EINVAL, "a batch needs one cutoff (set gexs_constants.cutoff > 0)");
    // line selection (Profiles/BaseProfile.cpp:152-157 as a filter on both sides):
    // start - HW <= nu <= end + HW over the nu-sorted list (nu == 0 was dropped at load)
    const double slo = grid->start - c->k.half_width, shi = grid->end + c->k.half_width;
    if (!(c->sel_lo == slo && c->sel_hi == shi && c->sel_gen == c->gen)) {
        if ((rc = select_range(c, slo, shi, &c->sel_l0, &c->sel_l1)) != GEXS_OK) return rc;
        c->sel_lo = slo;
        c->sel_hi = shi;
        c->sel_gen = c->gen;
    }
    const int64_t l0 = c->sel_l0, l1 = c->sel_l1;
    {
        TPParams ptp = make_tp(c, tp[0].T, tp[0].P, 1.0, grid->start, dfreq, 0, grid->npoints);
        if ((rc = ensure_pos(c, ptp, l0, l1 - l0)) != GEXS_OK) return rc;
    }
    const size_t np = (size_t)grid->npoints;
    const int n = (int)c->pos_key.n;
    HIPCHK(c, c->batch_out.ensure(np * (size_t)ntp * 8));
    c->last = gexs_stats{};
    c->last.n_lines = n;
    if (c->k.precision == GEXS_FP64) {
        std::vector<double> Qs((size_t)ntp);
        std::vector<double> needT;
        for (int i = 0; i < ntp; ++i)
            if (!(tp[i].Q > 0.0)) needT.push_back(tp[i].T);
        std::vector<double> q(needT.size());
        if (!needT.empty() && (rc = gexs_partition(c, needT.data(), (int)needT.size(), q.data())) != GEXS_OK)
            return rc;
        for (int i = 0, w = 0; i < ntp; ++i) Qs[i] = tp[i].Q > 0.0 ? tp[i].Q : q[(size_t)w++];
        HIPCHK(c, gxk::fill_f64(c->stream, c->batch_out.as<double>(), (int64_t)np * ntp, 0.0));
        if ((rc = next_events(c)) != GEXS_OK) return rc;
        HIPCHK(c, hipEventRecord(c->ev_start, c->stream));
        for (int i = 0; i < ntp; ++i) {
            TPParams t = make_tp(c, tp[i].T, tp[i].P, Qs[i], grid->start, dfreq, 0, grid->npoints);
            if ((rc = run_fp64(c, t, c->batch_out.as<double>() + (size_t)i * np, i == 0, i == 0)) != GEXS_OK)
                return rc;
        }
        HIPCHK(c, hipEventRecord(c->ev_k1, c->stream));
        if (out) HIPCHK(c, hipMemcpyAsync(out, c->batch_out.p, np * (size_t)ntp * 8, hipMemcpyDefault, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        unsigned long long ne = 0;
        HIPCHK(c, hipMemcpy(&ne, c->neval.p, 8, hipMemcpyDeviceToHost));
        c->last.n_eval = (int64_t)ne;  // summed over the batch
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_k0, c->ev_k1) == hipSuccess) c->last.kernel_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev_start, c->ev_k1) == hipSuccess) c->last.total_ms = ms;
        c->last_neval_pending = false;
        return GEXS_OK;
    }
    // ---- FP32C: host tables in pinned memory
    const int per = tiles_per_job(c);
    PassTables t;
    if ((rc = pass_tables(c, ntp, &t)) != GEXS_OK) return rc;
    if ((rc = pinned(c, &c->h_tab, &c->h_tab_cap, t.bytes + (size_t)ntp * 8)) != GEXS_OK) return rc;
    TPParams* htp = (TPParams*)c->h_tab;
    JobDesc* hjd = (JobDesc*)((char*)c->h_tab + t.jd_off);
    Job* hjobs = (Job*)((char*)c->h_tab + t.jobs_off);
    double* hT = (double*)((char*)c->h_tab + t.bytes);
    HIPCHK(c, c->rec.ensure((size_t)std::max(n, 1) * ntp * sizeof(VRec)));
    HIPCHK(c, c->recx.ensure((size_t)std::max(n, 1) * ntp * sizeof(VRecX)));
    HIPCHK(c, c->rec_parts.ensure((size_t)ntp * gxk::kRecRedBytes));  // before any capture
    HIPCHK(c, c->counter.ensure(256));
    HIPCHK(c, c->q_T.ensure((size_t)ntp * 8));
    HIPCHK(c, c->q_part.ensure((size_t)ntp * 256 * 8));
    HIPCHK(c, c->q_Q.ensure((size_t)ntp * 8));
    // previous graph launches have completed (run_batch is synchronous), so
    // the pinned tables may be rewritten
    int qs = 0;
    for (int j = 0; j < ntp; ++j) {
        htp[j] = make_tp(c, tp[j].T, tp[j].P, tp[j].Q > 0.0 ? tp[j].Q : 1.0, grid->start, dfreq, 0, grid->npoints);
        JobDesc d{};
        d.l0 = l0;
        d.pos_off = 0;
        d.rec_off = (int64_t)j * n;
        d.recx_off = (int64_t)j * n;
        d.n = n;
        d.items_off = j * per;
        d.ntiles = per;
        d.q_slot = tp[j].Q > 0.0 ? -1 : qs;
        if (d.q_slot >= 0) hT[qs++] = tp[j].T;
        hjd[j] = d;
        Job jb{};
        jb.pos = c->pos.as<LinePos>();
        jb.rec = c->rec.as<VRec>() + (size_t)j * n;
        jb.recx = c->recx.as<VRecX>() + (size_t)j * n;
        jb.out = c->batch_out.as<double>() + (size_t)j * np;
        jb.n = n;
        jb.profile = c->profile;
        jb.scale_slot = j;
        jb.accumulate = 0;  // every point of every grid is written once
        hjobs[j] = jb;
    }
    // the tiles kernel reads window and job tables from the device copy
    HIPCHK(c, hipMemcpyAsync(c->tab.p, c->h_tab, t.bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = ensure_items(c, ntp, htp, hjd)) != GEXS_OK) return rc;
    if ((rc = next_events(c)) != GEXS_OK) return rc;
    HIPCHK(c, hipEventRecord(c->ev_start, c->stream));
    if (c->use_graph) {
        // everything the captured work depends on: launch shape, sizes and
        // every device address it reads or writes (a buffer that grew since
        // the capture has moved)
        char key[512];
        std::snprintf(key, sizeof key,
                      "%d|%d|%llu|%d|%d|%d|%d|%lld|%lld|%d|%d|%a|%a|%a|%p|%p|%p|%p|%p|%p|%p|%p|%p|"
                      "%p|%p|%p|%p|%p",
                      ntp, qs, (unsigned long long)c->gen, c->ppl, c->split, c->order,
                      c->persistent_blocks, (long long)l0, (long long)n, grid->npoints, per,
                      grid->start, grid->end, cutoff, c->rec.p, c->recx.p, c->batch_out.p,
                      c->items.p, c->pos.p, c->tab.p, c->small.p, c->counter.p, c->h_tab,
                      c->st_E.p, c->st_g.p, c->nu.p, c->q_T.p, c->q_part.p);
        const std::string k2 = std::string(key) + "|" + std::to_string((uintptr_t)c->q_Q.p) + "|" +
                               std::to_string((uintptr_t)c->rec_parts.p);
        if (!c->graph_exec || c->graph_key != k2) {
            if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
            if (c->graph) (void)hipGraphDestroy(c->graph);
            c->graph_exec = nullptr;
            c->graph = nullptr;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_batch(c, ntp, qs, t);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc != GEXS_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return c->hipfail(e, "hipStreamEndCapture");
            c->graph = g;
            HIPCHK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
            c->graph_key = k2;
        }
        HIPCHK(c, hipEventRecord(c->ev_k0, c->stream));
        HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    } else {
        HIPCHK(c, hipEventRecord(c->ev_k0, c->stream));
        if ((rc = enqueue_batch(c, ntp, qs, t)) != GEXS_OK) return rc;
    }
    HIPCHK(c, hipEventRecord(c->ev_k1, c->stream));
    c->last.n_tiles = c->n_items;
    c->last_items = c->items.as<WorkItem>();
    c->last_kmax = t.kmax;
    if (out) HIPCHK(c, hipMemcpyAsync(out, c->batch_out.p, np * (size_t)ntp * 8, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<unsigned long long> ne((size_t)ntp, 0ull);
    HIPCHK(c, hipMemcpy(ne.data(), t.neval, ne.size() * 8, hipMemcpyDeviceToHost));
    int64_t tot = 0;
    for (auto v : ne) tot += (int64_t)v;
    c->last.n_eval = tot;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev_k0, c->ev_k1) == hipSuccess) c->last.kernel_ms = ms;
    if (hipEventElapsedTime(&ms, c->ev_start, c->ev_k1) == hipSuccess) c->last.total_ms = ms;
    c->last_neval_pending = false;
    return GEXS_OK;
}

int gexs_batch_device_out(gexs_ctx* c, const double** out, int64_t* count) {
    if (!c || !out) return GEXS_EINVAL;
    *out = c->batch_out.as<double>();
    if (count) *count = (int64_t)(c->batch_out.bytes / 8);
    return GEXS_OK;
}

// ------------------------------------------------------- streamed host list
// Column layout of one chunk (pinned staging and device copy alike):
// upper | lower | A | gamma | n, each at a 256-B aligned offset, for `cap`
// transitions.
struct ChunkLayout {
    size_t up, lo, A, gam, nex, bytes;
    int asz, gsz, nsz;
};

static ChunkLayout chunk_layout(const gexs_transitions* tr, bool has_gamma, int64_t cap) {
    ChunkLayout L{};
    L.asz = tr->A64 ? 8 : 4;
    L.gsz = has_gamma ? (tr->gamma64 ? 8 : 4) : 0;
    L.nsz = has_gamma ? (tr->nexp64 ? 8 : 4) : 0;
    const size_t n = (size_t)cap;
    L.up = 0;
    L.lo = al256(n * 4);
    L.A = L.lo + al256(n * 4);
    L.gam = L.A + al256(n * L.asz);
    L.nex = L.gam + al256(n * L.gsz);
    L.bytes = L.nex + al256(n * L.nsz);
    return L;
}

// Copies transitions [a, a + m) of tr into the staging columns at h
// (threads share the range) and returns the largest state index among them.
static uint32_t stage_chunk(const gexs_transitions* tr, bool has_gamma, const ChunkLayout& L, char* h,
                            int64_t a, int64_t m) {
    const int64_t per_thread = 1 << 20;
    int nthr = (int)std::min<int64_t>((m + per_thread - 1) / per_thread,
                                      std::max(1u, std::min(8u, std::thread::hardware_concurrency())));
    nthr = std::max(nthr, 1);
    std::vector<uint32_t> mx((size_t)nthr, 0u);
    auto work = [&](int t) {
        const int64_t b = a + m * t / nthr, e = a + m * (t + 1) / nthr, o = b - a;
        uint32_t hi = 0;
        uint32_t* up = (uint32_t*)(h + L.up) + o;
        uint32_t* lo = (uint32_t*)(h + L.lo) + o;
        for (int64_t i = b; i < e; ++i) {
            const uint32_t u = tr->upper[i], l = tr->lower[i];
            up[i - b] = u;
            lo[i - b] = l;
            hi = std::max(hi, std::max(u, l));
        }
        mx[(size_t)t] = hi;
        auto col = [&](size_t off, int sz, const void* p32, const void* p64) {
            const void* src = sz == 8 ? p64 : p32;
            std::memcpy(h + off + (size_t)o * sz, (const char*)src + (size_t)b * sz, (size_t)(e - b) * sz);
        };
        col(L.A, L.asz, tr->A, tr->A64);
        if (has_gamma) {
            col(L.gam, L.gsz, tr->gamma, tr->gamma64);
            col(L.nex, L.nsz, tr->nexp, tr->nexp64);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthr; ++t) {
        try {
            th.emplace_back(work, t);
        } catch (...) {
            work(t);  // no thread: this one does the share
        }
    }
    work(0);
    for (auto& x : th) x.join();
    uint32_t hi = 0;
    for (uint32_t v : mx) hi = std::max(hi, v);
    return hi;
}

// Joins the staging thread on every way out of gexs_stream_xsec.
struct Joiner {
    std::thread t;
    ~Joiner() {
        if (t.joinable()) t.join();
    }
};

// The current lines of c are those of sorted chunk set `set` while one
// exists (the pass launchers read them at enqueue time).
struct ChunkLines {
    gexs_ctx* c;
    StreamBufs* S;
    int set;
    int64_t n_saved;
    bool gamma_saved;
    ChunkLines(gexs_ctx* c_, StreamBufs* S_, int set_, int64_t n, bool has_gamma)
        : c(c_), S(S_), set(set_), n_saved(c_->n), gamma_saved(c_->has_gamma) {
        swap_all();
        c->n = n;
        c->has_gamma = has_gamma;
        c->gen++;
    }
    ~ChunkLines() {
        swap_all();
        c->n = n_saved;
        c->has_gamma = gamma_saved;
        c->gen++;  // caches built on the chunk's lines are stale
    }
    void swap_all() {
        swap_arr(c->nu, S->nu[set]);
        swap_arr(c->E, S->E[set]);
        swap_arr(c->A, S->A[set]);
        swap_arr(c->gamma, S->gamma[set]);
        swap_arr(c->nexp, S->nexp[set]);
        swap_arr(c->g, S->g[set]);
    }
};

// The pass of one chunk (its n selected lines are the current lines) on
// c->stream: every (T,P) adds into its grid of c->batch_out.  FP32C: the
// counters before the N_eval words are cleared, N_eval adds up over the chunks.
static int stream_chunk_pass(gexs_ctx* c, const std::vector<TPParams>& htp, int64_t n, int np) {
    const int ntp = (int)htp.size();
    hipStream_t s = c->stream;
    int rc;
    if ((rc = ensure_pos(c, htp[0], 0, n)) != GEXS_OK) return rc;
    if (c->k.precision == GEXS_FP64) {
        for (int j = 0; j < ntp; ++j)
            if ((rc = run_fp64(c, htp[(size_t)j], c->batch_out.as<double>() + (size_t)j * np, false, false)) !=
                GEXS_OK)
                return rc;
        return GEXS_OK;
    }
    const int per = tiles_per_job(c);
    PassTables t;
    if ((rc = pass_tables(c, ntp, &t)) != GEXS_OK) return rc;
    std::vector<char> host(t.bytes);
    TPParams* htab = (TPParams*)host.data();
    JobDesc* hjd = (JobDesc*)(host.data() + t.jd_off);
    Job* hjobs = (Job*)(host.data() + t.jobs_off);
    for (int j = 0; j < ntp; ++j) {
        htab[j] = htp[(size_t)j];
        JobDesc d{};
        d.l0 = 0;
        d.pos_off = 0;
        d.rec_off = (int64_t)j * n;
        d.recx_off = (int64_t)j * n;
        d.n = (int32_t)n;
        d.items_off = j * per;
        d.ntiles = per;
        d.q_slot = -1;
        hjd[j] = d;
        Job jb{};
        jb.pos = c->pos.as<LinePos>();
        jb.rec = c->rec.as<VRec>() + (size_t)j * n;
        jb.recx = c->recx.as<VRecX>() + (size_t)j * n;
        jb.out = c->batch_out.as<double>() + (size_t)j * np;
        jb.n = (int32_t)n;
        jb.profile = c->profile;
        jb.scale_slot = j;
        jb.accumulate = 1;  // the chunks add up
        hjobs[j] = jb;
    }
    HIPCHK(c, c->staging.upload(c->tab.p, host.data(), t.bytes, s));
    if ((rc = ensure_items(c, ntp, htab, hjd)) != GEXS_OK) return rc;
    HIPCHK(c, gxk::fill_u64(s, (unsigned long long*)c->small.p, 4 + (int64_t)ntp, 0ull));  // xq | scale
    HIPCHK(c, gxk::fill_u64(s, (unsigned long long*)t.kmax, ntp, 0ull));
    const double* gam = c->has_gamma ? c->gamma.as<double>() : nullptr;
    const double* nex = c->has_gamma ? c->nexp.as<double>() : nullptr;
    HIPCHK(c, c->rec_parts.ensure((size_t)ntp * gxk::kRecRedBytes));
    HIPCHK(c, gxk::line_rec(s, c->nu.as<double>(), c->E.as<double>(), c->A.as<double>(), c->g.as<int32_t>(), gam,
                            nex, c->pos.as<LinePos>(), t.jd, ntp, n, t.tp, t.scale, c->rec.as<VRec>(),
                            c->recx.as<VRecX>(), t.kmax, t.neval, 0, c->rec_parts.p));
    HIPCHK(c, gxk::profile_tiles(s, c->profile, c->ppl, c->split, c->items.as<WorkItem>(), c->n_items, t.jobs,
                                 t.scale, t.kmax, t.xq, c->counter.as<unsigned int>(), launch_mode(c),
                                 launch_blocks(c)));
    c->last.n_tiles = c->n_items;
    c->last_items = c->items.as<WorkItem>();
    c->last_kmax = t.kmax;
    return GEXS_OK;
}

int gexs_stream_xsec(gexs_ctx* c, const gexs_states* st, const gexs_transitions* tr, const gexs_tp* tp,
                     int ntp, const gexs_grid* grid, int64_t mem_budget, double* out,
                     gexs_stream_stats* stats) {
    const auto wall0 = std::chrono::steady_clock::now();
    if (!c || !st || !tr || !tp || ntp <= 0 || !grid || grid->npoints <= 0 || !(grid->end > grid->start) ||
        mem_budget <= 0)
        return GEXS_EINVAL;
    if (st->n <= 0 || !st->E || !st->g || tr->n < 0) return GEXS_EINVAL;
    if (tr->n > 0 && (!tr->upper || !tr->lower || (!tr->A && !tr->A64))) return GEXS_EINVAL;
    if (!c->have_k) return c->fail(GEXS_ESTATE, "stream_xsec before init_constants");
    if (c->server) return c->fail(GEXS_ESTATE, "stream_xsec while serving");
    if (st->n > 0xFFFFFFFFll) return c->fail(GEXS_EINVAL, "more than 2^32 states");
    const bool has_gamma = (tr->gamma != nullptr || tr->gamma64 != nullptr) &&
                           (tr->nexp != nullptr || tr->nexp64 != nullptr);
    if ((c->profile == GEXS_VOIGT || c->profile == GEXS_VOIGT_QUAD) && !has_gamma)
        return c->fail(GEXS_EINVAL, "Voigt profiles need gamma and n per transition");
    for (int i = 0; i < ntp; ++i)
        if (!(tp[i].T > 0.0) || !(tp[i].P >= 0.0)) return c->fail(GEXS_EINVAL, "T must be positive, P >= 0");
    const double cutoff = effective_cutoff(c->k, tp[0].P);
    for (int i = 1; i < ntp; ++i)
        if (effective_cutoff(c->k, tp[i].P) != cutoff)
            return c->fail(GEXS_EINVAL, "a batch needs one cutoff (set gexs_constants.cutoff > 0)");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipStreamSynchronize(c->pre));
    const bool f64 = c->k.precision == GEXS_FP64;
    const int np = grid->npoints;
    const double dfreq = (grid->end - grid->start) / (double)np;  // BaseProfile.cpp:10
    const double slo = grid->start - c->k.half_width, shi = grid->end + c->k.half_width;
    const int64_t N = tr->n;

    // chunk size under the budget: per transition two staging/raw sets, the
    // sort keys, two sorted sets and the pass's windows and records
    const ChunkLayout L1 = chunk_layout(tr, has_gamma, 1);
    const int64_t raw_b = 8 + L1.asz + L1.gsz + L1.nsz;
    const int64_t sorted_b = 28 + (has_gamma ? 16 : 0);
    const int64_t sort_b = 25;  // keys, values, their second buffers, digit counts
    const int64_t pass_b = (int64_t)sizeof(LinePos) +
                           (f64 ? (int64_t)sizeof(LineF64) : (int64_t)ntp * (int64_t)(sizeof(VRec) + sizeof(VRecX)));
    const int64_t per_tr = 2 * raw_b + 2 * sorted_b + sort_b + pass_b;
    const int64_t tiles = (int64_t)((np + GX_WAVE * c->ppl - 1) / (GX_WAVE * c->ppl)) * ntp;
    const int64_t fixed = st->n * 12 + (int64_t)ntp * np * 8 + tiles * (int64_t)(2 * sizeof(WorkItem) + 17) +
                          (int64_t)ntp * 8 * 256 + (4 << 20);
    // at least 4096 transitions per chunk (or the whole list, if shorter)
    const int64_t min_chunk = std::min<int64_t>(std::max<int64_t>(N, 1), 4096);
    if (mem_budget < fixed + per_tr * min_chunk)
        return c->fail(GEXS_ENOMEM, "mem_budget below " + std::to_string(fixed + per_tr * min_chunk) +
                                        " bytes (" + std::to_string(min_chunk) + " transitions per chunk)");
    int64_t cap = (mem_budget - fixed) / per_tr;
    cap = std::min<int64_t>(cap, std::max<int64_t>(N, 1));
    cap = std::min<int64_t>(cap, 0x7FFFFFFF);
    const int64_t nchunks = N == 0 ? 0 : (N + cap - 1) / cap;
    const ChunkLayout L = chunk_layout(tr, has_gamma, cap);

    // buffers, all sized before the first chunk (nothing is freed while a
    // chunk is in flight)
    if (!c->strm) {
        c->strm.reset(new (std::nothrow) StreamBufs());
        if (!c->strm) return GEXS_ENOMEM;
        hipError_t e = c->strm->init();
        if (e != hipSuccess) {
            c->strm.reset();
            return c->hipfail(e, "stream buffers");
        }
    }
    StreamBufs* S = c->strm.get();
    HIPCHK(c, hipStreamSynchronize(S->cp));
    hipStream_t s = c->stream, P = c->pre, cp = S->cp;
    HIPCHK(c, S->st_E.ensure((size_t)st->n * 8));
    HIPCHK(c, S->st_g.ensure((size_t)st->n * 4));
    HIPCHK(c, hipMemcpyAsync(S->st_E.p, st->E, (size_t)st->n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(S->st_g.p, st->g, (size_t)st->n * 4, hipMemcpyHostToDevice, s));
    const size_t cb = (size_t)cap;
    for (int i = 0; i < 2; ++i) {
        HIPCHK(c, S->raw[i].ensure(L.bytes));
        HIPCHK(c, S->nu[i].ensure(cb * 8));
        HIPCHK(c, S->E[i].ensure(cb * 8));
        HIPCHK(c, S->A[i].ensure(cb * 8));
        HIPCHK(c, S->g[i].ensure(cb * 4));
        if (has_gamma) {
            HIPCHK(c, S->gamma[i].ensure(cb * 8));
            HIPCHK(c, S->nexp[i].ensure(cb * 8));
        }
        if (S->h_bytes < L.bytes) {
            if (S->h[i]) (void)hipHostFree(S->h[i]);
            S->h[i] = nullptr;
            HIPCHK(c, hipHostMalloc(&S->h[i], L.bytes, hipHostMallocDefault));
        }
    }
    S->h_bytes = std::max(S->h_bytes, L.bytes);
    HIPCHK(c, S->keys.ensure(cb * 8));
    HIPCHK(c, S->vals.ensure(cb * 4));
    HIPCHK(c, S->dropped.ensure(64));
    HIPCHK(c, gxk::sort_buffers_alloc(S->sb, (int64_t)cb, 8));
    HIPCHK(c, c->pos.ensure(cb * sizeof(LinePos)));
    if (f64) {
        HIPCHK(c, c->lf64.ensure(cb * sizeof(LineF64)));
        HIPCHK(c, c->neval.ensure(8));
        HIPCHK(c, gxk::fill_u64(s, c->neval.as<unsigned long long>(), 1, 0ull));
    } else {
        HIPCHK(c, c->rec.ensure(cb * (size_t)ntp * sizeof(VRec)));
        HIPCHK(c, c->recx.ensure(cb * (size_t)ntp * sizeof(VRecX)));
        HIPCHK(c, c->rec_parts.ensure((size_t)ntp * gxk::kRecRedBytes));
        HIPCHK(c, c->counter.ensure(256));
        PassTables t;
        if ((rc = pass_tables(c, ntp, &t)) != GEXS_OK) return rc;
        HIPCHK(c, gxk::fill_u64(s, (unsigned long long*)c->small.p, t.zero_u64, 0ull));
    }
    HIPCHK(c, c->batch_out.ensure((size_t)np * ntp * 8));
    HIPCHK(c, gxk::fill_f64(s, c->batch_out.as<double>(), (int64_t)np * ntp, 0.0));

    // (T,P) tables; Q(T) from the streamed states where not given
    std::vector<TPParams> htp((size_t)ntp);
    {
        std::vector<double> needT;
        for (int i = 0; i < ntp; ++i)
            if (!(tp[i].Q > 0.0)) needT.push_back(tp[i].T);
        std::vector<double> q(needT.size());
        if (!needT.empty()) {
            const int nq = (int)needT.size(), nblk = 512;
            HIPCHK(c, S->q_T.ensure((size_t)nq * 8));
            HIPCHK(c, S->q_part.ensure((size_t)nq * nblk * 8));
            HIPCHK(c, S->q_Q.ensure((size_t)nq * 8));
            HIPCHK(c, hipMemcpyAsync(S->q_T.p, needT.data(), (size_t)nq * 8, hipMemcpyHostToDevice, s));
            HIPCHK(c, gxk::partition(s, S->st_E.as<double>(), S->st_g.as<int32_t>(), st->n, S->q_T.as<double>(), nq,
                                     S->q_part.as<double>(), nblk, S->q_Q.as<double>()));
            HIPCHK(c, hipMemcpyAsync(q.data(), S->q_Q.p, (size_t)nq * 8, hipMemcpyDeviceToHost, s));
        }
        HIPCHK(c, hipStreamSynchronize(s));
        for (int i = 0, w = 0; i < ntp; ++i)
            htp[(size_t)i] = make_tp(c, tp[i].T, tp[i].P, tp[i].Q > 0.0 ? tp[i].Q : q[(size_t)w++], grid->start,
                                     dfreq, 0, np);
    }

    while ((int64_t)S->tev.size() < 6 * nchunks) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        S->tev.push_back(e);
    }
    auto ev = [&](int64_t k, int i) { return S->tev[(size_t)(6 * k + i)]; };

    // the pipeline
    c->last = gexs_stats{};
    double stage_ms = 0.0;
    int64_t n_lines = 0;
    bool h2d_used[2] = {false, false};
    int stage_rc = GEXS_OK;
    Joiner ld;
    auto start_stage = [&](int64_t k) {
        const int set = (int)(k & 1);
        const bool wait = h2d_used[set];
        // host copy of chunk k into staging set k mod 2 (free once the DMA of
        // chunk k - 2 has read it), then its DMA into raw set k mod 2 (free
        // once the gather of chunk k - 2 has read it), enqueued from here
        auto job = [&, k, set, wait]() {
            if (hipSetDevice(c->device) != hipSuccess ||
                (wait && hipEventSynchronize(S->h2d_done[set]) != hipSuccess)) {
                stage_rc = GEXS_EHIP;
                return;
            }
            const auto h0 = std::chrono::steady_clock::now();
            const int64_t a = k * cap, m = std::min(cap, N - a);
            const uint32_t hi = stage_chunk(tr, has_gamma, L, (char*)S->h[set], a, m);
            stage_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
            if ((int64_t)hi >= st->n) {  // checked before the chunk reaches the device
                stage_rc = GEXS_EINVAL;
                return;
            }
            char* dev = (char*)S->raw[set].p;
            const char* h = (const char*)S->h[set];
            bool ok = k < 2 || hipStreamWaitEvent(cp, S->gathered[set], 0) == hipSuccess;
            if (k == 0) ok = ok && hipEventRecord(S->t0, cp) == hipSuccess;
            ok = ok && hipEventRecord(ev(k, 0), cp) == hipSuccess;
            auto put = [&](size_t off, int sz) {
                ok = ok && hipMemcpyAsync(dev + off, h + off, (size_t)m * sz, hipMemcpyHostToDevice, cp) == hipSuccess;
            };
            put(L.up, 4);
            put(L.lo, 4);
            put(L.A, L.asz);
            if (has_gamma) {
                put(L.gam, L.gsz);
                put(L.nex, L.nsz);
            }
            ok = ok && hipEventRecord(ev(k, 1), cp) == hipSuccess;
            ok = ok && hipEventRecord(S->h2d_done[set], cp) == hipSuccess;
            h2d_used[set] = true;
            stage_rc = ok ? GEXS_OK : GEXS_EHIP;
        };
        try {
            ld.t = std::thread(job);
        } catch (...) {
            job();
        }
    };
    if (nchunks > 0) start_stage(0);
    else HIPCHK(c, hipEventRecord(S->t0, s));
    for (int64_t k = 0; k < nchunks; ++k) {
        const int set = (int)(k & 1);
        const int64_t m = std::min(cap, N - k * cap);
        if (ld.t.joinable()) ld.t.join();
        if (stage_rc == GEXS_EINVAL) return c->fail(GEXS_EINVAL, "transition index outside the state table");
        if (stage_rc != GEXS_OK) return c->fail(stage_rc, "staging or copying a chunk");
        char* dev = (char*)S->raw[set].p;
        // chunk k + 1 is staged and copied meanwhile
        if (k + 1 < nchunks) start_stage(k + 1);
        // keys, sort and selection of chunk k (pre stream)
        const uint32_t* up = (const uint32_t*)(dev + L.up);
        const uint32_t* lo = (const uint32_t*)(dev + L.lo);
        HIPCHK(c, hipStreamWaitEvent(P, S->h2d_done[set], 0));
        HIPCHK(c, hipEventRecord(ev(k, 2), P));
        HIPCHK(c, gxk::fill_u64(P, S->dropped.as<unsigned long long>(), 2, 0ull));
        HIPCHK(c, gxk::chunk_keys(P, S->st_E.as<double>(), up, lo, m, slo, shi, S->keys.as<uint64_t>(),
                                  S->vals.as<uint32_t>(), S->dropped.as<unsigned long long>()));
        // the selected count and the digits the sort passes over: the one
        // wait of the chunk (the sort and gather then run unattended)
        HIPCHK(c, hipMemcpyAsync(S->h_small, S->dropped.p, 16, hipMemcpyDeviceToHost, P));
        HIPCHK(c, hipStreamSynchronize(P));
        const int64_t kept = m - (int64_t)S->h_small[0];
        const unsigned long long diff = S->h_small[1];
        HIPCHK(c, gxk::radix_sort_u64(P, S->keys.as<uint64_t>(), S->vals.as<uint32_t>(), m, S->sb, 0, 64, &diff));
        // the selected lines, by nu, into sorted set k mod 2 once pass k - 2 has read it
        if (k >= 2) HIPCHK(c, hipStreamWaitEvent(P, S->pass_done[set], 0));
        gxk::ChunkCols cols{};
        if (L.asz == 8) cols.A64 = (const double*)(dev + L.A); else cols.A32 = (const float*)(dev + L.A);
        if (has_gamma) {
            if (L.gsz == 8) cols.gam64 = (const double*)(dev + L.gam); else cols.gam32 = (const float*)(dev + L.gam);
            if (L.nsz == 8) cols.nex64 = (const double*)(dev + L.nex); else cols.nex32 = (const float*)(dev + L.nex);
        }
        HIPCHK(c, gxk::chunk_gather(P, S->vals.as<uint32_t>(), kept, up, lo, S->st_E.as<double>(),
                                    S->st_g.as<int32_t>(), cols, S->nu[set].as<double>(), S->E[set].as<double>(),
                                    S->g[set].as<int32_t>(), S->A[set].as<double>(),
                                    has_gamma ? S->gamma[set].as<double>() : nullptr,
                                    has_gamma ? S->nexp[set].as<double>() : nullptr));
        HIPCHK(c, hipEventRecord(ev(k, 3), P));
        HIPCHK(c, hipEventRecord(S->gathered[set], P));
        // the pass of chunk k (main stream)
        HIPCHK(c, hipStreamWaitEvent(s, S->gathered[set], 0));
        HIPCHK(c, hipEventRecord(ev(k, 4), s));
        if (kept > 0) {
            ChunkLines cur(c, S, set, kept, has_gamma);
            if ((rc = stream_chunk_pass(c, htp, kept, np)) != GEXS_OK) return rc;
        }
        HIPCHK(c, hipEventRecord(ev(k, 5), s));
        HIPCHK(c, hipEventRecord(S->pass_done[set], s));
        n_lines += kept;
    }
    HIPCHK(c, hipEventRecord(S->t1, s));
    if (out) HIPCHK(c, hipMemcpyAsync(out, c->batch_out.p, (size_t)np * ntp * 8, hipMemcpyDefault, s));
    HIPCHK(c, hipStreamSynchronize(s));
    int64_t ne = 0;
    if (f64) {
        unsigned long long v = 0;
        HIPCHK(c, hipMemcpy(&v, c->neval.p, 8, hipMemcpyDeviceToHost));
        ne = (int64_t)v;
    } else {
        PassTables t;
        if ((rc = pass_tables(c, ntp, &t)) != GEXS_OK) return rc;
        std::vector<unsigned long long> v((size_t)ntp, 0ull);
        HIPCHK(c, hipMemcpy(v.data(), t.neval, v.size() * 8, hipMemcpyDeviceToHost));
        for (auto x : v) ne += (int64_t)x;
    }
    float dms = 0.f;
    if (hipEventElapsedTime(&dms, S->t0, S->t1) != hipSuccess) dms = 0.f;
    double stage_sum[3] = {0.0, 0.0, 0.0};  // copy, sort, pass
    for (int64_t k = 0; k < nchunks; ++k)
        for (int i = 0; i < 3; ++i) {
            float x = 0.f;
            if (hipEventElapsedTime(&x, ev(k, 2 * i), ev(k, 2 * i + 1)) == hipSuccess) stage_sum[i] += x;
        }
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    c->last.n_eval = ne;
    c->last.n_lines = n_lines;
    c->last.kernel_ms = dms;
    c->last.total_ms = wall;
    c->last_neval_pending = false;
    if (stats) {
        *stats = gexs_stream_stats{};
        stats->n_transitions = N;
        stats->n_lines = n_lines;
        stats->n_eval = ne;
        stats->chunk = cap;
        stats->device_bytes = fixed + per_tr * cap;
        stats->n_chunks = (int32_t)nchunks;
        stats->wall_ms = wall;
        stats->device_ms = dms;
        stats->stage_ms = stage_ms;
        stats->copy_ms = stage_sum[0];
        stats->sort_ms = stage_sum[1];
        stats->pass_ms = stage_sum[2];
    }
    return GEXS_OK;
}

int gexs_set_launch(gexs_ctx* c, int persistent_blocks, int use_graph) {
    if (!c || persistent_blocks < 0) return GEXS_EINVAL;
    c->persistent_blocks = persistent_blocks;
    c->use_graph = use_graph != 0;
    return GEXS_OK;
}

int gexs_autotune(gexs_ctx* c, const gexs_tp* tp, int ntp, const gexs_grid* grid, gexs_tuning* best,
                  gexs_tuning* report, int report_cap, int* n_report) {
    if (!c || !tp || ntp <= 0 || !grid || grid->npoints <= 0) return GEXS_EINVAL;
    if (c->k.precision != GEXS_FP32C) return c->fail(GEXS_ESTATE, "autotune applies to the FP32C path");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
    const int cus = std::max(1, prop.multiProcessorCount);
    DevArr tmp;
    HIPCHK(c, tmp.ensure((size_t)grid->npoints * ntp * 8));
    const int ppls[4] = {1, 2, 4, 8};
    const int pers[3] = {0, 4 * cus, 8 * cus};
    const bool graph = c->use_graph;
    c->use_graph = false;
    gexs_tuning top{};
    top.kernel_ms = 1e300;
    int nrep = 0;
    for (int pi = 0; pi < 4; ++pi)
        for (int split = 1; split <= 16; split *= 4)
        for (int order = 0; order < 4; ++order)
            for (int pb = 0; pb < 3; ++pb) {
                if (order == 2 && pb == 0) continue;  // per-XCD queues are persistent
                if (order == 3 && pb != 0) continue;  // dispatch order applies to launch slots
                c->ppl = ppls[pi];
                c->split = split;
                c->order = order;
                c->persistent_blocks = pers[pb];
                double ms = 0.0;
                for (int rep = 0; rep < 2; ++rep) {  // first run builds positions and tiles
                    if ((rc = gexs_run_batch(c, tp, ntp, grid, tmp.as<double>())) != GEXS_OK) {
                        c->use_graph = graph;
                        return rc;
                    }
                    ms = c->last.kernel_ms;
                }
                gexs_tuning t{};
                t.points_per_lane = ppls[pi];
                t.waves_per_tile = split;
                t.order = order;
                t.persistent_blocks = pers[pb];
                t.kernel_ms = ms;
                if (report && nrep < report_cap) report[nrep] = t;
                nrep++;
                if (ms < top.kernel_ms) top = t;
            }
    c->ppl = top.points_per_lane;
    c->split = top.waves_per_tile;
    c->order = top.order;
    c->persistent_blocks = top.persistent_blocks;
    c->use_graph = graph;
    if (best) *best = top;
    if (n_report) *n_report = nrep;
    return GEXS_OK;
}

// ------------------------------------------------------ no-allocation pass
int gexs_run_fixed(gexs_ctx* c, const gexs_lines_dev* L, const gexs_tp* tp, const gexs_grid* grid,
                   double* d_out, int64_t* n_eval) {
    if (!c || !L || !tp || !grid || !d_out) return GEXS_EINVAL;
    if (!c->have_k) return c->fail(GEXS_ESTATE, "run_fixed before init_constants");
    if (c->k.precision != GEXS_FP32C) return c->fail(GEXS_EINVAL, "run_fixed runs the FP32C path only");
    const int64_t n = L->n;
    if (n < 0 || (n > 0 && (!L->nu || !L->E || !L->A || !L->g))) return GEXS_EINVAL;
    const bool voigt = c->profile == GEXS_VOIGT || c->profile == GEXS_VOIGT_QUAD;
    if (voigt && n > 0 && (!L->gamma || !L->nexp))
        return c->fail(GEXS_EINVAL, "Voigt profiles need gamma and n per line");
    if (n > kFixedMaxLines)
        return c->fail(GEXS_EINVAL, "run_fixed: " + std::to_string((long long)n) +
                                        " lines exceed its maximum of " +
                                        std::to_string((long long)kFixedMaxLines));
    if (grid->npoints <= 0 || !(grid->end > grid->start)) return GEXS_EINVAL;
    const int span = GX_WAVE * c->ppl;
    const int64_t ntiles = ((int64_t)grid->npoints + span - 1) / span;
    if (ntiles > kFixedMaxTiles)
        return c->fail(GEXS_EINVAL, "run_fixed: " + std::to_string((long long)ntiles) + " tiles of " +
                                        std::to_string(span) + " points exceed its maximum of " +
                                        std::to_string((long long)kFixedMaxTiles));
    if (!(tp->T > 0.0) || !(tp->P >= 0.0) || !(tp->Q > 0.0))
        return c->fail(GEXS_EINVAL, "run_fixed needs T > 0, P >= 0 and Q > 0");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    char* base = c->device < kMaxDevices ? (char*)g_scratch_ptr[c->device] : nullptr;
    if (!base) {
        const hipError_t e = c->device < kMaxDevices ? g_scratch_err[c->device] : hipErrorInvalidDevice;
        return c->hipfail(e == hipSuccess ? hipErrorInvalidSymbol : e, "hipGetSymbolAddress(scratch)");
    }
    TPParams* dtp = (TPParams*)(base + kFixedTabOff);
    JobDesc* djd = (JobDesc*)(base + kFixedTabOff + 256);
    Job* djob = (Job*)(base + kFixedTabOff + 512);
    unsigned long long* cnt = (unsigned long long*)(base + kFixedCntOff);
    unsigned int* xq = (unsigned int*)cnt;          // 8 queue counters (words 0-3)
    unsigned long long* scale = cnt + 4;
    unsigned long long* neval = cnt + 5;
    int* kmax = (int*)(cnt + 6);
    LinePos* pos = (LinePos*)(base + kFixedPosOff);
    VRec* rec = (VRec*)(base + kFixedRecOff);
    VRecX* recx = (VRecX*)(base + kFixedRecxOff);
    WorkItem* items = (WorkItem*)(base + kFixedItemsOff);
    const double dfreq = (grid->end - grid->start) / (double)grid->npoints;  // BaseProfile.cpp:10
    const TPParams t = make_tp(c, tp->T, tp->P, tp->Q, grid->start, dfreq, 0, grid->npoints);
    JobDesc d{};
    d.n = (int32_t)n;
    d.ntiles = (int32_t)ntiles;
    d.q_slot = -1;
    d.groups = 1;
    Job jb{};
    jb.pos = pos;
    jb.rec = rec;
    jb.recx = recx;
    jb.out = d_out;
    jb.n = (int32_t)n;
    jb.profile = c->profile;
    jb.accumulate = 0;  // every point of the grid is written once
    hipStream_t s = c->stream;
    HIPCHK(c, gxk::write_tables(s, t, d, jb, dtp, djd, djob, cnt, kFixedCntWords));
    if (n > 0) {
        HIPCHK(c, gxk::line_pos(s, L->nu, djd, 1, n, dtp, pos));
        HIPCHK(c, gxk::line_rec(s, L->nu, L->E, L->A, L->g, voigt ? L->gamma : nullptr,
                                voigt ? L->nexp : nullptr, pos, djd, 1, n, dtp, scale, rec, recx,
                                kmax, neval));
    }
    HIPCHK(c, gxk::tiles(s, pos, djd, 1, (int)ntiles, dtp, span, items, nullptr, nullptr));
    // launch shapes that need no host-built queue bounds or tile sort: one
    // slot per tile (grid or dispatch order) or persistent workgroups on one
    // counter; every shape of one tile width and split gives the same bits
    const int mode = (c->persistent_blocks > 0 && c->order != 2) ? 1 : (c->order == 0 ? 0 : 3);
    HIPCHK(c, gxk::profile_tiles(s, c->profile, c->ppl, c->split, items, (int)ntiles, djob, scale, kmax,
                                 xq, xq, mode, launch_blocks(c), true));
    if (n_eval) {
        unsigned long long ne = 0;
        HIPCHK(c, hipMemcpyAsync(&ne, neval, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        *n_eval = (int64_t)ne;
    }
    return GEXS_OK;
}

// --------------------------------------------------------------- serving
int gexs_serve_start(gexs_ctx* c, int max_inflight, int64_t mem_budget) {
    if (!c || max_inflight <= 0 || mem_budget <= 0) return GEXS_EINVAL;
    if (c->server) return c->fail(GEXS_ESTATE, "server already running");
    if (!c->have_k) return c->fail(GEXS_ESTATE, "serve_start before init_constants");
    if (c->n_states <= 0) return c->fail(GEXS_ESTATE, "serving needs a resident list (gexs_load_list)");
    if ((c->profile == GEXS_VOIGT || c->profile == GEXS_VOIGT_QUAD) && !c->has_gamma)
        return c->fail(GEXS_ESTATE, "Voigt serving needs gamma and n");
    int rc;
    if ((rc = set_device(c)) != GEXS_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    gxs::ListView v;
    v.nu = c->nu.as<double>();
    v.E = c->E.as<double>();
    v.A = c->A.as<double>();
    v.g = c->g.as<int32_t>();
    v.gamma = c->has_gamma ? c->gamma.as<double>() : nullptr;
    v.nexp = c->has_gamma ? c->nexp.as<double>() : nullptr;
    v.n = c->n;
    v.sE = c->st_E.as<double>();
    v.sg = c->st_g.as<int32_t>();
    v.n_states = c->n_states;
    v.nu_host.resize((size_t)c->n);
    if (c->n > 0) HIPCHK(c, hipMemcpy(v.nu_host.data(), c->nu.p, (size_t)c->n * 8, hipMemcpyDeviceToHost));
    gxs::Physics ph;
    ph.profile = c->profile;
    ph.mass = c->k.mean_mass;
    ph.t_ref = c->k.ref_temp;
    ph.p_scale = c->k.ref_press;
    ph.cutoff = c->k.cutoff;
    ph.half_width = c->k.half_width;
    ph.gauss_hwhm = c->k.gauss_hwhm;
    ph.width_conv = c->k.width_conv;
    ph.pressure_mode = c->k.pressure_mode;
    ph.ppl = c->ppl;
    ph.split = c->split;
    std::unique_ptr<gxs::Server> srv(new (std::nothrow) gxs::Server(c->device, std::move(v), ph, max_inflight,
                                                                    mem_budget));
    if (!srv) return GEXS_ENOMEM;
    std::string err;
    if ((rc = srv->start(&err)) != 0) return c->fail(rc, err);
    c->server = std::move(srv);
    return GEXS_OK;
}

int gexs_submit(gexs_ctx* c, const gexs_request* req, uint64_t* id) {
    if (!c || !req || !id) return GEXS_EINVAL;
    if (!c->server) return c->fail(GEXS_ESTATE, "serving not started");
    gxs::Request r{req->T, req->P, req->start, req->end, req->dfreq};
    std::string err;
    int rc = c->server->submit(r, id, &err);
    if (rc != 0) return c->fail(rc, err);
    return GEXS_OK;
}

static int fill_result(gexs_ctx* c, int rc, const gxs::ResultView& r, gexs_result* out) {
    if (rc == -6) return c->fail(GEXS_ENOTFOUND, "unknown request id");
    if (rc != 0) return c->fail(rc, "request lookup failed");
    out->status = r.status;
    out->npoints = r.npoints;
    out->sigma = r.sigma;
    out->latency_ms = r.latency_ms;
    out->device_ms = r.device_ms;
    out->tile_points = r.tile_points;
    out->waves_per_tile = r.waves_per_tile;
    out->line_groups = r.line_groups;
    out->reserved = 0;
    out->n_eval = r.n_eval;
    return GEXS_OK;
}

int gexs_poll(gexs_ctx* c, uint64_t id, gexs_result* out) {
    if (!c || !out) return GEXS_EINVAL;
    if (!c->server) return c->fail(GEXS_ESTATE, "serving not started");
    gxs::ResultView r;
    int rc = c->server->poll(id, &r);
    return fill_result(c, rc, r, out);
}

int gexs_wait(gexs_ctx* c, uint64_t id, gexs_result* out) {
    if (!c || !out) return GEXS_EINVAL;
    if (!c->server) return c->fail(GEXS_ESTATE, "serving not started");
    gxs::ResultView r;
    int rc = c->server->wait(id, &r);
    return fill_result(c, rc, r, out);
}

int gexs_release(gexs_ctx* c, uint64_t id) {
    if (!c) return GEXS_EINVAL;
    if (!c->server) return c->fail(GEXS_ESTATE, "serving not started");
    int rc = c->server->release(id);
    if (rc == -6) return c->fail(GEXS_ENOTFOUND, "unknown request id");
    if (rc != 0) return c->fail(GEXS_ESTATE, "request still running");
    return GEXS_OK;
}

int gexs_serve_stop(gexs_ctx* c) {
    if (!c) return GEXS_EINVAL;
    c->server.reset();  // drains in-flight batches and joins the scheduler
    return GEXS_OK;
}

}  // extern "C"
