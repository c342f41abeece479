// serve.hip -- continuous-batching request server (see serve.h).
#include "serve.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "kernels.h"

namespace gxs {

namespace {
constexpr size_t kAlign = 256;
constexpr int kMaxBatch = 64;
constexpr int kQBlocks = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

double cutoff_of(const Physics& ph, double P) {
    return ph.cutoff > 0.0 ? ph.cutoff : std::min(25.0 * P, 100.0);  // OpenMPManager.cpp:386
}

// Slot buffers grow by at least half their size (to a whole MiB), so that a
// run of slightly larger batches reallocates a few times only: hipFree waits
// for the whole device and pinning host pages costs milliseconds, both on
// the path of the batch being launched.
size_t grown(size_t cap, size_t want) {
    if (cap >= want) return cap;
    const size_t mib = size_t(1) << 20;
    return (std::max(want, cap + cap / 2) + mib - 1) / mib * mib;
}

hipError_t grow(void** p, size_t* cap, size_t want) {
    if (*p && *cap >= want) return hipSuccess;
    want = grown(*cap, std::max<size_t>(want, kAlign));
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    hipError_t e = hipMalloc(p, want);
    *cap = e == hipSuccess ? want : 0;
    return e;
}

hipError_t grow_pinned(void** p, size_t* cap, size_t want) {
    if (*p && *cap >= want) return hipSuccess;
    want = grown(*cap, std::max<size_t>(want, kAlign));
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    hipError_t e = hipHostMalloc(p, want, hipHostMallocDefault);
    *cap = e == hipSuccess ? want : 0;
    return e;
}

// Capacity each slot reserves at start: tables of kMaxBatch requests,
// 2^16 work items, 8 MiB of line-group partials and 2^21 output points
// (16 MiB pinned), which hold the batches of typical requests without
// growing.
constexpr size_t kReserveItems = size_t(1) << 16;
constexpr size_t kReservePart = size_t(8) << 20;
constexpr size_t kReservePoints = size_t(1) << 21;
}  // namespace

// ------------------------------------------------------------------ Arena
hipError_t Arena::init(size_t bytes) {
    cap_ = bytes / kAlign * kAlign;
    hipError_t e = hipMalloc((void**)&base_, std::max<size_t>(cap_, kAlign));
    if (e != hipSuccess) {
        base_ = nullptr;
        cap_ = 0;
        return e;
    }
    free_.clear();
    free_[0] = cap_;
    used_ = 0;
    return hipSuccess;
}

Arena::~Arena() {
    if (base_) (void)hipFree(base_);
}

bool Arena::alloc(size_t bytes, size_t* off) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= bytes) {
            *off = it->first;
            size_t rest = it->second - bytes;
            size_t noff = it->first + bytes;
            free_.erase(it);
            if (rest) free_[noff] = rest;
            used_ += bytes;
            return true;
        }
    }
    return false;
}

void Arena::free(size_t off, size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    // coalesce with the next and the previous free block
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

// ------------------------------------------------------------------ Server
Server::Server(int device, ListView view, Physics phys, int max_inflight, int64_t mem_budget)
    : device_(device), v_(std::move(view)), ph_(phys), max_inflight_(std::max(1, max_inflight)),
      budget_(mem_budget) {}

int Server::start(std::string* err) {
    if (hipSetDevice(device_) != hipSuccess) {
        *err = "hipSetDevice failed";
        return -3;
    }
    if (arena_.init((size_t)budget_) != hipSuccess) {
        *err = "server arena allocation failed";
        return -2;
    }
    for (int i = 0; i < max_inflight_; ++i) {
        // owned by slots_ before anything is created in it, so that the
        // destructor releases what a failed start leaves
        slots_.push_back(std::unique_ptr<Slot>(new Slot()));
        Slot* s = slots_.back().get();
        s->owner = this;
        if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
            *err = "stream/event creation failed";
            return -3;
        }
        const size_t tab_b = kMaxBatch * (sizeof(TPParams) + sizeof(JobDesc) + sizeof(Job) + 8) + 4 * kAlign;
        size_t host_b = 0;
        if (grow(&s->tab, &s->tab_cap, tab_b) != hipSuccess ||
            grow(&s->items, &s->items_cap, kReserveItems * sizeof(WorkItem)) != hipSuccess ||
            grow(&s->part, &s->part_cap, kReservePart) != hipSuccess ||
            grow_pinned(&s->stage, &s->stage_cap, tab_b) != hipSuccess ||
            grow_pinned((void**)&s->host_out, &host_b, kReservePoints * 8) != hipSuccess ||
            hipHostMalloc((void**)&s->host_neval, kMaxBatch * 8, hipHostMallocDefault) != hipSuccess) {
            s->host_cap = host_b / 8;
            *err = "slot buffer allocation failed";
            return -2;
        }
        s->host_cap = host_b / 8;
    }
    thr_ = std::thread(&Server::loop, this);
    return 0;
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (thr_.joinable()) thr_.join();
    (void)hipSetDevice(device_);
    for (auto& s : slots_) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        void* bufs[] = {s->tab, s->items, s->small, s->part};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (s->host_out) (void)hipHostFree(s->host_out);
        if (s->host_neval) (void)hipHostFree(s->host_neval);
        if (s->stage) (void)hipHostFree(s->stage);
        if (s->ev0) (void)hipEventDestroy(s->ev0);
        if (s->ev1) (void)hipEventDestroy(s->ev1);
        if (s->stream) (void)hipStreamDestroy(s->stream);
    }
}

size_t Server::footprint(const Req& q) const {
    return align_up((size_t)q.n * sizeof(LinePos)) + align_up((size_t)q.n * sizeof(VRec)) +
           align_up((size_t)q.n * sizeof(VRecX)) + align_up((size_t)q.npoints * 8);
}

int Server::submit(const Request& r, uint64_t* id, std::string* err) {
    if (!(r.T > 0.0) || !(r.end > r.start) || !(r.dfreq > 0.0) || !(r.P >= 0.0)) {
        *err = "request needs T > 0, P >= 0, end > start, dfreq > 0";
        return -1;
    }
    const double npf = std::round((r.end - r.start) / r.dfreq);
    if (!(npf >= 1.0) || npf > 2.0e9) {
        *err = "request grid has no points or too many";
        return -1;
    }
    std::unique_ptr<Req> q(new Req());
    q->r = r;
    q->npoints = (int)npf;
    request_shape(*q);
    // line selection, Profiles/BaseProfile.cpp:152-157
    const double lo = r.start - ph_.half_width, hi = r.end + ph_.half_width;
    auto a = std::lower_bound(v_.nu_host.begin(), v_.nu_host.end(), lo);
    auto b = std::upper_bound(a, v_.nu_host.end(), hi);
    q->l0 = a - v_.nu_host.begin();
    q->n = b - a;
    request_groups(*q);
    q->res.tile_points = GX_WAVE * q->ppl;
    q->res.waves_per_tile = q->split;
    q->res.line_groups = q->groups;
    q->foot = footprint(*q);
    if ((int64_t)q->foot > budget_) {
        *err = "request footprint exceeds the server memory budget";
        return -2;
    }
    q->arrival = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> g(mu_);
    if (stop_) {
        *err = "server stopping";
        return -4;
    }
    q->id = next_id_++;
    *id = q->id;
    pending_.push_back(q.get());
    reqs_[q->id] = std::move(q);
    cv_.notify_all();
    return 0;
}

namespace {
// the caller holds the server lock
void view_of(const Result& r, ResultView* v) {
    const int st = r.status;
    v->status = st;
    v->npoints = st == 0 ? r.npoints : 0;
    v->sigma = st == 0 ? r.sigma.data() : nullptr;
    v->latency_ms = r.latency_ms;
    v->device_ms = r.device_ms;
    v->tile_points = r.tile_points;
    v->waves_per_tile = r.waves_per_tile;
    v->line_groups = r.line_groups;
    v->n_eval = r.n_eval;
}
}  // namespace

int Server::poll(uint64_t id, ResultView* out) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return -6;
    view_of(it->second->res, out);
    return 0;
}

int Server::wait(uint64_t id, ResultView* out) {
    std::unique_lock<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return -6;
    Req* q = it->second.get();
    cv_done_.wait(g, [&] { return q->res.status != 1; });
    view_of(q->res, out);
    return 0;
}

int Server::release(uint64_t id) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return -6;
    if (it->second->res.status == 1) return -4;  // still running
    reqs_.erase(it);
    return 0;
}

void Server::finish(Slot& s) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, s.ev0, s.ev1);
    const auto now = std::chrono::steady_clock::now();
    size_t hoff = 0;
    int k = 0;
    for (Req* q : s.reqs) {
        q->res.n_eval = (int64_t)s.host_neval[k++];
        q->res.sigma.assign(s.host_out + hoff, s.host_out + hoff + q->npoints);
        hoff += (size_t)q->npoints;
        q->res.npoints = q->npoints;
        q->res.device_ms = ms;
        q->res.latency_ms = std::chrono::duration<double, std::milli>(now - q->arrival).count();
        q->res.status = 0;
        arena_.free(q->off, q->foot);
    }
    s.reqs.clear();
    s.busy = false;
}

// Launch shape of a request: the widest tile up to the configured one that
// still gives every SIMD of the device a wave (8 K waves), with the
// configured waves per tile; smaller requests spread their lines over more
// workgroups by line groups (request_groups), which measured faster than 16
// waves per tile (one workgroup holding a tile's near-centre lines waits on
// its slowest wave).  The shape depends on the request alone, so a result
// is the same in any batch.
void Server::request_shape(Req& q) const {
    const int64_t want_waves = 8 * 1024;
    const int split = std::max(ph_.split, 1);
    q.ppl = 1;
    for (int ppl = ph_.ppl; ppl > 1; ppl /= 2)
        if ((int64_t)q.npoints / (GX_WAVE * ppl) * split >= want_waves) {
            q.ppl = ppl;
            break;
        }
    q.ntiles = (q.npoints + GX_WAVE * q.ppl - 1) / (GX_WAVE * q.ppl);
    q.split = split;
}

// Line groups: a request whose tiles give the device fewer than kFillWaves
// waves (16 per SIMD) also cuts every tile's lines into groups, one
// workgroup each (their partial sums added in group order by k_group_sum),
// as long as a wave keeps kMinLinesPerWave lines.  A few hundred
// workgroups land one or two per CU, unevenly; more, smaller workgroups
// even that out (each group takes an equal share of a tile's near-centre
// lines too).  Lines per tile are estimated from the request's selected
// lines, its width and the cutoff window.
void Server::request_groups(Req& q) const {
    constexpr int64_t kFillWaves = 16384;
    constexpr int kMaxGroups = 256;
    constexpr double kMinLinesPerWave = 128.0;
    q.groups = 1;
    const int64_t waves = (int64_t)q.ntiles * q.split;
    if (waves >= kFillWaves || q.n <= 0) return;
    const double dnu = (q.r.end - q.r.start) / (double)q.npoints;
    const double span_cm = GX_WAVE * q.ppl * dnu;
    const double cut = cutoff_of(ph_, q.r.P);
    const double lines_per_tile = (double)q.n * std::min(1.0, (span_cm + 2.0 * cut) /
                                                                  (q.r.end - q.r.start + 2.0 * ph_.half_width));
    int g = (int)((kFillWaves + waves - 1) / waves);
    g = std::min(g, kMaxGroups);
    g = std::min(g, (int)(lines_per_tile / (kMinLinesPerWave * q.split)));
    q.groups = std::max(g, 1);
}

int Server::launch(Slot& s, std::vector<Req*>& batch, std::string* err) {
    const int nreq = (int)batch.size();
    hipStream_t st = s.stream;
    char* base = arena_.base();
    // requests grouped by launch shape (one tiles / profile launch per group)
    auto shape = [](const Req* q) { return q->ppl * 64 + q->split; };
    std::stable_sort(batch.begin(), batch.end(),
                     [&](const Req* a, const Req* b) { return shape(a) > shape(b); });
    // ---- host tables; each request's arena region holds its positions,
    // records and output grid
    std::vector<TPParams> tps((size_t)nreq);
    std::vector<JobDesc> jds((size_t)nreq);
    std::vector<Job> jobs((size_t)nreq);
    std::vector<double> Ts((size_t)nreq);
    int items_total = 0, max_tiles = 0;
    size_t part_total = 0;  // doubles
    int64_t max_n = 0;
    size_t host_total = 0;
    for (int k = 0; k < nreq; ++k) {
        Req* q = batch[(size_t)k];
        const size_t pos_b = q->off;
        const size_t rec_b = pos_b + align_up((size_t)q->n * sizeof(LinePos));
        const size_t recx_b = rec_b + align_up((size_t)q->n * sizeof(VRec));
        const size_t out_b = recx_b + align_up((size_t)q->n * sizeof(VRecX));
        TPParams& p = tps[(size_t)k];
        p = TPParams{};
        p.T = q->r.T;
        p.P = q->r.P;
        p.Q = 1.0;  // set on the device from the partition pass
        p.mass = ph_.mass;
        p.t_ref = ph_.t_ref;
        p.p_scale = ph_.p_scale;
        p.cutoff = cutoff_of(ph_, q->r.P);
        p.start = q->r.start;
        p.dfreq = (q->r.end - q->r.start) / (double)q->npoints;  // BaseProfile.cpp:10
        p.gauss_hwhm = ph_.gauss_hwhm;
        p.profile = ph_.profile;
        p.width_conv = ph_.width_conv;
        p.pressure_mode = ph_.pressure_mode;
        p.win_lo = 0;
        p.win_hi = q->npoints;
        JobDesc& d = jds[(size_t)k];
        d.l0 = q->l0;
        d.pos_off = (int64_t)(pos_b / sizeof(LinePos));
        d.rec_off = (int64_t)(rec_b / sizeof(VRec));
        d.recx_off = (int64_t)(recx_b / sizeof(VRecX));
        d.n = (int32_t)q->n;
        d.items_off = items_total;
        d.ntiles = q->ntiles;
        d.q_slot = k;
        d.groups = q->groups;
        Job& jb = jobs[(size_t)k];
        jb.pos = (const LinePos*)(base + pos_b);
        jb.rec = (const VRec*)(base + rec_b);
        jb.recx = (const VRecX*)(base + recx_b);
        jb.out = (double*)(base + out_b);
        jb.n = (int32_t)q->n;
        jb.profile = ph_.profile;
        jb.scale_slot = k;
        jb.accumulate = 0;  // each point is written exactly once
        jb.part = nullptr;  // set below, once the partial buffer is sized
        Ts[(size_t)k] = q->r.T;
        items_total += q->ntiles * q->groups;
        if (q->groups > 1) part_total += (size_t)q->ntiles * q->groups * GX_WAVE * q->ppl;
        max_tiles = std::max(max_tiles, q->ntiles);
        max_n = std::max<int64_t>(max_n, q->n);
        host_total += (size_t)q->npoints;
    }
    const size_t tp_b = align_up(tps.size() * sizeof(TPParams));
    const size_t jd_b = align_up(jds.size() * sizeof(JobDesc));
    const size_t jobs_b = align_up(jobs.size() * sizeof(Job));
    const size_t t_b = align_up(Ts.size() * 8);
    const size_t tab_b = tp_b + jd_b + jobs_b + t_b;
    // small arrays: scale, neval, kmax (kMaxBatch u64 each; cleared by one
    // fill), counter, Q, partition partials, line-record workgroup sums
    const size_t sm_scale = 0, sm_neval = kMaxBatch * 8, sm_kmax = 2 * kMaxBatch * 8,
                 sm_counter = 3 * kMaxBatch * 8, sm_q = sm_counter + 256,
                 sm_part = sm_q + kMaxBatch * 8,
                 sm_rp = align_up(sm_part + (size_t)kMaxBatch * kQBlocks * 8),
                 sm_b = sm_rp + (size_t)kMaxBatch * gxk::kRecRedBytes;
    hipError_t e = hipSuccess;
    size_t dummy_cap = 0;
    if (!s.small && (e = grow(&s.small, &dummy_cap, sm_b)) != hipSuccess) goto fail;
    if ((e = grow(&s.tab, &s.tab_cap, tab_b)) != hipSuccess) goto fail;
    if ((e = grow(&s.items, &s.items_cap, (size_t)std::max(items_total, 1) * sizeof(WorkItem))) !=
        hipSuccess)
        goto fail;
    if (part_total > 0 && (e = grow(&s.part, &s.part_cap, part_total * 8)) != hipSuccess) goto fail;
    {
        size_t po = 0;
        for (int k = 0; k < nreq; ++k) {
            const Req* q = batch[(size_t)k];
            if (q->groups <= 1) continue;
            jobs[(size_t)k].part = (double*)s.part + po;
            po += (size_t)q->ntiles * q->groups * GX_WAVE * q->ppl;
        }
    }
    if ((e = grow_pinned(&s.stage, &s.stage_cap, tab_b)) != hipSuccess) goto fail;
    if (s.host_cap < host_total) {
        size_t host_b = s.host_cap * 8;
        e = grow_pinned((void**)&s.host_out, &host_b, host_total * 8);
        s.host_cap = host_b / 8;
        if (e != hipSuccess) goto fail;
    }
    {
        // the slot's previous batch has completed (the slot was free), so its
        // staging buffer may be rewritten
        char* h = (char*)s.stage;
        std::memcpy(h, tps.data(), tps.size() * sizeof(TPParams));
        std::memcpy(h + tp_b, jds.data(), jds.size() * sizeof(JobDesc));
        std::memcpy(h + tp_b + jd_b, jobs.data(), jobs.size() * sizeof(Job));
        std::memcpy(h + tp_b + jd_b + jobs_b, Ts.data(), Ts.size() * 8);
        char* tab = (char*)s.tab;
        TPParams* d_tp = (TPParams*)tab;
        JobDesc* d_jd = (JobDesc*)(tab + tp_b);
        Job* d_jobs = (Job*)(tab + tp_b + jd_b);
        double* d_T = (double*)(tab + tp_b + jd_b + jobs_b);
        char* sm = (char*)s.small;
        unsigned long long* d_scale = (unsigned long long*)(sm + sm_scale);
        unsigned long long* d_neval = (unsigned long long*)(sm + sm_neval);
        unsigned int* d_counter = (unsigned int*)(sm + sm_counter);
        double* d_Q = (double*)(sm + sm_q);
        double* d_part = (double*)(sm + sm_part);
        int* d_kmax = (int*)(sm + sm_kmax);
        WorkItem* d_items = (WorkItem*)s.items;
        if ((e = hipEventRecord(s.ev0, st)) != hipSuccess) goto fail;
        if ((e = hipMemcpyAsync(tab, s.stage, tab_b, hipMemcpyHostToDevice, st)) != hipSuccess) goto fail;
        // Q(T) into every request's TPParams, and scale | neval | kmax zeroed
        if ((e = gxk::partition(st, v_.sE, v_.sg, v_.n_states, d_T, nreq, d_part, kQBlocks, d_Q, d_jd, nreq,
                                d_tp, d_scale, 3 * kMaxBatch)) != hipSuccess)
            goto fail;
        // windows and records of every request's lines in one launch; its
        // workgroup sums are added to the counters by the tile launches below
        if ((e = gxk::line_rec(st, v_.nu, v_.E, v_.A, v_.g, v_.gamma, v_.nexp, nullptr, d_jd,
                               nreq, max_n, d_tp, d_scale, (VRec*)base, (VRecX*)base, d_kmax, d_neval, 0,
                               sm + sm_rp, (LinePos*)base, false)) != hipSuccess)
            goto fail;
        const gxk::RecReduce red{sm + sm_rp, gxk::rec_blocks(max_n), d_scale, d_kmax, d_neval};
        const gxk::RecReduce* redp = max_n > 0 ? &red : nullptr;  // no launch, no sums
        for (int k0 = 0; k0 < nreq;) {  // tiles per tile-width group
            int k1 = k0, mt = 0;
            while (k1 < nreq && shape(batch[(size_t)k1]) == shape(batch[(size_t)k0]))
                mt = std::max(mt, batch[(size_t)k1++]->ntiles);
            if ((e = gxk::tiles(st, (const LinePos*)base, d_jd, k1 - k0, mt, d_tp,
                                GX_WAVE * batch[(size_t)k0]->ppl, d_items, nullptr, nullptr, k0, redp)) !=
                hipSuccess)
                goto fail;
            k0 = k1;
        }
        for (int k0 = 0; k0 < nreq;) {  // one profile launch per group, slots in dispatch order
            int k1 = k0, i1 = jds[(size_t)k0].items_off;
            bool grouped = false;
            while (k1 < nreq && shape(batch[(size_t)k1]) == shape(batch[(size_t)k0])) {
                grouped |= batch[(size_t)k1]->groups > 1;
                i1 += batch[(size_t)k1]->ntiles * batch[(size_t)k1]->groups;
                ++k1;
            }
            const int i0 = jds[(size_t)k0].items_off;
            if ((e = gxk::profile_tiles(st, ph_.profile, batch[(size_t)k0]->ppl, batch[(size_t)k0]->split,
                                        d_items + i0,
                                        i1 - i0, d_jobs, d_scale, d_kmax, d_counter, nullptr, 3, 0)) !=
                hipSuccess)
                goto fail;
            if (grouped &&
                (e = gxk::group_sum(st, d_items + i0, i1 - i0, d_jobs, GX_WAVE * batch[(size_t)k0]->ppl)) !=
                    hipSuccess)
                goto fail;
            k0 = k1;
        }
        size_t hoff = 0;
        for (int k = 0; k < nreq; ++k) {
            const size_t np = (size_t)batch[(size_t)k]->npoints;
            if ((e = hipMemcpyAsync(s.host_out + hoff, jobs[(size_t)k].out, np * 8, hipMemcpyDeviceToHost,
                                    st)) != hipSuccess)
                goto fail;
            hoff += np;
        }
        if ((e = hipMemcpyAsync(s.host_neval, d_neval, (size_t)nreq * 8, hipMemcpyDeviceToHost, st)) !=
            hipSuccess)
            goto fail;
        if ((e = hipEventRecord(s.ev1, st)) != hipSuccess) goto fail;
    }
    s.reqs = batch;
    s.busy = true;
    batches_++;
    return 0;
fail:
    *err = std::string("batch launch: ") + hipGetErrorString(e);
    return -3;
}

void Server::loop() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> g(mu_);
    for (;;) {
        bool inflight = false;
        for (auto& s : slots_) inflight |= s->busy;
        if (stop_ && !inflight && pending_.empty()) break;
        // completions
        for (auto& s : slots_) {
            if (s->busy && hipEventQuery(s->ev1) == hipSuccess) {
                finish(*s);
                cv_done_.notify_all();
            }
        }
        // admission: FIFO while the arena has room and a slot is free
        for (auto& s : slots_) {
            if (s->busy || pending_.empty()) continue;
            std::vector<Req*> batch;
            while (!pending_.empty() && (int)batch.size() < kMaxBatch) {
                Req* q = pending_.front();
                if (!arena_.alloc(q->foot, &q->off)) break;
                pending_.pop_front();
                batch.push_back(q);
            }
            if (batch.empty()) break;
            std::string err;
            // the launch (host tables, copies, kernel launches) runs without
            // the lock so that submit and poll do not wait for it; slots and
            // the arena's regions of this batch belong to this thread
            g.unlock();
            const int lrc = launch(*s, batch, &err);
            // copies and kernels enqueued before a failure may still be
            // writing the batch's arena regions and slot buffers: they are
            // given back only once the slot's stream has drained
            if (lrc != 0) (void)hipStreamSynchronize(s->stream);
            g.lock();
            if (lrc != 0) {
                for (Req* q : batch) {
                    q->res.status = -3;
                    arena_.free(q->off, q->foot);
                }
                loop_err_ = err;
                cv_done_.notify_all();
            }
        }
        inflight = false;
        for (auto& s : slots_) inflight |= s->busy;
        if (inflight)
            cv_.wait_for(g, std::chrono::microseconds(50));
        else if (!stop_ && pending_.empty())
            cv_.wait(g);
        else if (!pending_.empty())
            cv_.wait_for(g, std::chrono::microseconds(50));
    }
}

}  // namespace gxs
