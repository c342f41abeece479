// serve.h -- continuous-batching cross-section server over a resident line list.
//
// Requests (T, P, range, resolution) arrive at any time.  A scheduler thread
// admits pending requests in arrival order while their HBM footprint fits the
// server's device arena, groups the admitted ones into a batch and launches
// the batch on a free stream slot, so batches overlap on the device.  A batch
// is one launch sequence: partition function per request, line windows,
// FP32 records, tiles and one profile-kernel launch per tile width over the
// batch's tiles, then the copy-out.  Each request's tile width follows from
// its own grid size; every request owns its positions, records, tiles and
// output, and reductions never cross requests, so a result is bit-for-bit the
// same alone and in any mix.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gexs_internal.h"

namespace gxs {

// read-only view of the resident list (owned by the context)
struct ListView {
    const double* nu = nullptr;  // nu-sorted lines, device
    const double* E = nullptr;
    const double* A = nullptr;
    const int32_t* g = nullptr;
    const double* gamma = nullptr;
    const double* nexp = nullptr;
    int64_t n = 0;
    const double* sE = nullptr;  // states, device
    const int32_t* sg = nullptr;
    int64_t n_states = 0;
    std::vector<double> nu_host;  // host copy of nu for selection
};

struct Physics {
    int profile = GX_PROF_VOIGT;
    double mass = 18.0, t_ref = 296.0, p_scale = 1.0;
    double cutoff = 0.0;  // <= 0: min(25 P, 100)
    double half_width = 100.0;
    double gauss_hwhm = 0.1;
    int width_conv = 0, pressure_mode = 1;
    int ppl = 4;
    int split = 1;  // waves per tile
};

struct Request {
    double T, P, start, end, dfreq;
};

struct Result {
    int status = 1;  // 0 done, 1 pending/running, < 0 error
    int npoints = 0;
    std::vector<double> sigma;
    double latency_ms = 0.0, device_ms = 0.0;
    int tile_points = 0, waves_per_tile = 0, line_groups = 0;  // launch shape
    int64_t n_eval = 0;  // line-point evaluations of the request
};

// A request's state as one consistent record: poll and wait fill it while
// holding the server lock.  sigma points into the request's result, which
// is not written again once status is 0 and stays valid until release.
struct ResultView {
    int status = 1;
    int npoints = 0;
    const double* sigma = nullptr;
    double latency_ms = 0.0, device_ms = 0.0;
    int tile_points = 0, waves_per_tile = 0, line_groups = 0;
    int64_t n_eval = 0;
};

// first-fit allocator over one device allocation (offsets in bytes)
class Arena {
   public:
    hipError_t init(size_t bytes);
    ~Arena();
    bool alloc(size_t bytes, size_t* off);
    void free(size_t off, size_t bytes);
    char* base() const { return base_; }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }

   private:
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
};

class Server {
   public:
    Server(int device, ListView view, Physics phys, int max_inflight, int64_t mem_budget);
    ~Server();
    int start(std::string* err);
    int submit(const Request& r, uint64_t* id, std::string* err);
    int poll(uint64_t id, ResultView* out);
    int wait(uint64_t id, ResultView* out);
    int release(uint64_t id);
    // counters since start
    uint64_t batches() const { return batches_.load(); }

   private:
    struct Req {
        uint64_t id;
        Request r;
        int npoints = 0, ntiles = 0;
        int ppl = 1;    // tile width 64*ppl points,
        int split = 4;  // waves per tile and
        int groups = 1; // line groups per tile (workgroups per tile), all from the request's own size
        int64_t l0 = 0, n = 0;
        size_t foot = 0, off = 0;
        std::chrono::steady_clock::time_point arrival;
        Result res;
    };
    struct Slot {
        hipStream_t stream = nullptr;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        bool busy = false;
        std::vector<Req*> reqs;
        // device buffers of a batch (grow only; sized for the largest batch seen)
        void* tab = nullptr;  // TPParams | JobDesc | Job | T of the batch
        size_t tab_cap = 0;
        void* items = nullptr;  // work items
        size_t items_cap = 0;
        void* part = nullptr;  // partial rows of line-grouped tiles
        size_t part_cap = 0;
        void* small = nullptr;  // scale | neval | kmax | counter | Q | partition partials
        double* host_out = nullptr;  // pinned copy-out
        unsigned long long* host_neval = nullptr;  // pinned N_eval per request (kMaxBatch)
        size_t host_cap = 0;
        void* stage = nullptr;  // pinned staging of the tables
        size_t stage_cap = 0;
        Server* owner = nullptr;
    };
    void loop();
    void request_shape(Req& q) const;
    void request_groups(Req& q) const;
    int launch(Slot& s, std::vector<Req*>& batch, std::string* err);
    void finish(Slot& s);
    size_t footprint(const Req& q) const;

    int device_;
    ListView v_;
    Physics ph_;
    int max_inflight_;
    int64_t budget_;
    Arena arena_;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::mutex mu_;
    std::condition_variable cv_, cv_done_;
    std::deque<Req*> pending_;
    std::map<uint64_t, std::unique_ptr<Req>> reqs_;
    uint64_t next_id_ = 1;
    bool stop_ = false;
    std::thread thr_;
    std::atomic<uint64_t> batches_{0};
    std::string loop_err_;
};

}  // namespace gxs
