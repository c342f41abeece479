// kernels.h -- host-side launchers of the gfx950 kernels (kernels.hip, sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gexs_internal.h"

namespace gxk {

// Per-job launchers: blockIdx.y runs over the njobs jobs of jd / tp; max_n
// (max_tiles) bounds the lines (tiles) of any one job.
hipError_t line_pos(hipStream_t s, const double* nu, const JobDesc* jd, int njobs,
                    int64_t max_n, const TPParams* tp, LinePos* pos);
// FP32 records of every job's lines, with each amplitude as mantissa and
// exponent; emax[k] (zero before) receives job k's largest biased exponent,
// neval[k] its N_eval, kmax[2k] its largest k15 (Voigt profiles) and
// kmax[2k + 1] its largest kfar (Voigt) or reach (bin-Gauss), each zero
// before.
// max_blocks > 0 caps the blocks per job (a pass that runs beside another
// kernel gets its few blocks dispatched at once); 0: up to kRecMaxBlocks.
// parts (njobs * kRecRedBytes, or nullptr): each workgroup stores its sums
// there and a second kernel, one workgroup per job, adds them to the
// counters (four atomics per job); nullptr: four atomics per workgroup,
// which serialise on the job's counters (tens of ns each).
// pos_out (or nullptr): the launch also computes every line's window, as
// line_pos does, stores it there (at the jobs' pos_off) and uses it instead
// of reading pos: one launch and one dependency fewer.
// reduce = false (with parts): the second kernel is left out, and the
// caller passes the sums to the tiles launch (RecReduce) instead.
hipError_t line_rec(hipStream_t s, const double* nu, const double* E, const double* A,
                    const int* g, const double* gamma, const double* nexp, const LinePos* pos,
                    const JobDesc* jd, int njobs, int64_t max_n, const TPParams* tp,
                    unsigned long long* emax, VRec* rec, VRecX* recx, int* kmax,
                    unsigned long long* neval, int max_blocks = 0, void* parts = nullptr,
                    LinePos* pos_out = nullptr, bool reduce = true);
// the workgroups per job of a line_rec launch over max_n lines
int rec_blocks(int64_t max_n, int max_blocks = 0);
// Scratch of one job of a line-record launch: the sums of each of its
// workgroups, which a kernel of one workgroup per job adds to the counters.
// Launches that may run at the same time need scratch of their own.
constexpr int kRecMaxBlocks = 2048;
constexpr size_t kRecPartBytes = 32;
constexpr size_t kRecRedBytes = (size_t)kRecMaxBlocks * kRecPartBytes;
// line_rec of one job whose tables are kernel arguments (tp, jd, and jb,
// stored to djob for the profile kernel), its counters (emax, kmax, neval)
// zero before, added to with four atomics per pass (through red); the
// launch also zeroes clear[0 .. nclear), a later pass's counters (at least
// one block runs, also for jd.n = 0; at most kRecMaxBlocks).  Two kernels.
hipError_t line_rec_pass(hipStream_t s, const TPParams& tp, const JobDesc& jd, const Job& jb,
                         Job* djob, unsigned long long* clear, int nclear, const double* nu,
                         const double* E, const double* A, const int* g, const double* gamma,
                         const double* nexp, const LinePos* pos, unsigned long long* emax,
                         VRec* rec, VRecX* recx, int* kmax, unsigned long long* neval,
                         void* red, int max_blocks = 0);
// the segment bounds the profile kernel locates for every work item of tile
// width span from kmax (f2, s2, s3, f3, e2, e3; the bin-Gauss profiles f2,
// f3 only, the lines within
// reach of the tile), stored into the items: for inspection, the profile
// kernel does not read them.  pos: the lines of every job (pos_off 0).
hipError_t segments(hipStream_t s, WorkItem* items, int nitems, const LinePos* pos,
                    const int* kmax, int profile, int span);
// jobs job0 .. job0 + njobs - 1 of jd / tp, all with tiles of span points
// red (or nullptr): a line_rec launch's workgroup sums not yet reduced
// (line_rec with reduce = false); one more workgroup per job adds them to
// jobs job0 .. job0 + njobs - 1's counters, in place of line_rec's second kernel
struct RecReduce {
    const void* parts;  // line_rec's parts
    int nb;             // its workgroups per job (rec_blocks)
    unsigned long long* emax;
    int* kmax;
    unsigned long long* neval;
};
hipError_t tiles(hipStream_t s, const LinePos* pos, const JobDesc* jd, int njobs, int max_tiles,
                 const TPParams* tp, int span, WorkItem* items, uint32_t* cost_key,
                 uint32_t* item_idx, int job0 = 0, const RecReduce* red = nullptr);
// items of tiles cut into line groups (WorkItem::ngroups > 1): every point of
// such a tile = the sum of its groups' rows of Job::part (span = 64 * PPL)
hipError_t group_sum(hipStream_t s, const WorkItem* items, int nitems, const Job* jobs, int span);
hipError_t permute_items(hipStream_t s, const WorkItem* in, const uint32_t* idx, int n,
                         WorkItem* out);
// split 1: one wave per item; split 4: the four waves of a workgroup share an
// item's lines.  mode 0 / 3: one launch slot per item (XCD-contiguous /
// dispatch order); 1: `blocks` persistent workgroups on one counter (xq[0]);
// 2: `blocks` persistent workgroups on per-XCD queues (xq = 8 counters, zero
// at launch; xqb = 9 item bounds)
// fills the bin-averaged Gaussian's erfc table on the current device (once
// per device before its first Gaussian profile kernel; idempotent)
hipError_t init_gauss_tables(hipStream_t s);
// the profile kernel of one profile (tiles_voigt.hip, tiles_quad.hip,
// tiles_gauss.hip); single != 0: all items belong to jobs[0]
hipError_t profile_tiles_voigt(hipStream_t s, int ppl, int split, const WorkItem* items, int nitems,
                               const Job* jobs, int single, const unsigned long long* scale_bits,
                               const int* kmax, unsigned int* xq, const unsigned int* xqb, int mode, int blocks);
hipError_t profile_tiles_quad(hipStream_t s, int ppl, int split, const WorkItem* items, int nitems,
                              const Job* jobs, int single, const unsigned long long* scale_bits,
                              const int* kmax, unsigned int* xq, const unsigned int* xqb, int mode, int blocks);
hipError_t profile_tiles_gauss(hipStream_t s, int ppl, int split, const WorkItem* items, int nitems,
                               const Job* jobs, int single, const unsigned long long* scale_bits,
                               const int* kmax, unsigned int* xq, const unsigned int* xqb, int mode, int blocks);
// single_job: all items belong to jobs[0] (read without waiting for the item).
// kmax: the jobs' largest k15 and kfar (or reach) from line_rec, from which
// each wave locates its item's segments.
hipError_t profile_tiles(hipStream_t s, int profile, int ppl, int split, const WorkItem* items,
                         int nitems, const Job* jobs, const unsigned long long* scale_bits,
                         const int* kmax, unsigned int* xq, const unsigned int* xqb, int mode,
                         int blocks, bool single_job = false);
hipError_t line_f64(hipStream_t s, const double* nu, const double* E, const double* A,
                    const int* g, const double* gamma, const double* nexp, int n,
                    const LinePos* pos, const TPParams* tp, LineF64* out,
                    unsigned long long* neval);
hipError_t fp64_points(hipStream_t s, int profile, const double* nu, const LinePos* pos,
                       const LineF64* lf, int n, double start, double dfreq, int win_lo,
                       int win_hi, double* out);
// Q(T) of nT temperatures (dQ); with jd, also tp[j].Q for the jobs whose
// q_slot names a temperature, and clear[0 .. nclear) zeroed, in the same launch
hipError_t partition(hipStream_t s, const double* E, const int* g, int64_t n, const double* dT,
                     int nT, double* partial, int nblk, double* dQ, const JobDesc* jd = nullptr,
                     int njobs = 0, TPParams* tp = nullptr, unsigned long long* clear = nullptr,
                     int nclear = 0);
// one job's tables from the host values (kernel arguments) into device
// memory, and clear[0 .. nclear) zeroed, in one launch
hipError_t write_tables(hipStream_t s, const TPParams& tp, const JobDesc& jd, const Job& jb,
                        TPParams* dtp, JobDesc* djd, Job* djob, unsigned long long* clear,
                        int nclear);
hipError_t fill_u64(hipStream_t s, unsigned long long* p, int64_t n, unsigned long long v);
hipError_t fill_f64(hipStream_t s, double* p, int64_t n, double v);
hipError_t add_f64(hipStream_t s, double* dst, const double* src, int64_t n);

// ---- sort.hip: stable LSD radix sort of (key, uint32 value) pairs.
// Keys are uint64 (order-preserving image of a double) or uint32.
struct SortBuffers {
    void* keys_alt = nullptr;      // n keys
    uint32_t* vals_alt = nullptr;  // n values
    uint32_t* hist = nullptr;      // 256 * nblk
    uint32_t* totals = nullptr;    // 256 (device)
    uint32_t* totals_host = nullptr;  // 256 (pinned host)
    int64_t cap = 0;
    int key_bytes = 0;
};
size_t sort_scratch_bytes(int64_t n, int key_bytes);
// Sorts keys/vals in place (results end in keys/vals), skipping the passes
// whose digit is the same in every key.  Those digits come from diff (the OR
// of key ^ keys[0] over the keys) when given; otherwise one pass over the
// keys finds them and the call blocks on the stream once, before the passes.
hipError_t radix_sort_u64(hipStream_t s, uint64_t* keys, uint32_t* vals, int64_t n,
                          SortBuffers& buf, int begin_bit, int end_bit,
                          const unsigned long long* diff = nullptr);
hipError_t radix_sort_u32(hipStream_t s, uint32_t* keys, uint32_t* vals, int64_t n,
                          SortBuffers& buf, int begin_bit, int end_bit);
hipError_t sort_buffers_alloc(SortBuffers& buf, int64_t n, int key_bytes);
void sort_buffers_free(SortBuffers& buf);

// nu (double) -> order-preserving uint64 key, and the identity permutation;
// drop_zero sends nu == 0 lines to the end and counts them into *zeros
hipError_t make_nu_keys(hipStream_t s, const double* nu, int64_t n, uint64_t* keys,
                        uint32_t* vals, int drop_zero, unsigned long long* zeros);
// out[0..1] = [first nu >= lo, first nu > hi) of an ascending array
hipError_t nu_range(hipStream_t s, const double* nu, int64_t n, double lo, double hi,
                    long long* out);
// out[i] = in[perm[i]] for 8- and 4-byte elements
hipError_t gather_f64(hipStream_t s, const double* in, const uint32_t* perm, int64_t n,
                      double* out);
hipError_t gather_i32(hipStream_t s, const int32_t* in, const uint32_t* perm, int64_t n,
                      int32_t* out);
// assemble lines of a resident list from states + transitions
hipError_t assemble_lines(hipStream_t s, const double* sE, const int32_t* sg,
                          const uint32_t* up, const uint32_t* lo, int64_t n, double* nu,
                          double* E, int32_t* g);
hipError_t widen_f32(hipStream_t s, const float* in, int64_t n, double* out);

// ---- one chunk of a streamed list (gexs_stream_xsec)
// keys of the chunk's transitions by nu = E_u - E_l; those outside
// [nu_lo, nu_hi] or at nu == 0 sort last.  out[0] (zero before) counts them,
// out[1] (zero before) receives the OR of key ^ key[0] (radix_sort_u64's diff)
hipError_t chunk_keys(hipStream_t s, const double* sE, const uint32_t* up, const uint32_t* lo,
                      int64_t n, double nu_lo, double nu_hi, uint64_t* keys, uint32_t* vals,
                      unsigned long long* out);
// the chunk's A, gamma, n columns: per column one of the FP32 / FP64 arrays
// (gamma and n both NULL for a Doppler-only list)
struct ChunkCols {
    const float* A32;
    const double* A64;
    const float* gam32;
    const double* gam64;
    const float* nex32;
    const double* nex64;
};
// lines [0, n) of the sorted chunk: line i = transition perm[i] (gamma and
// nexp NULL: not written)
hipError_t chunk_gather(hipStream_t s, const uint32_t* perm, int64_t n, const uint32_t* up,
                        const uint32_t* lo, const double* sE, const int32_t* sg, const ChunkCols& in,
                        double* nu, double* E, int32_t* g, double* A, double* gamma, double* nexp);

}  // namespace gxk
