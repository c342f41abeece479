// sort.hip -- stable LSD radix sort on gfx950, plus gather/assembly helpers.
//
// One block = 256 threads x 16 rounds = 4096 keys.  Per 8-bit pass:
//   k_rs_hist     per-block digit histogram (LDS atomics on counts only)
//   k_rs_scan     per digit, exclusive scan of the block counts (fixed order)
//   k_rs_scatter  stable scatter: within a round, a lane's rank among the
//                 lanes of its wavefront holding the same digit comes from
//                 eight 64-bit ballots; waves and rounds are ordered through
//                 LDS counters, so equal keys keep their input order.
// The output order is a pure function of the input (no timing dependence).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "kernels.h"

#define RS_THREADS 256
#define RS_ROUNDS 16
#define RS_TILE (RS_THREADS * RS_ROUNDS)

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const K* __restrict__ keys, int64_t n,
                                                        int shift, uint32_t* __restrict__ hist,
                                                        int nblk) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        int64_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_rs_scan(uint32_t* __restrict__ hist, int nblk,
                                                 uint32_t* __restrict__ totals) {
    __shared__ uint32_t part[256];
    const int d = blockIdx.x;
    uint32_t* h = hist + (int64_t)d * nblk;
    const int seg = (nblk + 255) / 256;
    const int b0 = threadIdx.x * seg;
    const int b1 = min(b0 + seg, nblk);
    uint32_t s = 0;
    for (int b = b0; b < b1; ++b) s += h[b];
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 256 partial sums
    for (int o = 1; o < 256; o <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0u;
    for (int b = b0; b < b1; ++b) {
        uint32_t c = h[b];
        h[b] = run;
        run += c;
    }
    if (threadIdx.x == 255) totals[d] = part[255];
}

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const K* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           K* __restrict__ kout,
                                                           uint32_t* __restrict__ vout, int64_t n,
                                                           int shift,
                                                           const uint32_t* __restrict__ hist,
                                                           int nblk,
                                                           const uint32_t* __restrict__ totals) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[4][256];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // digit base = exclusive scan of totals + this block's prefix within the digit
    base[tid] = totals[tid];
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        uint32_t v = tid >= o ? base[tid - o] : 0u;
        __syncthreads();
        base[tid] += v;
        __syncthreads();
    }
    uint32_t excl = (tid > 0 ? base[tid - 1] : 0u) + hist[(int64_t)tid * nblk + blockIdx.x];
    __syncthreads();
    base[tid] = excl;
    run[tid] = 0;
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const int64_t blk0 = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = blk0 + r * RS_THREADS + tid;
        const bool valid = i < n;
        K key = valid ? kin[i] : (K)0;
        uint32_t val = valid ? vin[i] : 0u;
        const uint32_t d = (uint32_t)(key >> shift) & 0xFFu;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const int rank = __popcll(m & lt);
        const int cnt = __popcll(m);
#pragma unroll
        for (int w = 0; w < 4; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        if (valid && rank == cnt - 1) wcnt[wave][d] = (uint32_t)cnt;
        __syncthreads();
        if (valid) {
            uint32_t p = base[d] + run[d] + (uint32_t)rank;
            for (int w = 0; w < wave; ++w) p += wcnt[w][d];
            kout[p] = key;
            vout[p] = val;
        }
        __syncthreads();
        run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        __syncthreads();
    }
}

// out |= key ^ keys[0] over the keys: a digit that is zero in it is the same
// in every key (one atomic per wave)
template <typename K>
__global__ __launch_bounds__(256) void k_rs_diff(const K* __restrict__ keys, int64_t n,
                                                 unsigned long long* __restrict__ out) {
    const K ref = keys[0];
    unsigned long long x = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x |= (unsigned long long)(keys[i] ^ ref);
    for (int o = GX_WAVE / 2; o > 0; o >>= 1) x |= __shfl_xor(x, o);
    if ((threadIdx.x & (GX_WAVE - 1)) == 0 && x) atomicOr(out, x);
}

// With drop_zero, lines at nu == 0 get the largest key (they sort last) and
// are counted, so the caller can cut them off the end.
__global__ void k_nu_keys(const double* __restrict__ nu, int64_t n, uint64_t* __restrict__ keys,
                          uint32_t* __restrict__ vals, int drop_zero,
                          unsigned long long* __restrict__ zeros) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = nu[i];
    uint64_t u = (uint64_t)__double_as_longlong(v);
    u = (u >> 63) ? ~u : (u | 0x8000000000000000ull);
    if (drop_zero && v == 0.0) {
        u = ~0ull;
        atomicAdd(zeros, 1ull);
    }
    keys[i] = u;
    vals[i] = (uint32_t)i;
}

// [lower_bound(nu >= lo), upper_bound(nu <= hi)) of an ascending array
__global__ void k_nu_range(const double* __restrict__ nu, int64_t n, double lo, double hi,
                           long long* __restrict__ out) {
    if (threadIdx.x != 0) return;
    int64_t a = 0, b = n;
    while (a < b) {
        int64_t m = (a + b) >> 1;
        if (nu[m] < lo) a = m + 1; else b = m;
    }
    out[0] = a;
    int64_t c = a;
    b = n;
    while (c < b) {
        int64_t m = (c + b) >> 1;
        if (nu[m] <= hi) c = m + 1; else b = m;
    }
    out[1] = c;
}

template <typename T>
__global__ void k_gather(const T* __restrict__ in, const uint32_t* __restrict__ perm, int64_t n,
                         T* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[perm[i]];
}

__global__ void k_assemble(const double* __restrict__ sE, const int32_t* __restrict__ sg,
                           const uint32_t* __restrict__ up, const uint32_t* __restrict__ lo,
                           int64_t n, double* __restrict__ nu, double* __restrict__ E,
                           int32_t* __restrict__ g) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t u = up[i], l = lo[i];
        const double eu = sE[u], el = sE[l];
        nu[i] = eu - el;  // ExomolStateReader.cpp:160
        E[i] = el;        // :161
        g[i] = sg[u];     // :162
    }
}

__global__ void k_widen(const float* __restrict__ in, int64_t n, double* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

// One chunk of a streamed list: nu = E_u - E_l of each transition (as
// k_assemble), its order-preserving key and the identity permutation.  A
// transition outside the selection [lo, hi] or at nu == 0 takes the largest
// key, so it sorts behind every selected one.  out[0] += the number of such
// transitions, out[1] |= key ^ (key of transition 0) over the chunk (the
// digits the sort must pass over); one atomic each per wave.
__device__ __forceinline__ uint64_t chunk_key(const double* __restrict__ sE, uint32_t u, uint32_t l,
                                              double nu_lo, double nu_hi, bool& drop) {
    const double v = sE[u] - sE[l];
    drop = !(v >= nu_lo && v <= nu_hi) || v == 0.0;
    uint64_t k = (uint64_t)__double_as_longlong(v);
    k = (k >> 63) ? ~k : (k | 0x8000000000000000ull);
    return drop ? ~0ull : k;
}

__global__ __launch_bounds__(256) void k_chunk_keys(const double* __restrict__ sE,
                                                    const uint32_t* __restrict__ up,
                                                    const uint32_t* __restrict__ lo, int64_t n,
                                                    double nu_lo, double nu_hi,
                                                    uint64_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals,
                                                    unsigned long long* __restrict__ out) {
    bool d0;
    const uint64_t ref = chunk_key(sE, up[0], lo[0], nu_lo, nu_hi, d0);
    unsigned long long drops = 0, x = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        bool drop;
        const uint64_t k = chunk_key(sE, up[i], lo[i], nu_lo, nu_hi, drop);
        keys[i] = k;
        vals[i] = (uint32_t)i;
        drops += drop ? 1ull : 0ull;
        x |= k ^ ref;
    }
    for (int o = GX_WAVE / 2; o > 0; o >>= 1) {
        drops += __shfl_xor(drops, o);
        x |= __shfl_xor(x, o);
    }
    if ((threadIdx.x & (GX_WAVE - 1)) == 0) {
        if (drops) atomicAdd(&out[0], drops);
        if (x) atomicOr(&out[1], x);
    }
}

// The sorted lines of a streamed chunk: line i is transition perm[i], with
// nu, E'' = E_l and g' = g_u from the states (as k_assemble) and A, gamma, n
// from the chunk's FP32 (x32) or FP64 (x64) arrays, widened to FP64.
__device__ __forceinline__ double pick(const float* a32, const double* a64, uint32_t p) {
    return a64 ? a64[p] : (double)a32[p];
}
__global__ __launch_bounds__(256) void k_chunk_gather(
    const uint32_t* __restrict__ perm, int64_t n, const uint32_t* __restrict__ up,
    const uint32_t* __restrict__ lo, const double* __restrict__ sE, const int32_t* __restrict__ sg,
    const float* A32, const double* A64, const float* gam32, const double* gam64, const float* nex32,
    const double* nex64, double* __restrict__ nu, double* __restrict__ E, int32_t* __restrict__ g,
    double* __restrict__ A, double* __restrict__ gamma, double* __restrict__ nexp) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t p = perm[i];
        const uint32_t u = up[p], l = lo[p];
        const double eu = sE[u], el = sE[l];
        nu[i] = eu - el;
        E[i] = el;
        g[i] = sg[u];
        A[i] = pick(A32, A64, p);
        if (gamma) {
            gamma[i] = pick(gam32, gam64, p);
            nexp[i] = pick(nex32, nex64, p);
        }
    }
}

namespace gxk {

static int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 16384); }

size_t sort_scratch_bytes(int64_t n, int key_bytes) {
    int64_t nblk = (n + RS_TILE - 1) / RS_TILE;
    return (size_t)n * key_bytes + (size_t)n * 4 + (size_t)nblk * 256 * 4 + 1024;
}

hipError_t sort_buffers_alloc(SortBuffers& buf, int64_t n, int key_bytes) {
    if (buf.cap >= n && buf.key_bytes >= key_bytes && buf.keys_alt) return hipSuccess;
    sort_buffers_free(buf);
    int64_t nblk = (n + RS_TILE - 1) / RS_TILE;
    hipError_t e;
    if ((e = hipMalloc(&buf.keys_alt, (size_t)std::max<int64_t>(n, 1) * key_bytes)) != hipSuccess) return e;
    if ((e = hipMalloc(&buf.vals_alt, (size_t)std::max<int64_t>(n, 1) * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&buf.hist, (size_t)std::max<int64_t>(nblk, 1) * 256 * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&buf.totals, 256 * 4)) != hipSuccess) return e;
    if ((e = hipHostMalloc(&buf.totals_host, 256 * 4, hipHostMallocDefault)) != hipSuccess) return e;
    buf.cap = n;
    buf.key_bytes = key_bytes;
    return hipSuccess;
}

void sort_buffers_free(SortBuffers& buf) {
    if (buf.keys_alt) (void)hipFree(buf.keys_alt);
    if (buf.vals_alt) (void)hipFree(buf.vals_alt);
    if (buf.hist) (void)hipFree(buf.hist);
    if (buf.totals) (void)hipFree(buf.totals);
    if (buf.totals_host) (void)hipHostFree(buf.totals_host);
    buf = SortBuffers();
}

template <typename K>
static hipError_t radix_sort(hipStream_t s, K* keys, uint32_t* vals, int64_t n, SortBuffers& buf,
                             int begin_bit, int end_bit, const unsigned long long* diff_in) {
    if (n <= 1) return hipSuccess;
    hipError_t e = sort_buffers_alloc(buf, n, (int)sizeof(K));
    if (e != hipSuccess) return e;
    const int nblk = (int)((n + RS_TILE - 1) / RS_TILE);
    // the digits that vary: given, or found by one pass over the keys (the
    // stream blocks once, here, and never between the passes)
    unsigned long long diff;
    if (diff_in) {
        diff = *diff_in;
    } else {
        unsigned long long* d = (unsigned long long*)buf.totals;
        if ((e = fill_u64(s, d, 1, 0ull)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_rs_diff<K>), dim3(grid_for(n)), dim3(256), 0, s, keys, n, d);
        if ((e = hipMemcpyAsync(buf.totals_host, d, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        diff = *(const unsigned long long*)buf.totals_host;
    }
    K* kin = keys;
    uint32_t* vin = vals;
    K* kout = (K*)buf.keys_alt;
    uint32_t* vout = buf.vals_alt;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        if (((diff >> shift) & 0xFFull) == 0) continue;  // every key has the same digit: order unchanged
        hipLaunchKernelGGL((k_rs_hist<K>), dim3(nblk), dim3(RS_THREADS), 0, s, kin, n, shift,
                           buf.hist, nblk);
        hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(256), 0, s, buf.hist, nblk, buf.totals);
        hipLaunchKernelGGL((k_rs_scatter<K>), dim3(nblk), dim3(RS_THREADS), 0, s, kin, vin, kout,
                           vout, n, shift, buf.hist, nblk, buf.totals);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (kin != keys) {
        if ((e = hipMemcpyAsync(keys, kin, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(vals, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t radix_sort_u64(hipStream_t s, uint64_t* keys, uint32_t* vals, int64_t n,
                          SortBuffers& buf, int begin_bit, int end_bit,
                          const unsigned long long* diff) {
    return radix_sort<uint64_t>(s, keys, vals, n, buf, begin_bit, end_bit, diff);
}

hipError_t radix_sort_u32(hipStream_t s, uint32_t* keys, uint32_t* vals, int64_t n,
                          SortBuffers& buf, int begin_bit, int end_bit) {
    return radix_sort<uint32_t>(s, keys, vals, n, buf, begin_bit, end_bit, nullptr);
}

hipError_t make_nu_keys(hipStream_t s, const double* nu, int64_t n, uint64_t* keys,
                        uint32_t* vals, int drop_zero, unsigned long long* zeros) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_nu_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nu, n, keys,
                       vals, drop_zero, zeros);
    return hipGetLastError();
}

hipError_t nu_range(hipStream_t s, const double* nu, int64_t n, double lo, double hi,
                    long long* out) {
    hipLaunchKernelGGL(k_nu_range, dim3(1), dim3(64), 0, s, nu, n, lo, hi, out);
    return hipGetLastError();
}

hipError_t gather_f64(hipStream_t s, const double* in, const uint32_t* perm, int64_t n,
                      double* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(256), 0, s, in, perm, n, out);
    return hipGetLastError();
}

hipError_t gather_i32(hipStream_t s, const int32_t* in, const uint32_t* perm, int64_t n,
                      int32_t* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather<int32_t>, dim3(grid_for(n)), dim3(256), 0, s, in, perm, n, out);
    return hipGetLastError();
}

hipError_t assemble_lines(hipStream_t s, const double* sE, const int32_t* sg,
                          const uint32_t* up, const uint32_t* lo, int64_t n, double* nu,
                          double* E, int32_t* g) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_assemble, dim3(grid_for(n)), dim3(256), 0, s, sE, sg, up, lo, n, nu, E, g);
    return hipGetLastError();
}

hipError_t widen_f32(hipStream_t s, const float* in, int64_t n, double* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_widen, dim3(grid_for(n)), dim3(256), 0, s, in, n, out);
    return hipGetLastError();
}

hipError_t chunk_keys(hipStream_t s, const double* sE, const uint32_t* up, const uint32_t* lo,
                      int64_t n, double nu_lo, double nu_hi, uint64_t* keys, uint32_t* vals,
                      unsigned long long* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_chunk_keys, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0,
                       s, sE, up, lo, n, nu_lo, nu_hi, keys, vals, out);
    return hipGetLastError();
}

hipError_t chunk_gather(hipStream_t s, const uint32_t* perm, int64_t n, const uint32_t* up,
                        const uint32_t* lo, const double* sE, const int32_t* sg, const ChunkCols& in,
                        double* nu, double* E, int32_t* g, double* A, double* gamma, double* nexp) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_chunk_gather, dim3(grid_for(n)), dim3(256), 0, s, perm, n, up, lo, sE, sg,
                       in.A32, in.A64, in.gam32, in.gam64, in.nex32, in.nex64, nu, E, g, A, gamma,
                       nexp);
    return hipGetLastError();
}

}  // namespace gxk
