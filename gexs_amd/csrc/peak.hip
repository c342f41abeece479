// peak.hip -- on-device VALU rate probes (libgexs_peak.so), run by bench.py.
//
// Three kernels of independent FP32 dependency chains, sized to fill every
// SIMD: FMA chains (the FP32 vector rate this build can reach: hipcc is told
// not to form packed FP32, so one FMA per lane per instruction), v_rcp_f32
// chains (the issue cost of the reciprocal), and the region-I evaluation
// sequence of the profile kernel (x, s, den, num, rcp, accumulate: six VALU
// instructions per point) with its operands in SGPRs, as in the kernel's
// staged line loop.
#include <hip/hip_runtime.h>

namespace {

constexpr int kChains = 16;

__global__ __launch_bounds__(256) void k_fma_chains(float* out, float b, float c, int iters) {
    float a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = threadIdx.x * 1e-3f + k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) a[k] = fmaf(a[k], b, c);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_rcp_chains(float* out, int iters) {
    float a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = 1.5f + threadIdx.x * 1e-3f + k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) a[k] = __builtin_amdgcn_rcpf(a[k]);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// per point: x = fma(l, A, X0), s = fma(x, x, E2), den = fma(s, s, H),
// num = fma(s, Kl, Kl), acc = fma(num, rcp(den), acc); X0 moves by 64 A per
// step (one more instruction per 16 points), so nothing is loop-invariant
__global__ __launch_bounds__(256) void k_region1_chains(float* out, float A, float E2, float H,
                                                       float Kl, int iters) {
    float acc[kChains], l[kChains];
    float X0 = 40.0f + threadIdx.x * 1e-2f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
        acc[k] = 0.0f;
        l[k] = -(float)k;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
            const float x = fmaf(l[k], A, X0);
            const float s = fmaf(x, x, E2);
            const float den = fmaf(s, s, H);
            const float num = fmaf(s, Kl, Kl);
            acc[k] = fmaf(num, __builtin_amdgcn_rcpf(den), acc[k]);
        }
        X0 = fmaf(-64.0f, A, X0);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

// kind 0: FMA chains, 1: rcp chains, 2: region-I sequence.  Launches `blocks`
// 256-thread workgroups of `iters` steps of 16 chains once to warm up, then
// times `reps` launches with HIP events; *ms = average per launch.  Every
// thread does iters * 16 chain steps.  Returns a hipError_t.
extern "C" int gexs_peak_run(int kind, int blocks, int iters, int reps, float* ms) {
    float* out = nullptr;
    hipError_t e = hipMalloc(&out, (size_t)blocks * 256 * sizeof(float));
    if (e != hipSuccess) return (int)e;
    hipEvent_t a = nullptr, b = nullptr;
    hipEventCreate(&a);
    hipEventCreate(&b);
    auto launch = [&]() {
        if (kind == 0)
            k_fma_chains<<<blocks, 256>>>(out, 0.999f, 1e-3f, iters);
        else if (kind == 1)
            k_rcp_chains<<<blocks, 256>>>(out, iters);
        else
            k_region1_chains<<<blocks, 256>>>(out, 1.37f, 63.5f, 128.0f, 0.01f, iters);
    };
    launch();
    hipEventRecord(a, nullptr);
    for (int r = 0; r < reps; ++r) launch();
    hipEventRecord(b, nullptr);
    e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipGetLastError();
    float t = 0.0f;
    hipEventElapsedTime(&t, a, b);
    *ms = t / (float)reps;
    hipEventDestroy(a);
    hipEventDestroy(b);
    hipFree(out);
    return (int)e;
}
