// tiles_gauss.hip -- the bin-averaged Gaussian instantiations of the profile
// kernel (Doppler and the fixed-HWHM Gaussian).
#include "profile_tiles.h"

// piece k of the erfc table (FP64 on the device, stored in FP32)
__global__ __launch_bounds__(256) void k_erfc_pieces() {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= gx::kErfcPieces) return;
    const double4 c = gx::erfc_piece(k);
    gx_erfc_pieces[k] = make_float4((float)c.x, (float)c.y, (float)c.z, (float)c.w);
}

namespace gxk {

hipError_t init_gauss_tables(hipStream_t s) {
    hipLaunchKernelGGL(k_erfc_pieces, dim3((gx::kErfcPieces + 255) / 256), dim3(256), 0, s);
    return hipGetLastError();
}

hipError_t profile_tiles_gauss(hipStream_t s, int ppl, int split, const WorkItem* items, int nitems,
                               const Job* jobs, int single, const unsigned long long* scale_bits,
                               const int* kmax, unsigned int* xq, const unsigned int* xqb, int mode, int blocks) {
    return tile_launch::launch_prof<GX_PROF_DOPPLER>(s, ppl, split, items, nitems, jobs, single, scale_bits, kmax, xq,
                                                     xqb, mode, blocks);
}

}  // namespace gxk
