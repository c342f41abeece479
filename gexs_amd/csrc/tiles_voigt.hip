// tiles_voigt.hip -- the Voigt (Humlicek W4) instantiations of the profile kernel.
#include "profile_tiles.h"

namespace gxk {

hipError_t profile_tiles_voigt(hipStream_t s, int ppl, int split, const WorkItem* items, int nitems,
                               const Job* jobs, int single, const unsigned long long* scale_bits,
                               const int* kmax, unsigned int* xq, const unsigned int* xqb, int mode, int blocks) {
    return tile_launch::launch_prof<GX_PROF_VOIGT>(s, ppl, split, items, nitems, jobs, single, scale_bits, kmax, xq,
                                                   xqb, mode, blocks);
}

}  // namespace gxk
