// launch_adjoint.hip -- launcher of the reverse stage kernel.
#include "launch.hpp"

namespace hrw {
namespace {

template <int R, bool AMAX>
static void ladj(const AdjArgs &g, hipStream_t s)
{
    const unsigned grid = (unsigned)((g.nchunks * 64 + kBlock - 1) / kBlock);
    adjoint_stage_kernel<R, AMAX><<<grid, kBlock, 0, s>>>(g);
}

}  // namespace

void launch_adjoint_stage(int R, bool alphamax, const AdjArgs &g, hipStream_t s)
{
    if (R == 4) {
        if (alphamax) ladj<4, true>(g, s);
        else ladj<4, false>(g, s);
    } else {
        if (alphamax) ladj<2, true>(g, s);
        else ladj<2, false>(g, s);
    }
}

}  // namespace hrw
