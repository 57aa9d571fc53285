// launch_adjoint.hip -- launchers of the reverse stage kernels.
#include "launch.hpp"

namespace hrw {
namespace {

template <int R, bool AMAX>
static void ladj(const AdjArgs &g, hipStream_t s)
{
    const unsigned grid = (unsigned)((g.nchunks * 64 + kBlock - 1) / kBlock);
    adjoint_stage_kernel<R, AMAX><<<grid, kBlock, 0, s>>>(g);
}

template <int K, int FLUX, bool PER, bool AMAX>
static void lgen(const GenAdjArgs &g, hipStream_t s)
{
    const unsigned grid = (unsigned)((g.n + kGenTile - 1) / kGenTile);
    rhs_vjp_kernel<K, FLUX, PER, AMAX><<<grid, kGenTile, 0, s>>>(g);
}
template <int K>
static void lgen_k(int flux, bool periodic, bool amax, const GenAdjArgs &g, hipStream_t s)
{
    if (flux == LF) {
        if (periodic) {
            if (amax) lgen<K, LF, true, true>(g, s);
            else lgen<K, LF, true, false>(g, s);
        } else {
            if (amax) lgen<K, LF, false, true>(g, s);
            else lgen<K, LF, false, false>(g, s);
        }
    } else {
        if (periodic) lgen<K, GODUNOV, true, false>(g, s);
        else lgen<K, GODUNOV, false, false>(g, s);
    }
}

}  // namespace

void launch_rhs_vjp(int k, int flux, bool periodic, bool alphamax, const GenAdjArgs &g,
                    hipStream_t s)
{
    if (k == 1) lgen_k<1>(flux, periodic, alphamax, g, s);
    else if (k == 2) lgen_k<2>(flux, periodic, alphamax, g, s);
    else lgen_k<3>(flux, periodic, alphamax, g, s);
}

template <int K>
static void lvjp5(int flux, bool periodic, bool alphamax, const BvjpArgs &g, hipStream_t s)
{
    const unsigned grid = (unsigned)((g.nchunks + 3) / 4);
    if (flux == LF) {
        if (alphamax) {
            if (periodic) rhs_vjp5_kernel<K, LF, true, true><<<grid, 256, 0, s>>>(g);
            else rhs_vjp5_kernel<K, LF, false, true><<<grid, 256, 0, s>>>(g);
        } else {
            if (periodic) rhs_vjp5_kernel<K, LF, true, false><<<grid, 256, 0, s>>>(g);
            else rhs_vjp5_kernel<K, LF, false, false><<<grid, 256, 0, s>>>(g);
        }
    } else {
        if (periodic) rhs_vjp5_kernel<K, GODUNOV, true, false><<<grid, 256, 0, s>>>(g);
        else rhs_vjp5_kernel<K, GODUNOV, false, false><<<grid, 256, 0, s>>>(g);
    }
}

void launch_rhs_vjp5(int k, int flux, bool periodic, bool alphamax, const BvjpArgs &g,
                     hipStream_t s)
{
    if (g.nchunks <= 0) return;
    if (k == 2) lvjp5<2>(flux, periodic, alphamax, g, s);
    else lvjp5<3>(flux, periodic, alphamax, g, s);
}

void launch_rhs_vjp5_fixup(const BvjpArgs &g, hipStream_t s)
{
    const int nblocks = (int)((g.nchunks + 3) / 4);
    rhs_vjp_alpha_fixup<<<1, kGenTile, 0, s>>>(g.partial, nblocks, g.argmax, g.aset, g.s, g.out, g.ch);
}

void launch_rhs_vjp_fixup(const GenAdjArgs &g, hipStream_t s)
{
    const int nblocks = (int)((g.n + kGenTile - 1) / kGenTile);
    rhs_vjp_alpha_fixup<<<1, kGenTile, 0, s>>>(g.partial, nblocks, g.argmax, g.aset, g.s, g.out, g.ch);
}

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
