// launch_kernels.hip -- launchers of the LDS-tile 1D stage, the 2D stage and the batch stage.
#include "launch.hpp"

namespace hrw {
namespace {

constexpr int kCPT = 4;   // cells per thread in the 1D stage kernel

template <int K, int FLUX, bool PER, bool FAST, bool AMAX>
static void l1d(const Stage1DArgs &a, hipStream_t s)
{
    constexpr int TILE = kBlock * kCPT - 2;
    const unsigned grid = (unsigned)((a.n + TILE - 1) / TILE);
    stage1d_kernel<K, FLUX, PER, FAST, AMAX, kCPT><<<grid, kBlock, 0, s>>>(a);
}
template <int K, int FLUX, bool PER, bool FAST>
static void l1d_a(bool amax, const Stage1DArgs &a, hipStream_t s)
{
    if constexpr (FLUX == LF) {
        if (amax) return l1d<K, FLUX, PER, FAST, true>(a, s);
    }
    l1d<K, FLUX, PER, FAST, false>(a, s);
}
template <int K, int FLUX, bool PER>
static void l1d_f(bool fast, bool amax, const Stage1DArgs &a, hipStream_t s)
{
    if (fast) l1d_a<K, FLUX, PER, true>(amax, a, s);
    else l1d_a<K, FLUX, PER, false>(amax, a, s);
}
template <int K, int FLUX>
static void l1d_p(bool per, bool fast, bool amax, const Stage1DArgs &a, hipStream_t s)
{
    if (per) l1d_f<K, FLUX, true>(fast, amax, a, s);
    else l1d_f<K, FLUX, false>(fast, amax, a, s);
}
template <int K>
static void l1d_x(int flux, bool per, bool fast, bool amax, const Stage1DArgs &a, hipStream_t s)
{
    if (flux == LF) l1d_p<K, LF>(per, fast, amax, a, s);
    else l1d_p<K, GODUNOV>(per, fast, amax, a, s);
}
// Rows per wavefront: the tuned value, else the fewest (at least 8) that let every wavefront
// of the launch be resident at once.
template <int K, int BC, bool FAST>
static void l2d(Stage2DArgs a, hipStream_t s)
{
    if (a.ty <= 0) {
        static int per_cu = 0;
        if (!per_cu) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pbe2d_kernel<K, BC, FAST>, kBlock, 0) !=
                    hipSuccess || per_cu <= 0)
                per_cu = 4;
        }
        const long long slots = (long long)per_cu * device_cus() * (kBlock / 64);
        const long long bands = std::max<long long>(1, slots / a.nstrips);
        a.ty = (int)std::max<long long>(8, (a.n2 + bands - 1) / bands);
    }
    const long long nsy = (a.n2 + a.ty - 1) / a.ty;
    const long long waves = nsy * a.nstrips;
    const unsigned grid = (unsigned)((waves * 64 + kBlock - 1) / kBlock);
    pbe2d_kernel<K, BC, FAST><<<grid, kBlock, 0, s>>>(a);
}
template <int K>
static void l2d_b(int bc, bool fast, const Stage2DArgs &a, hipStream_t s)
{
    if (bc == HRW_BC_ZERO_FLUX) {
        if (fast) l2d<K, 0, true>(a, s);
        else l2d<K, 0, false>(a, s);
    } else {
        if (fast) l2d<K, 1, true>(a, s);
        else l2d<K, 1, false>(a, s);
    }
}
template <int FLUX, bool AMAX, int R>
static void lbatch_r(const BatchArgs &a, hipStream_t s)
{
    const unsigned grid = (unsigned)((a.nchunks * 64 + kBlock - 1) / kBlock);
    switch (stage_npw(a.stage)) {
    case 1: batch_stage_kernel<FLUX, AMAX, R, 1><<<grid, kBlock, 0, s>>>(a); break;
    default: batch_stage_kernel<FLUX, AMAX, R, 0><<<grid, kBlock, 0, s>>>(a); break;
    }
}
// Persistent form: as many blocks as fit on the device at once (occupancy of this
// instantiation, asked once), never more than there are chunks.
template <int FLUX, bool AMAX, int R, int NPW>
static void lbatch_pf_n(const BatchArgs &a, hipStream_t s)
{
    static int per_cu = 0;
    if (!per_cu) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, reinterpret_cast<const void *>(&batch_pf_kernel<FLUX, AMAX, R, NPW>), kBlock, 0) != hipSuccess ||
            nb <= 0)
            nb = 2;
        per_cu = nb;
    }
    const long long want = (a.nchunks + kBlock / 64 - 1) / (kBlock / 64);
    const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(want, (long long)per_cu * device_cus()));
    batch_pf_kernel<FLUX, AMAX, R, NPW><<<grid, kBlock, 0, s>>>(a);
}
template <int FLUX, bool AMAX, int R>
static void lbatch_pf_r(const BatchArgs &a, hipStream_t s)
{
    switch (stage_npw(a.stage)) {
    case 1: lbatch_pf_n<FLUX, AMAX, R, 1>(a, s); break;
    default: lbatch_pf_n<FLUX, AMAX, R, 0>(a, s); break;
    }
}
template <int FLUX, bool AMAX>
static void lbatch(int R, int form, const BatchArgs &a, hipStream_t s)
{
    if (form == HRW_BATCH_PERSISTENT) {
        if (R == 4) lbatch_pf_r<FLUX, AMAX, 4>(a, s);
        else if (R == 6) lbatch_pf_r<FLUX, AMAX, 6>(a, s);
        else lbatch_pf_r<FLUX, AMAX, 8>(a, s);
        return;
    }
    if (R == 4) lbatch_r<FLUX, AMAX, 4>(a, s);
    else if (R == 6) lbatch_r<FLUX, AMAX, 6>(a, s);
    else lbatch_r<FLUX, AMAX, 8>(a, s);
}

}  // namespace

void launch_tile_1d(int k, int flux, bool periodic, bool fast, bool alphamax, const Stage1DArgs &a,
                    hipStream_t s)
{
    if (k == 1) l1d_x<1>(flux, periodic, fast, alphamax, a, s);
    else if (k == 2) l1d_x<2>(flux, periodic, fast, alphamax, a, s);
    else l1d_x<3>(flux, periodic, fast, alphamax, a, s);
}

void launch_pbe2d(int k, int bc, bool fast, const Stage2DArgs &a, hipStream_t s)
{
    if (k == 1) l2d_b<1>(bc, fast, a, s);
    else if (k == 2) l2d_b<2>(bc, fast, a, s);
    else l2d_b<3>(bc, fast, a, s);
}

void launch_batch_stage(int flux, bool alphamax, int R, int form, const BatchArgs &a, hipStream_t s)
{
    if (flux == LF && alphamax) lbatch<LF, true>(R, form, a, s);
    else if (flux == LF) lbatch<LF, false>(R, form, a, s);
    else lbatch<GODUNOV, false>(R, form, a, s);
}

}  // namespace hrw
