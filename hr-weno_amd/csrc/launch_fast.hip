// launch_fast.hip -- launchers of the register-blocked 1D stage kernels for one WENO order,
// HRW_K (the build compiles this file once per K = 1, 2, 3, in parallel).
#include "launch.hpp"

#ifndef HRW_K
#error "compile with -DHRW_K=1, 2 or 3"
#endif

namespace hrw {
namespace {

template <int K, int FLUX, bool PER, bool AMAX, int R, bool CW>
static void l1f(const Stage1DArgs &a, hipStream_t s)
{
    constexpr long long per_block = (long long)kOutLanes * R * (kBlock / 64);
    const unsigned grid = (unsigned)((a.n + per_block - 1) / per_block);
    switch (stage_npw(a.stage)) {
    case 2: stage1d_fast_kernel<K, FLUX, PER, AMAX, R, 2, CW><<<grid, kBlock, 0, s>>>(a); break;
    case 1: stage1d_fast_kernel<K, FLUX, PER, AMAX, R, 1, CW><<<grid, kBlock, 0, s>>>(a); break;
    default: stage1d_fast_kernel<K, FLUX, PER, AMAX, R, 0, CW><<<grid, kBlock, 0, s>>>(a); break;
    }
}
// Persistent form: as many blocks as fit on the device at once (occupancy of this
// instantiation, asked once), never more than there are chunks.
template <int K, int FLUX, bool PER, bool AMAX, int R, int NPW, bool CW>
static void l1p_n(const Stage1DArgs &a, hipStream_t s)
{
    static int per_cu = 0;
    if (!per_cu) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, reinterpret_cast<const void *>(&stage1d_pf_kernel<K, FLUX, PER, AMAX, R, NPW, CW>), kBlock, 0) != hipSuccess ||
            nb <= 0)
            nb = 2;
        per_cu = nb;
    }
    const long long chunks = (a.n + (long long)kOutLanes * R - 1) / ((long long)kOutLanes * R);
    const long long want = (chunks + kBlock / 64 - 1) / (kBlock / 64);
    const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(want, (long long)per_cu * device_cus()));
    stage1d_pf_kernel<K, FLUX, PER, AMAX, R, NPW, CW><<<grid, kBlock, 0, s>>>(a);
}
template <int K, int FLUX, bool PER, bool AMAX, int R, bool CW>
static void l1p(const Stage1DArgs &a, hipStream_t s)
{
    switch (stage_npw(a.stage)) {
    case 2: l1p_n<K, FLUX, PER, AMAX, R, 2, CW>(a, s); break;
    case 1: l1p_n<K, FLUX, PER, AMAX, R, 1, CW>(a, s); break;
    default: l1p_n<K, FLUX, PER, AMAX, R, 0, CW>(a, s); break;
    }
}
template <int K, int FLUX, bool PER, bool AMAX, bool CW>
static void l1f_r(int form, int r, const Stage1DArgs &a, hipStream_t s)
{
    if (form != FORM_CHUNK) {
        if (r == 4) l1p<K, FLUX, PER, AMAX, 4, CW>(a, s);
        else if (r == 6) l1p<K, FLUX, PER, AMAX, 6, CW>(a, s);
        else l1p<K, FLUX, PER, AMAX, 8, CW>(a, s);
        return;
    }
    if (r == 4) l1f<K, FLUX, PER, AMAX, 4, CW>(a, s);
    else if (r == 6) l1f<K, FLUX, PER, AMAX, 6, CW>(a, s);
    else l1f<K, FLUX, PER, AMAX, 8, CW>(a, s);
}
template <int K, int FLUX, bool PER, bool AMAX>
static void l1f_w(bool constw, int form, int r, const Stage1DArgs &a, hipStream_t s)
{
    if (constw) l1f_r<K, FLUX, PER, AMAX, true>(form, r, a, s);
    else l1f_r<K, FLUX, PER, AMAX, false>(form, r, a, s);
}
template <int K, int FLUX, bool PER>
static void l1f_a(bool amax, bool constw, int form, int r, const Stage1DArgs &a, hipStream_t s)
{
    if constexpr (FLUX == LF) {
        if (amax) return l1f_w<K, FLUX, PER, true>(constw, form, r, a, s);
    }
    l1f_w<K, FLUX, PER, false>(constw, form, r, a, s);
}
template <int K>
void l1f_x(int flux, bool per, bool amax, bool constw, int form, int r, const Stage1DArgs &a,
           hipStream_t s)
{
    if (flux == LF) {
        if (per) l1f_a<K, LF, true>(amax, constw, form, r, a, s);
        else l1f_a<K, LF, false>(amax, constw, form, r, a, s);
    } else {
        if (per) l1f_a<K, GODUNOV, true>(amax, constw, form, r, a, s);
        else l1f_a<K, GODUNOV, false>(amax, constw, form, r, a, s);
    }
}


}  // namespace

#define HRW_CAT2(a, b) a##b
#define HRW_CAT(a, b) HRW_CAT2(a, b)
void HRW_CAT(launch_fast_1d_k, HRW_K)(int flux, bool periodic, bool alphamax, bool constw,
                                      int form, int r, const Stage1DArgs &a, hipStream_t s)
{
    l1f_x<HRW_K>(flux, periodic, alphamax, constw, form, r, a, s);
}

}  // namespace hrw
