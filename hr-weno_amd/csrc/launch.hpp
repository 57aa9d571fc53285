// launch.hpp -- host launchers of the stage kernels, each compiled in a translation unit of
// its own (they instantiate most of the device code) and called by the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/hrweno_capi.h"
#include "kernels_1d.hpp"
#include "kernels_1d_fast.hpp"
#include "kernels_2d.hpp"
#include "kernels_adjoint.hpp"
#include "kernels_adjoint_general.hpp"
#include "kernels_adjoint_blocked.hpp"

namespace hrw {

// Compute units of the current device (asked once).
inline int device_cus()
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

// Register-blocked 1D stage (FAST form).  form: one chunk per wavefront (stage1d_fast_kernel)
// or persistent with the next chunk's loads issued ahead (stage1d_pf_kernel); r cells per lane
// (4, 6, 8); constw: every edge of the linear grid is exact, so every width is rx.  One
// definition per K (launch_fast_k<K>.o).
enum { FORM_CHUNK = 0, FORM_PERSISTENT = 1 };
void launch_fast_1d_k1(int flux, bool periodic, bool alphamax, bool constw, int form, int r,
                       const Stage1DArgs &a, hipStream_t s);
void launch_fast_1d_k2(int flux, bool periodic, bool alphamax, bool constw, int form, int r,
                       const Stage1DArgs &a, hipStream_t s);
void launch_fast_1d_k3(int flux, bool periodic, bool alphamax, bool constw, int form, int r,
                       const Stage1DArgs &a, hipStream_t s);
// LDS-tile 1D stage (EXACT form, and the FAST LDS variant).
void launch_tile_1d(int k, int flux, bool periodic, bool fast, bool alphamax, const Stage1DArgs &a,
                    hipStream_t s);
// 2D population balance stage.
void launch_pbe2d(int k, int bc, bool fast, const Stage2DArgs &a, hipStream_t s);
// One stage of every line of a batch (form HRW_BATCH_CHUNK or HRW_BATCH_PERSISTENT).
void launch_batch_stage(int flux, bool alphamax, int R, int form, const BatchArgs &a, hipStream_t s);
// The reverse of one stage (periodic Lax-Friedrichs lines, register-blocked).
void launch_adjoint_stage(int R, bool alphamax, const AdjArgs &g, hipStream_t s);
// out = sum e_i x_i + ch J(s)^T (b1 y1 + b2 y2) for any 1D Burgers composition on a linear grid,
// and (alpha = max|s|) the alpha-bar correction once that pass has finished.
void launch_rhs_vjp(int k, int flux, bool periodic, bool alphamax, const GenAdjArgs &g,
                    hipStream_t s);
void launch_rhs_vjp_fixup(const GenAdjArgs &g, hipStream_t s);
// The same for K = 2 or 3, register-blocked: one wavefront per 60*kBvjpR cells (and its
// alpha-bar correction for alpha = max|s|).
void launch_rhs_vjp5(int k, int flux, bool periodic, bool alphamax, const BvjpArgs &g,
                     hipStream_t s);
void launch_rhs_vjp5_fixup(const BvjpArgs &g, hipStream_t s);

}  // namespace hrw
