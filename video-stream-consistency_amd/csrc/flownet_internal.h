// Launch helpers of the flow network kernels (flownet_ops.hip), used by the
// network driver (flownet.hip).  Element kinds are EK_F32 / EK_BF16.
#pragma once

#include "vsc_internal.h"

namespace vsc {

// the most (first, second) pairs one decode batches: the forward and backward
// flows of up to 4 frames
constexpr int FLOWNET_MAX_PAIRS = 8;

// f1_c0 >= 0: also copies f1 (C channels) into `out` at channel f1_c0;
// f1, f2: N <= FLOWNET_MAX_PAIRS items; bf16_products (FP32 storage): the
// products on the BF16 MFMA (features rounded to BF16 as staged)
int launch_corr_nhwc(int ek, const void* const* f1, const void* const* f2, int cs1, int cs2, int C, int N, int H,
                     int W, void* out, int ocs, int f1_c0, hipStream_t s, bool bf16_products = false);
// upflow fused into the warp (moduleUpflow, pwcnet.py:472, then Backward)
struct UpflowArgs {
    const float* in;           // coarser flow, [N][H/2][W/2][2] FP32
    const float* w;            // ConvTranspose2d weights [2][2][4][4]
    const float* b;            // bias [2]
    float* out;                // upsampled flow, [N][H][W][2] FP32
    void* cat; int ccs, cc0;   // its copy in the concatenation buffer (element kind)
};
// up: nullptr, or the warp flow is computed from up->in (flow/fcs unused)
int launch_warp_nhwc(int ek, const void* const* src, int scs, const float* flow, int fcs, float scale, void* out,
                     int ocs, int C, int N, int H, int W, hipStream_t s, const UpflowArgs* up = nullptr);
int launch_preprocess(int ek, const uint8_t* rgba, int W, int H, int Wd, int Hd, int Wp, int Hp, void* out, int ocs,
                      hipStream_t s);
// fl: N (1 .. FLOWNET_MAX_PAIRS) level-3 flows [N][h][w][2]; out[i]: the i-th flow at Wd x Hd
int launch_flow_final(const float* fl, int N, int h, int w, int Hp, int Wp, float* const* out, int Wd, int Hd,
                      int out_ch, int swap_scale, hipStream_t s);

}  // namespace vsc
