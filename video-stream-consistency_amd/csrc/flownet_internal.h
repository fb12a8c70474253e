// Launch helpers of the flow network kernels (flownet_ops.hip), used by the
// network driver (flownet.hip).  Element kinds are EK_F32 / EK_BF16.
#pragma once

#include "vsc_internal.h"

namespace vsc {

// f1_c0 >= 0: also copies f1 (C channels) into `out` at channel f1_c0
int launch_corr_nhwc(int ek, const void* const f1[2], const void* const f2[2], int cs1, int cs2, int C, int N, int H,
                     int W, void* out, int ocs, int f1_c0, hipStream_t s);
int launch_warp_nhwc(int ek, const void* const src[2], int scs, const float* flow, int fcs, float scale, void* out,
                     int ocs, int C, int N, int H, int W, hipStream_t s);
int launch_upflow(int ek, const float* in, int N, int Hi, int Wi, const float* w, const float* b, float* out, void* cat,
                  int ccs, int cc0, hipStream_t s);
int launch_preprocess(int ek, const uint8_t* rgba, int W, int H, int Wd, int Hd, int Wp, int Hp, void* out, int ocs,
                      hipStream_t s);
int launch_flow_final(const float* fl, int h, int w, int Hp, int Wp, float* out, int Wd, int Hd, int out_ch,
                      int swap_scale, hipStream_t s);

}  // namespace vsc
