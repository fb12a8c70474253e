// Reduced-precision correlation on MFMA (BF16 / F16 / FP8 e4m3, FP32 accumulation).
#include "vsc_internal.h"

namespace vsc {

int corr_mfma_dispatch(const void* f1, const void* f2, float* out, int N, int C, int H, int W, int md, int legacy,
                       int in_dtype, const float* chan_scale, hipStream_t s)
{
    (void)f1; (void)f2; (void)out; (void)N; (void)C; (void)H; (void)W; (void)md; (void)legacy; (void)in_dtype;
    (void)chan_scale; (void)s;
    return set_error(VSC_E_UNSUPPORTED, "reduced-precision correlation is not built yet");
}

}  // namespace vsc

extern "C" int vsc_quantize_fp8_pair(const float*, const float*, uint8_t*, uint8_t*, float*, int, int, int, int,
                                     vsc_hip_stream)
{
    return vsc::set_error(VSC_E_UNSUPPORTED, "vsc_quantize_fp8_pair is not built yet");
}

extern "C" int vsc_convert_f32(const float*, void*, size_t, int, vsc_hip_stream)
{
    return vsc::set_error(VSC_E_UNSUPPORTED, "vsc_convert_f32 is not built yet");
}
