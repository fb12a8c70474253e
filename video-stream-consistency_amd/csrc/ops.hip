// ORT custom-op kernels for gfx950: "Correlation" (9x9 cost volume) and
// "Warp" (bilinear backward warp), NCHW in / NCHW out, with the attribute and
// shape contract of src/ort_custom_ops.
#include "vsc_internal.h"

namespace vsc {

// ---------------------------------------------------------------------------
// Correlation, FP32, sequential-in-c accumulation (the order of the reference
// CPU op, correlation.cc:163-177: *dst = 0; *dst += v1*v2 for c = 0..C-1).
//
// A 256-thread block owns a 64 x 4 tile of output pixels of one image and all
// (2*MD+1)^2 displacements; thread (tx, ty) keeps the P*P accumulators of its
// pixel in registers.  Channels are streamed in chunks of CC: the f2 tile with
// an MD-pixel halo, (4+2MD) x (64+2MD) x CC, is staged in LDS with zeros
// outside the image (an out-of-range term of the reference contributes nothing,
// and acc + (+-0) == acc for the accumulator values that occur, so zero terms
// leave every bit unchanged).  The transpose to the LDS layout replaces the
// reference's channels-last rearrange kernels and their per-call cudaMalloc
// (correlation_cuda.cu:32-95,365-370).  Output rows are written coalesced
// along W for each displacement plane.
// ---------------------------------------------------------------------------
static constexpr int CT_W = 64, CT_H = 4, CT_CC = 8;

template <int MD, bool LEGACY>
__global__ __launch_bounds__(256) void k_corr_f32(const float* __restrict__ f1, const float* __restrict__ f2,
                                                  float* __restrict__ out, int C, int H, int W)
{
    constexpr int P = 2 * MD + 1;
    constexpr int SW = CT_W + 2 * MD, SH = CT_H + 2 * MD;
    __shared__ float s2[CT_CC][SH][SW];

    const int n = blockIdx.z;
    const int tx = threadIdx.x % CT_W, ty = threadIdx.x / CT_W;
    const int w0 = blockIdx.x * CT_W, h0 = blockIdx.y * CT_H;
    const int w = w0 + tx, h = h0 + ty;
    const long long HW = (long long)H * W;
    const float* a = f1 + (long long)n * C * HW;
    const float* b = f2 + (long long)n * C * HW;
    const bool valid = (w < W) && (h < H);

    float acc[P * P];
#pragma unroll
    for (int d = 0; d < P * P; ++d) acc[d] = 0.0f;

    for (int c0 = 0; c0 < C; c0 += CT_CC) {
        __syncthreads();
        for (int e = threadIdx.x; e < CT_CC * SH * SW; e += 256) {
            const int cc = e / (SH * SW);
            const int rem = e - cc * SH * SW;
            const int r = rem / SW, q = rem - r * SW;
            const int gh = h0 - MD + r, gw = w0 - MD + q, gc = c0 + cc;
            float v = 0.0f;
            if (gc < C && gh >= 0 && gh < H && gw >= 0 && gw < W) v = b[gc * HW + (long long)gh * W + gw];
            s2[cc][r][q] = v;
        }
        __syncthreads();
        const int cn = min(CT_CC, C - c0);
        for (int cc = 0; cc < cn; ++cc) {
            const float v1 = valid ? a[(c0 + cc) * HW + (long long)h * W + w] : 0.0f;
#pragma unroll
            for (int dy = 0; dy < P; ++dy)
#pragma unroll
                for (int dx = 0; dx < P; ++dx) acc[dy * P + dx] += v1 * s2[cc][ty + dy][tx + dx];
        }
    }
    if (!valid) return;
    float* o = out + (long long)n * P * P * HW + (long long)h * W + w;
#pragma unroll
    for (int d = 0; d < P * P; ++d) o[d * HW] = LEGACY ? acc[d] / (float)C : acc[d];
}

// Any displacement, FP32 or FP64: one thread per output element, sequential
// in c (correlation.cc:147-182).  Used for max_displacement > 4 and FP64.
template <typename T, bool LEGACY>
__global__ __launch_bounds__(256) void k_corr_generic(const T* __restrict__ f1, const T* __restrict__ f2,
                                                      T* __restrict__ out, int N, int C, int H, int W, int md)
{
    const int P = 2 * md + 1;
    const long long HW = (long long)H * W;
    const long long total = (long long)N * P * P * HW;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int w = (int)(i % W);
    const int h = (int)((i / W) % H);
    const long long nd = i / HW;
    const int d = (int)(nd % (P * P));
    const int n = (int)(nd / (P * P));
    const int dy = d / P - md, dx = d % P - md;
    const int h2 = h + dy, w2 = w + dx;
    T acc = 0;
    if (h2 >= 0 && h2 < H && w2 >= 0 && w2 < W) {
        const T* a = f1 + (long long)n * C * HW + (long long)h * W + w;
        const T* b = f2 + (long long)n * C * HW + (long long)h2 * W + w2;
        for (int c = 0; c < C; ++c) acc += a[c * HW] * b[c * HW];
    }
    out[i] = LEGACY ? acc / (T)C : acc;
}

template <int MD>
static void launch_corr_f32(const float* f1, const float* f2, float* out, int N, int C, int H, int W, bool legacy,
                            hipStream_t s)
{
    dim3 grid((W + CT_W - 1) / CT_W, (H + CT_H - 1) / CT_H, N);
    if (legacy)
        hipLaunchKernelGGL((k_corr_f32<MD, true>), grid, dim3(256), 0, s, f1, f2, out, C, H, W);
    else
        hipLaunchKernelGGL((k_corr_f32<MD, false>), grid, dim3(256), 0, s, f1, f2, out, C, H, W);
}

// ---------------------------------------------------------------------------
// Warp (warp.cc:71-134 / warp_cuda.cu:28-84).  One thread per output pixel of
// one image: the flow is read and the bilinear weights and in-bounds mask are
// computed once, then the C planes are gathered (the reference re-reads the
// flow once per channel, grid.y = C).  FP32 weights; the reference evaluates
// weights in double for float tensors, so FP32 agrees to rounding, and the
// keep/zero decision (double)mask > 0.999 becomes mask >= 0.999f, the same
// threshold on the float grid (0.999f is the first float above 0.999).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_warp(const T* __restrict__ in, const T* __restrict__ flow, T* __restrict__ out,
                                              int C, int H, int W)
{
    const int x = blockIdx.x * 64 + threadIdx.x % 64;
    const int y = blockIdx.y * 4 + threadIdx.x / 64;
    const int n = blockIdx.z;
    if (x >= W || y >= H) return;
    const long long HW = (long long)H * W;
    const long long p = (long long)y * W + x;
    const T xf = (T)x + flow[(long long)n * 2 * HW + p];
    const T yf = (T)y + flow[(long long)n * 2 * HW + HW + p];
    const T xL = floor(xf), yT = floor(yf);
    const T alpha = xf - xL, beta = yf - yT;
    const T xR = xL + (T)1, yB = yT + (T)1;
    const bool mL = (T)0 <= xL && xL <= (T)(W - 1);
    const bool mR = (T)0 <= xR && xR <= (T)(W - 1);
    const bool mT = (T)0 <= yT && yT <= (T)(H - 1);
    const bool mB = (T)0 <= yB && yB <= (T)(H - 1);
    const T wTL = ((T)1 - alpha) * ((T)1 - beta);
    const T wTR = alpha * ((T)1 - beta);
    const T wBL = ((T)1 - alpha) * beta;
    const T wBR = alpha * beta;
    T mask = (T)0;
    mask = mask + (mT && mL ? wTL : (T)0);
    mask = mask + (mT && mR ? wTR : (T)0);
    mask = mask + (mB && mL ? wBL : (T)0);
    mask = mask + (mB && mR ? wBR : (T)0);
    bool keep;
    if constexpr (sizeof(T) == 4) keep = mask >= 0.999f;
    else keep = mask > 0.999;
    const long long base = (long long)n * C * HW;
    T* o = out + base + p;
    if (!keep) {
        for (int c = 0; c < C; ++c) o[c * HW] = (T)0;
        return;
    }
    const int iL = (int)xL, iR = (int)xR, iT = (int)yT, iB = (int)yB;
    const long long oTL = (long long)iT * W + iL, oTR = (long long)iT * W + iR;
    const long long oBL = (long long)iB * W + iL, oBR = (long long)iB * W + iR;
    const T* src = in + base;
    for (int c = 0; c < C; ++c) {
        const T* pl = src + c * HW;
        T v = (T)0;
        if (mT && mL) v = v + wTL * pl[oTL];
        if (mT && mR) v = v + wTR * pl[oTR];
        if (mB && mL) v = v + wBL * pl[oBL];
        if (mB && mR) v = v + wBR * pl[oBR];
        o[c * HW] = v;
    }
}

}  // namespace vsc

using namespace vsc;

extern "C" int vsc_correlation_forward(const void* f1, const void* f2, void* out, int N, int C, int H, int W,
                                       int max_displacement, int legacy, int in_dtype, const float* chan_scale,
                                       vsc_hip_stream stream)
{
    if (!f1 || !f2 || !out) return set_error(VSC_E_INVALID, "vsc_correlation_forward: null pointer");
    if (N < 1 || C < 1 || H < 1 || W < 1) return set_error(VSC_E_INVALID, "vsc_correlation_forward: empty tensor");
    if (max_displacement < 0) return set_error(VSC_E_INVALID, "vsc_correlation_forward: max_displacement < 0");
    if (legacy && max_displacement > 4)
        return set_error(VSC_E_INVALID, "vsc_correlation_forward: legacy correlation pads by 4, max_displacement <= 4");
    const int md = max_displacement;
    hipStream_t s = as_stream(stream);
    if (in_dtype == VSC_F32) {
        const float* a = static_cast<const float*>(f1);
        const float* b = static_cast<const float*>(f2);
        float* o = static_cast<float*>(out);
        switch (md) {
            case 0: launch_corr_f32<0>(a, b, o, N, C, H, W, legacy, s); break;
            case 1: launch_corr_f32<1>(a, b, o, N, C, H, W, legacy, s); break;
            case 2: launch_corr_f32<2>(a, b, o, N, C, H, W, legacy, s); break;
            case 3: launch_corr_f32<3>(a, b, o, N, C, H, W, legacy, s); break;
            case 4: launch_corr_f32<4>(a, b, o, N, C, H, W, legacy, s); break;
            default: {
                const long long total = (long long)N * (2 * md + 1) * (2 * md + 1) * H * W;
                hipLaunchKernelGGL((k_corr_generic<float, false>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                                   s, a, b, o, N, C, H, W, md);
            }
        }
        return check_launch("correlation f32");
    }
    if (in_dtype == VSC_F64) {
        const long long total = (long long)N * (2 * md + 1) * (2 * md + 1) * H * W;
        const dim3 g((unsigned)((total + 255) / 256));
        const double* a = static_cast<const double*>(f1);
        const double* b = static_cast<const double*>(f2);
        double* o = static_cast<double*>(out);
        if (legacy)
            hipLaunchKernelGGL((k_corr_generic<double, true>), g, dim3(256), 0, s, a, b, o, N, C, H, W, md);
        else
            hipLaunchKernelGGL((k_corr_generic<double, false>), g, dim3(256), 0, s, a, b, o, N, C, H, W, md);
        return check_launch("correlation f64");
    }
    if (in_dtype == VSC_BF16 || in_dtype == VSC_F16 || in_dtype == VSC_FP8E4M3) {
        return corr_mfma_dispatch(f1, f2, static_cast<float*>(out), N, C, H, W, md, legacy, in_dtype, chan_scale, s);
    }
    return set_error(VSC_E_DTYPE, "vsc_correlation_forward: unsupported input type (float, double, bf16, f16, fp8e4m3)");
}

extern "C" int vsc_warp_forward(const void* in, const void* flow, void* out, int N, int C, int H, int W, int dtype,
                                vsc_hip_stream stream)
{
    if (!in || !flow || !out) return set_error(VSC_E_INVALID, "vsc_warp_forward: null pointer");
    if (N < 1 || C < 1 || H < 1 || W < 1) return set_error(VSC_E_INVALID, "vsc_warp_forward: empty tensor");
    const dim3 grid((W + 63) / 64, (H + 3) / 4, N);
    hipStream_t s = as_stream(stream);
    if (dtype == VSC_F32) {
        hipLaunchKernelGGL(k_warp<float>, grid, dim3(256), 0, s, static_cast<const float*>(in),
                           static_cast<const float*>(flow), static_cast<float*>(out), C, H, W);
        return check_launch("warp f32");
    }
    if (dtype == VSC_F64) {
        hipLaunchKernelGGL(k_warp<double>, grid, dim3(256), 0, s, static_cast<const double*>(in),
                           static_cast<const double*>(flow), static_cast<double*>(out), C, H, W);
        return check_launch("warp f64");
    }
    return set_error(VSC_E_DTYPE, "vsc_warp_forward: unsupported input type (float or double)");
}
