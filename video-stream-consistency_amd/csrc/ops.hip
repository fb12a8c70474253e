// ORT custom-op kernels for gfx950: "Correlation" (9x9 cost volume) and
// "Warp" (bilinear backward warp), NCHW in / NCHW out, with the attribute and
// shape contract of src/ort_custom_ops.
#include "vsc_internal.h"

#include <mutex>
#include <string>

namespace vsc {

// ---------------------------------------------------------------------------
// Correlation, FP32, sequential-in-c accumulation (the order of the reference
// CPU op, correlation.cc:163-177: *dst = 0; *dst += v1*v2 for c = 0..C-1).
//
// A block owns a tile of 8 rows x 8*QX pixels of one image and DG of the
// vertical displacements (a 576-thread block all 9; the small levels' tiles
// spread them over 3 blocks of 3 waves).  Wave d handles vertical
// displacement dy = d - MD; its lane (row, quad)
// keeps QX pixels x (2*MD+1) horizontal displacements in registers (QX = 4
// for large images: 12 f2 loads from LDS feed 36 multiply-adds; QX = 1 for
// small pyramid levels, where parallelism matters more than reuse).  Channels
// stream through LDS in chunks (8 channels for QX = 4, 32 for QX = 1): the f1 tile and the f2 tile (a
// 4-pixel halo in x, the rows the block's displacements reach), zeros
// outside the image (an out-of-range term of the
// reference contributes nothing, and acc + (+-0) == acc for the accumulator
// values that occur, so zero terms leave every bit unchanged).  Every thread
// issues all its global loads of the next chunk before computing the current
// one, so the chunk loop overlaps memory and arithmetic.  This staging replaces the reference's
// channels-last rearrange kernels and their per-call cudaMalloc
// (correlation_cuda.cu:32-95,365-370).  Each (dy, dx) output row segment of
// 8*QX pixels is written contiguously.
// ---------------------------------------------------------------------------
static constexpr int CF_TH = 8, CF_HALO = 4;
// vertical displacements per block of the 1-pixel tile (md = 4: three blocks
// of 3 waves per tile; 1920x1088 pyramid 134.0 us against 141.3 us for one
// block of 9 waves and 162.7 us for nine blocks of one wave)
static constexpr int CF_DG1 = 3;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

// DG: vertical displacements (waves) per block; the P of them are split over
// P / DG blocks (blockIdx.z = item * (P / DG) + group), each staging only the
// CF_TH + DG - 1 f2 rows its displacements read.
template <int MD, bool LEGACY, int QX, int CC, int DG>
__global__ __launch_bounds__(64 * DG) void k_corr_f32(const float* __restrict__ f1, const float* __restrict__ f2,
                                                      float* __restrict__ out, int C, int H, int W)
{
    constexpr int P = 2 * MD + 1;
    constexpr int NTH = 64 * DG, G = DG >= P ? 1 : P / DG;
    static_assert(DG >= P || P % DG == 0, "displacement groups of equal size");
    constexpr int TW = 8 * QX;
    constexpr int SW = TW + 2 * CF_HALO, SH = CF_TH + DG - 1;
    // f2 row pitch: the 1-pixel tile's lanes read 8 rows x 8 consecutive
    // pixels at once; a pitch of 24 floats puts those rows on distinct banks
    // (16 would stack them on the same 16)
    constexpr int SWP = QX == 1 ? 24 : SW;
    constexpr int N1 = CC * CF_TH * TW, N2 = CC * SH * SW;
    constexpr int L1 = (N1 + NTH - 1) / NTH, L2 = (N2 + NTH - 1) / NTH;
    __shared__ __attribute__((aligned(16))) float s1[CC][CF_TH][TW];
    __shared__ __attribute__((aligned(16))) float s2[CC][SH][SWP];

    const int n = blockIdx.z / G, dg = blockIdx.z % G;
    const int tid = threadIdx.x;
    const int dl = tid / 64, d = dg * DG + dl, rem = tid % 64;   // local and global displacement row
    const int row = rem / 8, quad = rem % 8;
    const int w0 = blockIdx.x * TW, h0 = blockIdx.y * CF_TH;
    const long long HW = (long long)H * W;
    const float* a = f1 + (long long)n * C * HW;
    const float* b = f2 + (long long)n * C * HW;

    float acc[QX][P];
#pragma unroll
    for (int x = 0; x < QX; ++x)
#pragma unroll
        for (int k = 0; k < P; ++k) acc[x][k] = 0.0f;

    // chunk staging: global loads from clamped addresses into registers (all
    // in flight together), LDS stores after the previous chunk's compute
    float r1[L1], r2[L2];
    auto load_chunk = [&](int c0) {
#pragma unroll
        for (int k = 0; k < L1; ++k) {
            const int e = tid + NTH * k;
            const int x = e % TW, r = (e / TW) % CF_TH, c = e / (TW * CF_TH);
            const int gh = min(h0 + r, H - 1), gw = min(w0 + x, W - 1), gc = min(c0 + c, C - 1);
            const bool ok = e < N1 && c0 + c < C && h0 + r < H && w0 + x < W;
            const float v = a[gc * HW + (long long)gh * W + gw];
            r1[k] = ok ? v : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < L2; ++k) {
            const int e = tid + NTH * k;
            const int x = e % SW, r = (e / SW) % SH, c = e / (SW * SH);
            const int ghr = h0 - MD + dg * DG + r, gwr = w0 - CF_HALO + x;
            const int gh = min(max(ghr, 0), H - 1), gw = min(max(gwr, 0), W - 1), gc = min(c0 + c, C - 1);
            const bool ok = e < N2 && c0 + c < C && ghr >= 0 && ghr < H && gwr >= 0 && gwr < W;
            const float v = b[gc * HW + (long long)gh * W + gw];
            r2[k] = ok ? v : 0.0f;
        }
    };
    load_chunk(0);
    for (int c0 = 0; c0 < C; c0 += CC) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < L1; ++k) {
            const int e = tid + NTH * k;
            if (e < N1) (&s1[0][0][0])[e] = r1[k];
        }
#pragma unroll
        for (int k = 0; k < L2; ++k) {
            const int e = tid + NTH * k;
            if (e < N2) {
                if constexpr (SWP == SW) (&s2[0][0][0])[e] = r2[k];
                else s2[e / (SW * SH)][(e / SW) % SH][e % SW] = r2[k];
            }
        }
        __syncthreads();
        if (c0 + CC < C) load_chunk(c0 + CC);   // next chunk in flight during the compute
        if (d < P) {
            const int cn = min(CC, C - c0);
            const int br = row + dl;
            for (int cc = 0; cc < cn; ++cc) {
                float va[QX], vb[QX + 2 * MD];
#pragma unroll
                for (int x = 0; x < QX; ++x) va[x] = s1[cc][row][quad * QX + x];
#pragma unroll
                for (int x = 0; x < QX + 2 * MD; ++x) vb[x] = s2[cc][br][quad * QX + CF_HALO - MD + x];
#pragma unroll
                for (int x = 0; x < QX; ++x)
#pragma unroll
                    for (int k = 0; k < P; ++k) acc[x][k] += va[x] * vb[x + k];
            }
        }
    }
    const int h = h0 + row;
    if (d >= P || h >= H) return;
    float* o = out + ((long long)n * P * P + (long long)d * P) * HW + (long long)h * W + w0 + quad * QX;
    // four pixels of a plane row as one 16-byte store when they are in the
    // image and 16-byte aligned (a 16-byte aligned output and W a multiple of
    // 4: every plane row then starts aligned)
    if (QX == 4 && (W & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0 && w0 + quad * QX + 3 < W) {
#pragma unroll
        for (int k = 0; k < P; ++k) {
            float v[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) v[x] = LEGACY ? acc[x % QX][k] / (float)C : acc[x % QX][k];
            __builtin_nontemporal_store(*reinterpret_cast<const f32x4_t*>(v), reinterpret_cast<f32x4_t*>(o + k * HW));
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < P; ++k)
#pragma unroll
        for (int x = 0; x < QX; ++x)
            if (w0 + quad * QX + x < W)
                __builtin_nontemporal_store(LEGACY ? acc[x][k] / (float)C : acc[x][k], o + k * HW + x);
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

template <int MD, int QX, int CC, int DG>
static void launch_corr_f32_q(const float* f1, const float* f2, float* out, int N, int C, int H, int W, bool legacy,
                              hipStream_t s)
{
    constexpr int P = 2 * MD + 1, G = DG >= P ? 1 : P / DG;
    dim3 grid((W + 8 * QX - 1) / (8 * QX), (H + CF_TH - 1) / CF_TH, N * G);
    if (legacy)
        hipLaunchKernelGGL((k_corr_f32<MD, true, QX, CC, DG>), grid, dim3(64 * DG), 0, s, f1, f2, out, C, H, W);
    else
        hipLaunchKernelGGL((k_corr_f32<MD, false, QX, CC, DG>), grid, dim3(64 * DG), 0, s, f1, f2, out, C, H, W);
}

// 4 pixels per thread when that gives at least half a machine of blocks
// (1920x1088 pyramid, with the 16-byte stores: at 240x136, 136 blocks of the
// 4-pixel tile take 29.8 us against 40.2 us for 510 of the 1-pixel tile; at
// 120x68, 36 against 135 blocks, 40 against 24 us)
template <int MD>
static void launch_corr_f32(const float* f1, const float* f2, float* out, int N, int C, int H, int W, bool legacy,
                            hipStream_t s)
{
    const long long blocks4 = (long long)((W + 31) / 32) * ((H + CF_TH - 1) / CF_TH) * N;
    if (blocks4 >= 128) launch_corr_f32_q<MD, 4, 8, 9>(f1, f2, out, N, C, H, W, legacy, s);
    else launch_corr_f32_q<MD, 1, 32, ((2 * MD + 1) % CF_DG1 == 0 ? CF_DG1 : 2 * MD + 1)>(f1, f2, out, N, C, H, W,
                                                                                           legacy, s);
}

// ---------------------------------------------------------------------------
// Warp (warp.cc:71-134 / warp_cuda.cu:28-84).  One thread per output pixel and
// group of 8 channels: the flow is read and the bilinear weights and in-bounds
// mask are computed once per group, then the 8 planes are gathered with
// independent loads (the reference re-reads the flow once per channel,
// grid.y = C, and serialises nothing else).  FP32 weights; the reference evaluates
// weights in double for float tensors, so FP32 agrees to rounding, and the
// keep/zero decision (double)mask > 0.999 becomes mask >= 0.999f, the same
// threshold on the float grid (0.999f is the first float above 0.999).
// ---------------------------------------------------------------------------
static constexpr int WG_CH = 8;

template <typename T>
__global__ __launch_bounds__(256) void k_warp(const T* __restrict__ in, const T* __restrict__ flow, T* __restrict__ out,
                                              int C, int H, int W)
{
    const int x = blockIdx.x * 64 + threadIdx.x % 64;
    const int y = blockIdx.y * 4 + threadIdx.x / 64;
    const int groups = (C + WG_CH - 1) / WG_CH;
    const int n = blockIdx.z / groups;
    const int c0 = (blockIdx.z % groups) * WG_CH;
    const int c1 = min(C, c0 + WG_CH);
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
        for (int c = c0; c < c1; ++c) o[c * HW] = (T)0;
        return;
    }
    // All 4 x WG_CH gathers are issued before any is used: from addresses
    // clamped into the image and into the group's channels, so every load is
    // valid; a corner outside the image is still left out of the sum below,
    // which keeps the reference's terms and order.
    const int iL = min(max((int)xL, 0), W - 1), iR = min(max((int)xR, 0), W - 1);
    const int iT = min(max((int)yT, 0), H - 1), iB = min(max((int)yB, 0), H - 1);
    const long long oTL = (long long)iT * W + iL, oTR = (long long)iT * W + iR;
    const long long oBL = (long long)iB * W + iL, oBR = (long long)iB * W + iR;
    const T* src = in + base;
    T vTL[WG_CH], vTR[WG_CH], vBL[WG_CH], vBR[WG_CH];
#pragma unroll
    for (int j = 0; j < WG_CH; ++j) {
        const T* pl = src + (long long)min(c0 + j, c1 - 1) * HW;
        vTL[j] = pl[oTL];
        vTR[j] = pl[oTR];
        vBL[j] = pl[oBL];
        vBR[j] = pl[oBR];
    }
#pragma unroll
    for (int j = 0; j < WG_CH; ++j) {
        if (c0 + j >= c1) break;
        T v = (T)0;
        if (mT && mL) v = v + wTL * vTL[j];
        if (mT && mR) v = v + wTR * vTR[j];
        if (mB && mL) v = v + wBL * vBL[j];
        if (mB && mR) v = v + wBR * vBR[j];
        o[(long long)(c0 + j) * HW] = v;
    }
}

}  // namespace vsc

using namespace vsc;

namespace {

// workspace: corr_mfma_workspace_bytes() bytes for the reduced-precision
// types (NULL: allocated stream-ordered for the call); unused otherwise
int correlation_forward(const void* f1, const void* f2, void* out, int N, int C, int H, int W, int max_displacement,
                        int legacy, int in_dtype, const float* chan_scale, void* workspace, vsc_hip_stream stream)
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
        return corr_mfma_dispatch(f1, f2, static_cast<float*>(out), N, C, H, W, md, legacy, in_dtype, chan_scale,
                                  workspace, s);
    }
    return set_error(VSC_E_DTYPE, "vsc_correlation_forward: unsupported input type (float, double, bf16, f16, fp8e4m3)");
}

}  // namespace

extern "C" int vsc_correlation_forward(const void* f1, const void* f2, void* out, int N, int C, int H, int W,
                                       int max_displacement, int legacy, int in_dtype, const float* chan_scale,
                                       vsc_hip_stream stream)
{
    return correlation_forward(f1, f2, out, N, C, H, W, max_displacement, legacy, in_dtype, chan_scale, nullptr,
                               stream);
}

// ---------------------------------------------------------------------------
// The Correlation op without any allocation: the channels-last operand copies
// of the reduced-precision types live in a file-scope device array (one copy
// per GPU, held until the process ends) instead of a stream-ordered
// allocation.  Its address on each device is looked up once, with that
// device current; host code uses that address, never the array's name.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[VSC_CORRELATION_STATIC_SCRATCH];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device
int static_scratch(void** out)
{
    int dev = 0;
    if (check_hip(hipGetDevice(&dev), "vsc_correlation_forward_noalloc")) return VSC_E_HIP;
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(VSC_E_UNSUPPORTED, "vsc_correlation_forward_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "vsc_correlation_forward_noalloc scratch"))
            return VSC_E_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return VSC_OK;
}

}  // namespace

extern "C" int vsc_correlation_forward_noalloc(const void* f1, const void* f2, void* out, int N, int C, int H, int W,
                                               int max_displacement, int legacy, int in_dtype, const float* chan_scale,
                                               vsc_hip_stream stream)
{
    const size_t need = (N < 1 || C < 1 || H < 1 || W < 1 || corr_mfma_reads_nchw(f1, f2, W, in_dtype))
                            ? 0
                            : corr_mfma_workspace_bytes(N, C, H, W, in_dtype);
    if (need > VSC_CORRELATION_STATIC_SCRATCH)
        return set_error(VSC_E_INVALID, "vsc_correlation_forward_noalloc: the operands need " + std::to_string(need) +
                                            " bytes of scratch, more than the static " +
                                            std::to_string((size_t)VSC_CORRELATION_STATIC_SCRATCH));
    void* ws = nullptr;
    if (need) {
        int rc = static_scratch(&ws);
        if (rc) return rc;
    }
    return correlation_forward(f1, f2, out, N, C, H, W, max_displacement, legacy, in_dtype, chan_scale, ws, stream);
}

extern "C" size_t vsc_correlation_workspace_size(int N, int C, int H, int W, int in_dtype)
{
    if (N < 1 || C < 1 || H < 1 || W < 1) return 0;
    return corr_mfma_workspace_bytes(N, C, H, W, in_dtype);
}

extern "C" int vsc_correlation_forward_ws(const void* f1, const void* f2, void* out, int N, int C, int H, int W,
                                          int max_displacement, int legacy, int in_dtype, const float* chan_scale,
                                          void* workspace, size_t workspace_bytes, vsc_hip_stream stream)
{
    const size_t need =
        corr_mfma_reads_nchw(f1, f2, W, in_dtype) ? 0 : vsc_correlation_workspace_size(N, C, H, W, in_dtype);
    if (need && (!workspace || workspace_bytes < need))
        return set_error(VSC_E_INVALID, "vsc_correlation_forward_ws: workspace missing or smaller than "
                                        "vsc_correlation_workspace_size()");
    return correlation_forward(f1, f2, out, N, C, H, W, max_displacement, legacy, in_dtype, chan_scale,
                               need ? workspace : nullptr, stream);
}

extern "C" int vsc_warp_forward(const void* in, const void* flow, void* out, int N, int C, int H, int W, int dtype,
                                vsc_hip_stream stream)
{
    if (!in || !flow || !out) return set_error(VSC_E_INVALID, "vsc_warp_forward: null pointer");
    if (N < 1 || C < 1 || H < 1 || W < 1) return set_error(VSC_E_INVALID, "vsc_warp_forward: empty tensor");
    const dim3 grid((W + 63) / 64, (H + 3) / 4, N * ((C + WG_CH - 1) / WG_CH));
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
