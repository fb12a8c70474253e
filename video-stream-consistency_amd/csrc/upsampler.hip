// Learned flow upsampler for the FLOWDOWNSCALE path (the reference upsamples
// the low-resolution flow with get_bilinear only, flowmodel.cpp:158-161):
//   x   = [ bilinear x f of the low-res flow (2 ch), luma of the full-res frame ]
//   a1  = ReLU(conv3x3(x, 3 -> 16)),  a2 = ReLU(conv3x3(a1, 16 -> 16))
//   out = bilinear flow + gate * conv3x3(a2, 16 -> 2)
// Inference runs in FP16 storage on MFMA (FP32 accumulation); training keeps
// FP32 master weights and runs FP32 forward, backward (data and weight
// gradients) and Adam on the device.  The bilinear mapping is the reference's
// kernel_bilinear (flowconsistency.cu:58-74) with the flow vectors multiplied
// by the scale factor.
#include "vsc_internal.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <type_traits>
#include <vector>

namespace vsc {
namespace {

// master parameter layout (FP32, conv-kernel layout [cout][3][3][cin_pad])
constexpr int U_C1 = 4;                                   // layer-1 input channels, padded (x, y, luma, 0)
constexpr int U_W1 = 0, U_B1 = U_W1 + 16 * 9 * U_C1;
constexpr int U_W2 = U_B1 + 16, U_B2 = U_W2 + 16 * 9 * 16;
constexpr int U_W3 = U_B2 + 16, U_B3 = U_W3 + 2 * 9 * 16;
constexpr int U_NP = U_B3 + 2;                            // 3202 floats
constexpr int U_NREF = 16 * 3 * 9 + 16 + 16 * 16 * 9 + 16 + 2 * 16 * 9 + 2;  // torch layout: 3058
// FP16 inference weights of the fused kernel, MFMA A-operand rows [co][K]:
// layer 1 K = tap*4 + c (36 used of 64), layers 2 and 3 K = tap*16 + ci (144
// used of 160); layer 3 rows 2..15 are zero
constexpr int H_K1 = 64, H_K2 = 160;
constexpr int H_W1 = 0, H_W2 = 16 * H_K1, H_W3 = H_W2 + 16 * H_K2, H_N = H_W3 + 16 * H_K2;
// flipped / transposed FP32 weights for the data gradients
constexpr int T_W2 = 0, T_W3 = 16 * 9 * 16, T_N = T_W3 + 16 * 9 * 4;

}  // namespace
}  // namespace vsc

struct vsc_upsampler {
    float* p = nullptr;       // U_NP master params
    float* g = nullptr;       // gradients
    float* m = nullptr;       // Adam moments
    float* v = nullptr;
    uint16_t* w16 = nullptr;  // FP16 inference weights
    float* wT = nullptr;      // transposed weights for backward data
    int step = 0;
    // training workspace (grown on demand)
    size_t ws_px = 0, ws_part = 0;
    float *x = nullptr, *base = nullptr, *a1 = nullptr, *a2 = nullptr, *out = nullptr;
    float *dout = nullptr, *d2 = nullptr, *d1 = nullptr, *part = nullptr, *loss = nullptr;
};

namespace vsc {
namespace {

// ---------------------------------------------------------------------------
// input: bilinear x scale of the low-res flow, luma; FP32 residual base
// ---------------------------------------------------------------------------
template <int EK>
__global__ void k_ups_prep(const float* __restrict__ lo, int lw, int lh, int lch, long long lo_stride,
                           const uint8_t* __restrict__ rgba, int W, int H, float scale, void* xin, int xcs,
                           float* base, float* out3)
{
    using E = Elem<EK>;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, n = blockIdx.z;
    if (x >= W || y >= H) return;
    const BiTap t = bilinear_tap(x, y, lw, lh, W, H);
    View fl{const_cast<float*>(lo) + n * lo_stride, lw, lh, lch, 1, (long long)lch};
    const float u = bilinear_sample(fl, t, 0) * scale, v = bilinear_sample(fl, t, 1) * scale;
    const long long p = ((long long)n * H + y) * W + x;
    const uint8_t* px = rgba + p * 4;
    const float luma = (0.299f * (float)px[0] + 0.587f * (float)px[1] + 0.114f * (float)px[2]) * (1.0f / 255.0f);
    typename E::T* xi = static_cast<typename E::T*>(xin) + p * xcs;
    xi[0] = E::st(u);
    xi[1] = E::st(v);
    xi[2] = E::st(luma);
    for (int c = 3; c < xcs; ++c) xi[c] = E::st(0.0f);
    base[p * 2] = u;
    base[p * 2 + 1] = v;
    if (out3) out3[p * 3 + 2] = 0.0f;
}

// Charbonnier end-point loss: rho = sqrt(|out - gt|^2 + eps^2), loss = mean rho;
// dOut = (out - gt) / rho / count (4-channel rows, channels 2 and 3 zero).
__global__ void k_charb(const float* __restrict__ out, const float* __restrict__ gt, float* dout, float* part,
                        long long npx, float eps, float inv)
{
    __shared__ float red[256];
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (i < npx) {
        const float du = out[i * 2] - gt[i * 2], dv = out[i * 2 + 1] - gt[i * 2 + 1];
        const float rho = sqrtf(du * du + dv * dv + eps * eps);
        l = rho;
        dout[i * 4] = du / rho * inv;
        dout[i * 4 + 1] = dv / rho * inv;
        dout[i * 4 + 2] = 0.0f;
        dout[i * 4 + 3] = 0.0f;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_relu_bw(float* d, const float* a, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n && !(a[i] > 0.0f)) d[i] = 0.0f;
}

// weight and bias gradients of a 3x3 conv (stride 1, pad 1) over a 16 x 16
// pixel tile: part[block][co][tap][ci] = sum_p dY[p][co] * X[p + tap][ci],
// part[block][E + co] = sum_p dY[p][co].  X: [N][H][W][CI], dY: [N][H][W][DS].
template <int CI, int CO, int DS>
__global__ __launch_bounds__(256) void k_conv_bw_w(const float* __restrict__ X, const float* __restrict__ dY, int H,
                                                   int W, int tiles_x, int tiles_y, float* part)
{
    constexpr int E = CO * 9 * CI;
    __shared__ float sx[18 * 18 * CI];
    __shared__ float sd[256 * CO];
    const int tid = threadIdx.x, n = blockIdx.z;
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int y0 = ty * 16, x0 = tx * 16;
    for (int e = tid; e < 18 * 18 * CI; e += 256) {
        const int c = e % CI, p = e / CI, iy = y0 - 1 + p / 18, ix = x0 - 1 + p % 18;
        sx[e] = (iy >= 0 && iy < H && ix >= 0 && ix < W) ? X[(((long long)n * H + iy) * W + ix) * CI + c] : 0.0f;
    }
    for (int e = tid; e < 256 * CO; e += 256) {
        const int c = e % CO, p = e / CO, iy = y0 + p / 16, ix = x0 + p % 16;
        sd[e] = (iy < H && ix < W) ? dY[(((long long)n * H + iy) * W + ix) * DS + c] : 0.0f;
    }
    __syncthreads();
    float* out = part + ((long long)n * tiles_y * tiles_x + blockIdx.x) * (E + CO);
    for (int e = tid; e < E + CO; e += 256) {
        float s = 0.0f;
        if (e < E) {
            const int ci = e % CI, tap = (e / CI) % 9, co = e / (9 * CI), ky = tap / 3, kx = tap % 3;
            for (int p = 0; p < 256; ++p) s += sd[p * CO + co] * sx[((p / 16 + ky) * 18 + p % 16 + kx) * CI + ci];
        } else {
            for (int p = 0; p < 256; ++p) s += sd[p * CO + (e - E)];
        }
        out[e] = s;
    }
}

// out[e] = sum over rows b of part[b][e], in row order (deterministic)
__global__ void k_reduce_rows(const float* __restrict__ part, int rows, int cols, float* out)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= cols) return;
    float s = 0.0f;
    for (int b = 0; b < rows; ++b) s += part[(long long)b * cols + e];
    out[e] = s;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// FP16 inference weights and transposed FP32 weights from the master params
__global__ void k_ups_derive(const float* __restrict__ p, uint16_t* w16, float* wT)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < H_N) {
        float v = 0.0f;
        if (i < H_W2) {
            const int co = i / H_K1, k = i % H_K1, tap = k / 4, c = k % 4;
            v = (tap < 9 && c < 3) ? p[U_W1 + (co * 9 + tap) * U_C1 + c] : 0.0f;
        } else if (i < H_W3) {
            const int co = (i - H_W2) / H_K2, k = (i - H_W2) % H_K2;
            v = k < 144 ? p[U_W2 + co * 144 + k] : 0.0f;
        } else {
            const int co = (i - H_W3) / H_K2, k = (i - H_W3) % H_K2;
            v = (co < 2 && k < 144) ? p[U_W3 + co * 144 + k] : 0.0f;
        }
        w16[i] = f32_to_f16(v);
    }
    if (i < T_N) {
        float v = 0.0f;
        if (i < T_W3) {   // wT2[ci][t][co] = W2[co][8 - t][ci]
            const int co = i % 16, t = (i / 16) % 9, ci = i / 144;
            v = p[U_W2 + (co * 9 + 8 - t) * 16 + ci];
        } else {          // wT3[ci][t][co4] = W3[co][8 - t][ci], co < 2
            const int j = i - T_W3, co = j % 4, t = (j / 4) % 9, ci = j / 36;
            v = co < 2 ? p[U_W3 + (co * 9 + 8 - t) * 16 + ci] : 0.0f;
        }
        wT[i] = v;
    }
}

// ---------------------------------------------------------------------------
// Fused FP16 inference: the three convolutions of one 16 x 32 output tile in
// one block, activations in LDS.  The input tile carries a 3-pixel halo, the
// first activation a 2-pixel and the second a 1-pixel halo; activations of
// pixels outside the frame are zero (each layer's zero padding).  Each layer
// is an MFMA GEMM D[co][px] = W[co][K] * X[K][px] over 16-pixel groups, K =
// (tap, channel) pairs gathered from LDS: layer 1 in two 16x16x32 steps (taps
// 0..7, tap 8), layers 2 and 3 in five (tap pairs).  Weights stay in
// registers for the whole block.
// ---------------------------------------------------------------------------
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 up_f16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float up_f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t up_u32x4;

constexpr int UF_TH = 16, UF_TW = 32;
constexpr int UF_XH = UF_TH + 6, UF_XW = UF_TW + 6;      // input tile
constexpr int UF_1H = UF_TH + 4, UF_1W = UF_TW + 4;      // layer-1 output tile
constexpr int UF_2H = UF_TH + 2, UF_2W = UF_TW + 2;      // layer-2 output tile
constexpr int UF_AS = 24;                                 // LDS pixel stride of activations (halfs): 48 B

__device__ inline up_f32x4 up_mma(const up_u32x4& a, const up_u32x4& b, up_f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(up_f16x8, a), __builtin_bit_cast(up_f16x8, b), c,
                                                  0, 0, 0);
}

__device__ inline uint32_t pack_h2(float a, float b)
{
    return (uint32_t)f32_to_f16(a) | ((uint32_t)f32_to_f16(b) << 16);
}

// one 16-pixel group of a 16-channel layer: K = 144 (9 taps x 16) in five
// steps; off[m] is the lane's LDS offset of tap 2m + (lg >> 1) relative to
// the group pixel's top-left tap (negative: the tap does not exist)
__device__ inline up_f32x4 ups_l16(const uint16_t* src, const up_u32x4 (&w)[5], int base, const int (&off)[5])
{
    up_f32x4 acc = up_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    up_u32x4 xb[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        xb[m] = up_u32x4{0u, 0u, 0u, 0u};
        if (off[m] >= 0) xb[m] = *reinterpret_cast<const up_u32x4*>(&src[base + off[m]]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) acc = up_mma(w[m], xb[m], acc);
    return acc;
}

// blockIdx.y selects the flow (up to two flows of one frame per launch)
struct UpsFlows { const float* lo[2]; float* out[2]; };

__global__ __launch_bounds__(512) void k_ups_fused(UpsFlows flows, int lw, int lh, int lch,
                                                   const uint8_t* __restrict__ rgba, int W, int H, float scale,
                                                   const uint16_t* __restrict__ wf, const float* __restrict__ prm,
                                                   const float* gate, int out_ch, int tiles_x)
{
    const float* __restrict__ lo = flows.lo[blockIdx.y];
    float* __restrict__ out = flows.out[blockIdx.y];
    constexpr int NW = 8;   // waves per block; each keeps two 16-pixel groups in flight
    __shared__ __attribute__((aligned(16))) uint2 sx[UF_XH * UF_XW];
    __shared__ __attribute__((aligned(16))) uint16_t s1[UF_1H * UF_1W * UF_AS];
    __shared__ __attribute__((aligned(16))) uint16_t s2[UF_2H * UF_2W * UF_AS];
    __shared__ float2 sbase[UF_TH * UF_TW];          // FP32 bilinear flow of the output tile
    __shared__ int tx0[UF_XW], tx1[UF_XW], ty0[UF_XH], ty1[UF_XH];
    __shared__ float tfx[UF_XW], tfy[UF_XH];
    const int tid = threadIdx.x, q = tid / 64, lane = tid % 64, l16 = lane % 16, lg = lane / 16;
    const int y0 = (blockIdx.x / tiles_x) * UF_TH, x0 = (blockIdx.x % tiles_x) * UF_TW;
    // bilinear taps of the tile's columns and rows (kernel_bilinear mapping)
    if (tid < UF_XW) {
        const int ix = min(max(x0 - 3 + tid, 0), W - 1);
        const BiTap t = bilinear_tap(ix, 0, lw, lh, W, H);
        tx0[tid] = t.ix; tx1[tid] = t.ix1; tfx[tid] = t.fx;
    } else if (tid >= 64 && tid < 64 + UF_XH) {
        const int iy = min(max(y0 - 3 + tid - 64, 0), H - 1);
        const BiTap t = bilinear_tap(0, iy, lw, lh, W, H);
        ty0[tid - 64] = t.iy; ty1[tid - 64] = t.iy1; tfy[tid - 64] = t.fy;
    }

    // weights (A operand rows co = l16), biases for this lane's 4 output rows,
    // and the lane's tap offsets in the activation tiles
    up_u32x4 w1[2], w2[5], w3[5];
#pragma unroll
    for (int m = 0; m < 2; ++m) w1[m] = *reinterpret_cast<const up_u32x4*>(wf + H_W1 + l16 * H_K1 + 32 * m + 8 * lg);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        w2[m] = *reinterpret_cast<const up_u32x4*>(wf + H_W2 + l16 * H_K2 + 32 * m + 8 * lg);
        w3[m] = *reinterpret_cast<const up_u32x4*>(wf + H_W3 + l16 * H_K2 + 32 * m + 8 * lg);
    }
    float b1[4], b2[4], b3[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        b1[r] = prm[U_B1 + 4 * lg + r];
        b2[r] = prm[U_B2 + 4 * lg + r];
        b3[r] = 4 * lg + r < 2 ? prm[U_B3 + 4 * lg + r] : 0.0f;
    }
    int off1[5], off2[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const int tap = 2 * m + (lg >> 1);
        off1[m] = tap < 9 ? ((tap / 3) * UF_1W + tap % 3) * UF_AS + 8 * (lg & 1) : -1;
        off2[m] = tap < 9 ? ((tap / 3) * UF_2W + tap % 3) * UF_AS + 8 * (lg & 1) : -1;
    }
    const int t0 = 2 * lg, t1 = 2 * lg + 1;
    const int offx0 = (t0 / 3) * UF_XW + t0 % 3, offx1 = (t1 / 3) * UF_XW + t1 % 3, offx8 = 2 * UF_XW + 2;
    __syncthreads();

    // Tiles whose input tile (with its 3-pixel halo) lies inside the frame
    // skip every per-pixel frame test (block-uniform choice).
    const bool interior = y0 >= 3 && y0 + UF_TH + 3 <= H && x0 >= 3 && x0 + UF_TW + 3 <= W;
    auto run = [&](auto interior_c) {
        constexpr bool INTERIOR = decltype(interior_c)::value;
        // input tile: (u, v) = bilinear x scale, luma; zero outside the frame
        for (int e = tid; e < UF_XH * UF_XW; e += 64 * NW) {
            const int r = e / UF_XW, c = e % UF_XW;
            const int iy = y0 - 3 + r, ix = x0 - 3 + c;
            uint2 v = make_uint2(0u, 0u);
            if (INTERIOR || (iy >= 0 && iy < H && ix >= 0 && ix < W)) {
                // bilinear_sample's operation order, both channels from 8-byte loads
                const int a0 = tx0[c], a1 = tx1[c], r0 = ty0[r] * lw, r1 = ty1[r] * lw;
                const float fx = tfx[c], fy = tfy[r];
                float2 v00, v10, v01, v11;
                if (lch == 2) {
                    const float2* f2 = reinterpret_cast<const float2*>(lo);
                    v00 = f2[r0 + a0]; v10 = f2[r0 + a1]; v01 = f2[r1 + a0]; v11 = f2[r1 + a1];
                } else {
                    v00 = make_float2(lo[(r0 + a0) * lch], lo[(r0 + a0) * lch + 1]);
                    v10 = make_float2(lo[(r0 + a1) * lch], lo[(r0 + a1) * lch + 1]);
                    v01 = make_float2(lo[(r1 + a0) * lch], lo[(r1 + a0) * lch + 1]);
                    v11 = make_float2(lo[(r1 + a1) * lch], lo[(r1 + a1) * lch + 1]);
                }
                const float gx = 1.0f - fx, gy = 1.0f - fy;
                const float u = ((((v00.x * gx) * gy + (v10.x * fx) * gy) + (v01.x * gx) * fy) + (v11.x * fx) * fy) * scale;
                const float w = ((((v00.y * gx) * gy + (v10.y * fx) * gy) + (v01.y * gx) * fy) + (v11.y * fx) * fy) * scale;
                const uint8_t* px = rgba + ((long long)iy * W + ix) * 4;
                const float luma = (0.299f * (float)px[0] + 0.587f * (float)px[1] + 0.114f * (float)px[2]) * (1.0f / 255.0f);
                v = make_uint2(pack_h2(u, w), pack_h2(luma, 0.0f));
                if (r >= 3 && r < 3 + UF_TH && c >= 3 && c < 3 + UF_TW) sbase[(r - 3) * UF_TW + c - 3] = make_float2(u, w);
            }
            sx[e] = v;
        }
        __syncthreads();

        // layer 1 on the 20 x 36 tile, ReLU, zero outside the frame
        constexpr int N1 = UF_1H * UF_1W, G1 = (N1 + 15) / 16;
        auto l1 = [&](int p) {
            const int bx = (p / UF_1W) * UF_XW + p % UF_1W;
            const uint2 a = sx[bx + offx0], b = sx[bx + offx1];
            const uint2 e = lg == 0 ? sx[bx + offx8] : make_uint2(0u, 0u);
            up_f32x4 acc = up_mma(w1[0], up_u32x4{a.x, a.y, b.x, b.y}, up_f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            return up_mma(w1[1], up_u32x4{e.x, e.y, 0u, 0u}, acc);
        };
        auto put1 = [&](int g, const up_f32x4& acc) {
            const int p = 16 * g + l16;
            if (p >= N1) return;
            const int r = p / UF_1W, c = p % UF_1W, iy = y0 - 2 + r, ix = x0 - 2 + c;
            const bool in = INTERIOR || (iy >= 0 && iy < H && ix >= 0 && ix < W);
            float o[4];
    #pragma unroll
            for (int k = 0; k < 4; ++k) o[k] = in ? fmaxf(acc[k] + b1[k], 0.0f) : 0.0f;
            *reinterpret_cast<uint2*>(&s1[p * UF_AS + 4 * lg]) = make_uint2(pack_h2(o[0], o[1]), pack_h2(o[2], o[3]));
        };
        for (int g = q; g < G1; g += 2 * NW) {
            const int g2 = g + NW;
            const up_f32x4 a0 = l1(min(16 * g + l16, N1 - 1));
            const up_f32x4 a1 = l1(min(16 * g2 + l16, N1 - 1));
            put1(g, a0);
            if (g2 < G1) put1(g2, a1);
        }
        __syncthreads();

        // layer 2 on the 18 x 34 tile
        constexpr int N2 = UF_2H * UF_2W, G2 = (N2 + 15) / 16;
        auto put2 = [&](int g, const up_f32x4& acc) {
            const int p = 16 * g + l16;
            if (p >= N2) return;
            const int r = p / UF_2W, c = p % UF_2W, iy = y0 - 1 + r, ix = x0 - 1 + c;
            const bool in = INTERIOR || (iy >= 0 && iy < H && ix >= 0 && ix < W);
            float o[4];
    #pragma unroll
            for (int k = 0; k < 4; ++k) o[k] = in ? fmaxf(acc[k] + b2[k], 0.0f) : 0.0f;
            *reinterpret_cast<uint2*>(&s2[p * UF_AS + 4 * lg]) = make_uint2(pack_h2(o[0], o[1]), pack_h2(o[2], o[3]));
        };
        auto base2 = [&](int p) { return ((p / UF_2W) * UF_1W + p % UF_2W) * UF_AS; };
        for (int g = q; g < G2; g += 2 * NW) {
            const int g2 = g + NW;
            const up_f32x4 a0 = ups_l16(s1, w2, base2(min(16 * g + l16, N2 - 1)), off1);
            const up_f32x4 a1 = ups_l16(s1, w2, base2(min(16 * g2 + l16, N2 - 1)), off1);
            put2(g, a0);
            if (g2 < G2) put2(g2, a1);
        }
        __syncthreads();

        // layer 3 on the 16 x 32 output tile, added to the FP32 bilinear flow
        const float gv = gate ? *gate : 1.0f;
        constexpr int G3 = UF_TH * UF_TW / 16;
        auto put3 = [&](int g, const up_f32x4& acc) {
            const int p = 16 * g + l16, r = p / UF_TW, c = p % UF_TW, iy = y0 + r, ix = x0 + c;
            if (lg == 0 && (INTERIOR || (iy < H && ix < W))) {
                const float2 bs = sbase[p];
                float* o = out + ((long long)iy * W + ix) * out_ch;
                o[0] = bs.x + gv * (acc[0] + b3[0]);
                o[1] = bs.y + gv * (acc[1] + b3[1]);
                if (out_ch == 3) o[2] = 0.0f;
            }
        };
        auto base3 = [&](int p) { return ((p / UF_TW) * UF_2W + p % UF_TW) * UF_AS; };
        for (int g = q; g < G3; g += 2 * NW) {
            const int g2 = g + NW;
            const up_f32x4 a0 = ups_l16(s2, w3, base3(16 * g + l16), off2);
            const up_f32x4 a1 = ups_l16(s2, w3, base3(16 * min(g2, G3 - 1) + l16), off2);
            put3(g, a0);
            if (g2 < G3) put3(g2, a1);
        }
    };
    if (interior) run(std::true_type{});
    else run(std::false_type{});
}

// ---------------------------------------------------------------------------
// Forward-backward flow consistency (Sundaram, Brox & Keutzer 2010): f maps
// frame A to B, b maps B to A.  Pixel x is inconsistent when x + f(x) leaves
// the frame or |f(x) + b(x + f(x))|^2 >= 0.01 (|f(x)|^2 + |b(x + f(x))|^2) + 0.5,
// with b sampled bilinearly.  One count per frame; the gate is 1 (use the
// learned upsampler) when the inconsistent fraction is at most max_frac, else
// 0 (bilinear fallback for this frame).
// ---------------------------------------------------------------------------
__global__ void k_fb_zero(uint32_t* c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = 0u;
}

__global__ void k_fb_count(const float* __restrict__ f, const float* __restrict__ b, int w, int h, int ch,
                           uint32_t* count)
{
    uint32_t n = 0;
    const long long npx = (long long)w * h;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < npx; i += (long long)gridDim.x * 256) {
        bool bad;
        const int x = (int)(i % w), y = (int)(i / w);
        const float fx = f[i * ch], fy = f[i * ch + 1];
        const float tx = (float)x + fx, ty = (float)y + fy;
        if (!(tx >= 0.0f && ty >= 0.0f && tx <= (float)(w - 1) && ty <= (float)(h - 1))) {
            bad = true;
        } else {
            const int x0 = min((int)tx, w - 1), y0 = min((int)ty, h - 1);
            const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
            const float ax = tx - (float)x0, ay = ty - (float)y0;
            float bv[2];
            for (int c = 0; c < 2; ++c) {
                const float v00 = b[((long long)y0 * w + x0) * ch + c], v10 = b[((long long)y0 * w + x1) * ch + c];
                const float v01 = b[((long long)y1 * w + x0) * ch + c], v11 = b[((long long)y1 * w + x1) * ch + c];
                bv[c] = (1.0f - ay) * ((1.0f - ax) * v00 + ax * v10) + ay * ((1.0f - ax) * v01 + ax * v11);
            }
            const float sx = fx + bv[0], sy = fy + bv[1];
            const float lhs = sx * sx + sy * sy;
            const float rhs = 0.01f * (fx * fx + fy * fy + bv[0] * bv[0] + bv[1] * bv[1]) + 0.5f;
            bad = !(lhs < rhs);
        }
        n += bad ? 1u : 0u;
    }
    // block sum, one atomic per block (few hundred per frame)
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    __shared__ uint32_t wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x / 64] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (t) atomicAdd(count, t);
    }
}

__global__ void k_fb_gate(const uint32_t* count, long long npx, float max_frac, float* gate)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) gate[0] = ((double)count[0] <= (double)max_frac * (double)npx) ? 1.0f : 0.0f;
}

int derive(vsc_upsampler* u, hipStream_t s)
{
    hipLaunchKernelGGL(k_ups_derive, dim3((H_N + 255) / 256), dim3(256), 0, s, u->p, u->w16, u->wT);
    return check_launch("upsampler derive");
}

// reference (torch) layout <-> master layout
void ref_to_master(const float* r, std::vector<float>& m)
{
    m.assign(U_NP, 0.0f);
    const float* q = r;
    for (int co = 0; co < 16; ++co)
        for (int ci = 0; ci < 3; ++ci)
            for (int t = 0; t < 9; ++t) m[U_W1 + (co * 9 + t) * U_C1 + ci] = q[(co * 3 + ci) * 9 + t];
    q += 16 * 3 * 9;
    for (int co = 0; co < 16; ++co) m[U_B1 + co] = q[co];
    q += 16;
    for (int co = 0; co < 16; ++co)
        for (int ci = 0; ci < 16; ++ci)
            for (int t = 0; t < 9; ++t) m[U_W2 + (co * 9 + t) * 16 + ci] = q[(co * 16 + ci) * 9 + t];
    q += 16 * 16 * 9;
    for (int co = 0; co < 16; ++co) m[U_B2 + co] = q[co];
    q += 16;
    for (int co = 0; co < 2; ++co)
        for (int ci = 0; ci < 16; ++ci)
            for (int t = 0; t < 9; ++t) m[U_W3 + (co * 9 + t) * 16 + ci] = q[(co * 16 + ci) * 9 + t];
    q += 2 * 16 * 9;
    for (int co = 0; co < 2; ++co) m[U_B3 + co] = q[co];
}

void master_to_ref(const std::vector<float>& m, float* r)
{
    float* q = r;
    for (int co = 0; co < 16; ++co)
        for (int ci = 0; ci < 3; ++ci)
            for (int t = 0; t < 9; ++t) q[(co * 3 + ci) * 9 + t] = m[U_W1 + (co * 9 + t) * U_C1 + ci];
    q += 16 * 3 * 9;
    for (int co = 0; co < 16; ++co) q[co] = m[U_B1 + co];
    q += 16;
    for (int co = 0; co < 16; ++co)
        for (int ci = 0; ci < 16; ++ci)
            for (int t = 0; t < 9; ++t) q[(co * 16 + ci) * 9 + t] = m[U_W2 + (co * 9 + t) * 16 + ci];
    q += 16 * 16 * 9;
    for (int co = 0; co < 16; ++co) q[co] = m[U_B2 + co];
    q += 16;
    for (int co = 0; co < 2; ++co)
        for (int ci = 0; ci < 16; ++ci)
            for (int t = 0; t < 9; ++t) q[(co * 16 + ci) * 9 + t] = m[U_W3 + (co * 9 + t) * 16 + ci];
    q += 2 * 16 * 9;
    for (int co = 0; co < 2; ++co) q[co] = m[U_B3 + co];
}

int conv(const void* in, int cin, const void* wt, const float* bias, void* out, int out_cs, int cout, int act,
         const void* add, int add_cs, const float* gate, int ek, int eko, int N, int H, int W, hipStream_t s)
{
    ConvDesc d{};
    d.in = Nhwc{const_cast<void*>(in), N, H, W, cin, cin, 0};
    d.out = Nhwc{out, N, H, W, cout, out_cs, 0};
    d.add = Nhwc{const_cast<void*>(add), N, H, W, cout, add_cs, 0};
    d.wt = wt; d.bias = bias; d.gate = gate;
    d.stride = 1; d.dil = 1; d.pad = 1; d.act = act;
    d.ek_in = ek; d.ek_out = eko;
    return launch_conv3x3(d, s);
}

int grow(vsc_upsampler* u, size_t px, size_t part)
{
    if (px <= u->ws_px && part <= u->ws_part) return VSC_OK;
    float** bufs[] = {&u->x, &u->base, &u->a1, &u->a2, &u->out, &u->dout, &u->d2, &u->d1, &u->part, &u->loss};
    if (hipDeviceSynchronize() != hipSuccess) return set_error(VSC_E_HIP, "upsampler: sync");
    for (float** b : bufs) { (void)hipFree(*b); *b = nullptr; }
    const size_t sz[] = {px * 4, px * 2, px * 16, px * 16, px * 2, px * 4, px * 16, px * 16, part, px / 256 + 2};
    for (int i = 0; i < 10; ++i)
        if (hipMalloc(reinterpret_cast<void**>(bufs[i]), sz[i] * 4) != hipSuccess) {
            u->ws_px = u->ws_part = 0;
            return set_error(VSC_E_NOMEM, "upsampler: workspace allocation failed");
        }
    u->ws_px = px;
    u->ws_part = part;
    return VSC_OK;
}

// forward (FP32) and backward into u->g; optional loss (host, synchronises)
int fwd_bwd(vsc_upsampler* u, int B, const float* lo, int lw, int lh, const uint8_t* rgba, const float* gt, int W,
            int H, float scale, float* loss_host, hipStream_t s)
{
    const size_t px = (size_t)B * W * H;
    const int tx = (W + 15) / 16, ty = (H + 15) / 16;
    const size_t part = (size_t)B * tx * ty * (16 * 9 * 16 + 16);
    int rc = grow(u, px, part);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ups_prep<EK_F32>, dim3((W + 255) / 256, H, B), dim3(256), 0, s, lo, lw, lh, 2,
                       (long long)lw * lh * 2, rgba, W, H, scale, u->x, U_C1, u->base, nullptr);
    if ((rc = check_launch("upsampler prep"))) return rc;
    // forward
    if ((rc = conv(u->x, U_C1, u->p + U_W1, u->p + U_B1, u->a1, 16, 16, 2, nullptr, 0, nullptr, EK_F32, EK_F32, B, H, W, s))) return rc;
    if ((rc = conv(u->a1, 16, u->p + U_W2, u->p + U_B2, u->a2, 16, 16, 2, nullptr, 0, nullptr, EK_F32, EK_F32, B, H, W, s))) return rc;
    if ((rc = conv(u->a2, 16, u->p + U_W3, u->p + U_B3, u->out, 2, 2, 0, u->base, 2, nullptr, EK_F32, EK_F32, B, H, W, s))) return rc;
    // loss and dOut
    const unsigned lb = (unsigned)((px + 255) / 256);
    hipLaunchKernelGGL(k_charb, dim3(lb), dim3(256), 0, s, u->out, gt, u->dout, u->loss, (long long)px, 1e-3f,
                       1.0f / (float)px);
    if ((rc = check_launch("upsampler loss"))) return rc;
    const int nb = B * tx * ty;
    dim3 gw(tx * ty, 1, B);
    // layer 3
    hipLaunchKernelGGL((k_conv_bw_w<16, 2, 4>), gw, dim3(256), 0, s, u->a2, u->dout, H, W, tx, ty, u->part);
    hipLaunchKernelGGL(k_reduce_rows, dim3((2 * 9 * 16 + 2 + 255) / 256), dim3(256), 0, s, u->part, nb, 2 * 9 * 16 + 2,
                       u->g + U_W3);
    if ((rc = conv(u->dout, 4, u->wT + T_W3, nullptr, u->d2, 16, 16, 0, nullptr, 0, nullptr, EK_F32, EK_F32, B, H, W, s))) return rc;
    hipLaunchKernelGGL(k_relu_bw, dim3((unsigned)((px * 16 + 255) / 256)), dim3(256), 0, s, u->d2, u->a2, (long long)px * 16);
    // layer 2
    hipLaunchKernelGGL((k_conv_bw_w<16, 16, 16>), gw, dim3(256), 0, s, u->a1, u->d2, H, W, tx, ty, u->part);
    hipLaunchKernelGGL(k_reduce_rows, dim3((16 * 9 * 16 + 16 + 255) / 256), dim3(256), 0, s, u->part, nb,
                       16 * 9 * 16 + 16, u->g + U_W2);
    if ((rc = conv(u->d2, 16, u->wT + T_W2, nullptr, u->d1, 16, 16, 0, nullptr, 0, nullptr, EK_F32, EK_F32, B, H, W, s))) return rc;
    hipLaunchKernelGGL(k_relu_bw, dim3((unsigned)((px * 16 + 255) / 256)), dim3(256), 0, s, u->d1, u->a1, (long long)px * 16);
    // layer 1
    hipLaunchKernelGGL((k_conv_bw_w<U_C1, 16, 16>), gw, dim3(256), 0, s, u->x, u->d1, H, W, tx, ty, u->part);
    hipLaunchKernelGGL(k_reduce_rows, dim3((16 * 9 * U_C1 + 16 + 255) / 256), dim3(256), 0, s, u->part, nb,
                       16 * 9 * U_C1 + 16, u->g + U_W1);
    if ((rc = check_launch("upsampler backward"))) return rc;
    if (loss_host) {
        std::vector<float> parts(lb);
        if (check_hip(hipMemcpyAsync(parts.data(), u->loss, lb * 4, hipMemcpyDeviceToHost, s), "upsampler loss") ||
            check_hip(hipStreamSynchronize(s), "upsampler loss"))
            return VSC_E_HIP;
        double t = 0.0;
        for (float x : parts) t += x;
        *loss_host = (float)(t / (double)px);
    }
    return VSC_OK;
}

int check_train_args(const vsc_upsampler* u, int B, const float* lo, int lw, int lh, const uint8_t* rgba,
                     const float* gt, int W, int H)
{
    if (!u || !lo || !rgba || !gt) return set_error(VSC_E_INVALID, "vsc_upsampler: NULL argument");
    if (B <= 0 || W <= 0 || H <= 0 || lw <= 0 || lh <= 0 || (size_t)B * W * H > (1ull << 28))
        return set_error(VSC_E_INVALID, "vsc_upsampler: bad batch shape");
    return VSC_OK;
}

}  // namespace
}  // namespace vsc

using namespace vsc;

extern "C" size_t vsc_upsampler_param_count(void) { return U_NREF; }

extern "C" int vsc_upsampler_create(unsigned long long seed, vsc_upsampler** out)
{
    if (!out) return set_error(VSC_E_INVALID, "vsc_upsampler_create: out is NULL");
    *out = nullptr;
    vsc_upsampler* u = new (std::nothrow) vsc_upsampler();
    if (!u) return set_error(VSC_E_NOMEM, "vsc_upsampler_create: host memory");
    void** bufs[] = {(void**)&u->p, (void**)&u->g, (void**)&u->m, (void**)&u->v, (void**)&u->w16, (void**)&u->wT};
    const size_t sz[] = {U_NP * 4, U_NP * 4, U_NP * 4, U_NP * 4, H_N * 2, T_N * 4};
    for (int i = 0; i < 6; ++i)
        if (hipMalloc(bufs[i], sz[i]) != hipSuccess || hipMemset(*bufs[i], 0, sz[i]) != hipSuccess) {
            vsc_upsampler_destroy(u);
            return set_error(VSC_E_NOMEM, "vsc_upsampler_create: device allocation failed");
        }
    // Kaiming normal (fan_in, ReLU gain sqrt 2), zero biases; the last layer
    // starts at zero so the untrained network is the bilinear upsampler
    std::vector<float> ref(U_NREF, 0.0f);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 77;
    auto nxt = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return ((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0) + 0.5 / 9007199254740992.0;
    };
    auto normal = [&]() { const double a = nxt(), b = nxt(); return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b); };
    for (int i = 0; i < 16 * 3 * 9; ++i) ref[i] = (float)(std::sqrt(2.0 / 27.0) * normal());
    for (int i = 0; i < 16 * 16 * 9; ++i) ref[16 * 3 * 9 + 16 + i] = (float)(std::sqrt(2.0 / 144.0) * normal());
    int rc = vsc_upsampler_set_params(u, ref.data());
    if (rc) { vsc_upsampler_destroy(u); return rc; }
    *out = u;
    return VSC_OK;
}

extern "C" int vsc_upsampler_destroy(vsc_upsampler* u)
{
    if (!u) return VSC_OK;
    void* bufs[] = {u->p, u->g, u->m, u->v, u->w16, u->wT, u->x, u->base, u->a1, u->a2, u->out, u->dout, u->d2, u->d1, u->part, u->loss};
    for (void* b : bufs) (void)hipFree(b);
    delete u;
    return VSC_OK;
}

extern "C" int vsc_upsampler_get_params(const vsc_upsampler* u, float* host)
{
    if (!u || !host) return set_error(VSC_E_INVALID, "vsc_upsampler_get_params: NULL argument");
    std::vector<float> m(U_NP);
    if (check_hip(hipDeviceSynchronize(), "vsc_upsampler_get_params") ||
        check_hip(hipMemcpy(m.data(), u->p, U_NP * 4, hipMemcpyDeviceToHost), "vsc_upsampler_get_params"))
        return VSC_E_HIP;
    master_to_ref(m, host);
    return VSC_OK;
}

extern "C" int vsc_upsampler_set_params(vsc_upsampler* u, const float* host)
{
    if (!u || !host) return set_error(VSC_E_INVALID, "vsc_upsampler_set_params: NULL argument");
    std::vector<float> m;
    ref_to_master(host, m);
    if (check_hip(hipDeviceSynchronize(), "vsc_upsampler_set_params") ||
        check_hip(hipMemcpy(u->p, m.data(), U_NP * 4, hipMemcpyHostToDevice), "vsc_upsampler_set_params") ||
        check_hip(hipMemset(u->m, 0, U_NP * 4), "vsc_upsampler_set_params") ||
        check_hip(hipMemset(u->v, 0, U_NP * 4), "vsc_upsampler_set_params"))
        return VSC_E_HIP;
    u->step = 0;
    int rc = derive(u, nullptr);
    if (rc) return rc;
    return check_hip(hipDeviceSynchronize(), "vsc_upsampler_set_params");
}

// the fused inference kernel keeps its activations in LDS: no workspace
extern "C" size_t vsc_upsampler_workspace_size(int width, int height)
{
    (void)width;
    (void)height;
    return 0;
}

namespace vsc {

int upsampler_run_flows(vsc_upsampler* u, int n, const float* const flow_lo[2], int lo_w, int lo_h, int lo_ch,
                        const uint8_t* rgba, int width, int height, float scale, float* const flow_out[2], int out_ch,
                        const float* gate, hipStream_t s)
{
    if (!u || !rgba || n < 1 || n > 2) return set_error(VSC_E_INVALID, "vsc_upsampler_run: NULL argument");
    for (int i = 0; i < n; ++i)
        if (!flow_lo[i] || !flow_out[i]) return set_error(VSC_E_INVALID, "vsc_upsampler_run: NULL argument");
    if (lo_w <= 0 || lo_h <= 0 || width <= 0 || height <= 0 || (lo_ch != 2 && lo_ch != 3) ||
        (out_ch != 2 && out_ch != 3))
        return set_error(VSC_E_INVALID, "vsc_upsampler_run: bad shape");
    UpsFlows f{};
    for (int i = 0; i < n; ++i) { f.lo[i] = flow_lo[i]; f.out[i] = flow_out[i]; }
    const int tx = (width + UF_TW - 1) / UF_TW, ty = (height + UF_TH - 1) / UF_TH;
    hipLaunchKernelGGL(k_ups_fused, dim3(tx * ty, n), dim3(512), 0, s, f, lo_w, lo_h, lo_ch, rgba, width, height,
                       scale, u->w16, u->p, gate, out_ch, tx);
    return check_launch("upsampler fused");
}

}  // namespace vsc

extern "C" int vsc_upsampler_run(vsc_upsampler* u, const float* flow_lo, int lo_w, int lo_h, int lo_ch,
                                 const uint8_t* rgba, int width, int height, float scale, float* flow_out, int out_ch,
                                 const float* gate, void* workspace, vsc_hip_stream stream)
{
    (void)workspace;
    const float* const lo[2] = {flow_lo, nullptr};
    float* const out[2] = {flow_out, nullptr};
    return upsampler_run_flows(u, 1, lo, lo_w, lo_h, lo_ch, rgba, width, height, scale, out, out_ch, gate,
                               as_stream(stream));
}

extern "C" int vsc_upsampler_loss_grad(vsc_upsampler* u, int batch, const float* flow_lo, int lo_w, int lo_h,
                                       const uint8_t* rgba, const float* flow_gt, int width, int height, float scale,
                                       float* loss, float* grad_host, vsc_hip_stream stream)
{
    int rc = check_train_args(u, batch, flow_lo, lo_w, lo_h, rgba, flow_gt, width, height);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    if ((rc = fwd_bwd(u, batch, flow_lo, lo_w, lo_h, rgba, flow_gt, width, height, scale, loss, s))) return rc;
    if (grad_host) {
        std::vector<float> g(U_NP);
        if (check_hip(hipMemcpyAsync(g.data(), u->g, U_NP * 4, hipMemcpyDeviceToHost, s), "vsc_upsampler_loss_grad") ||
            check_hip(hipStreamSynchronize(s), "vsc_upsampler_loss_grad"))
            return VSC_E_HIP;
        master_to_ref(g, grad_host);
    }
    return VSC_OK;
}

extern "C" int vsc_upsampler_train_step(vsc_upsampler* u, int batch, const float* flow_lo, int lo_w, int lo_h,
                                        const uint8_t* rgba, const float* flow_gt, int width, int height, float scale,
                                        float lr, float* loss, vsc_hip_stream stream)
{
    int rc = check_train_args(u, batch, flow_lo, lo_w, lo_h, rgba, flow_gt, width, height);
    if (rc) return rc;
    if (!(lr > 0.0f)) return set_error(VSC_E_INVALID, "vsc_upsampler_train_step: lr must be positive");
    hipStream_t s = as_stream(stream);
    if ((rc = fwd_bwd(u, batch, flow_lo, lo_w, lo_h, rgba, flow_gt, width, height, scale, loss, s))) return rc;
    u->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - (float)std::pow(0.9, u->step), bc2 = 1.0f - (float)std::pow(0.999, u->step);
    hipLaunchKernelGGL(k_adam, dim3((U_NP + 255) / 256), dim3(256), 0, s, u->p, u->g, u->m, u->v, U_NP, lr, b1, b2,
                       1e-8f, bc1, bc2);
    if ((rc = check_launch("upsampler adam"))) return rc;
    return derive(u, s);
}

extern "C" int vsc_flow_fb_gate(const float* fwd, const float* bwd, int width, int height, int channels,
                                float max_inconsistent, uint32_t* count, float* gate, vsc_hip_stream stream)
{
    if (!fwd || !bwd || !count || !gate) return set_error(VSC_E_INVALID, "vsc_flow_fb_gate: NULL argument");
    if (width <= 0 || height <= 0 || (channels != 2 && channels != 3))
        return set_error(VSC_E_INVALID, "vsc_flow_fb_gate: bad shape");
    hipStream_t s = as_stream(stream);
    const long long npx = (long long)width * height;
    hipLaunchKernelGGL(k_fb_zero, dim3(1), dim3(64), 0, s, count);
    hipLaunchKernelGGL(k_fb_count, dim3((unsigned)std::min<long long>(512, (npx + 255) / 256)), dim3(256), 0, s, fwd, bwd, width, height,
                       channels, count);
    hipLaunchKernelGGL(k_fb_gate, dim3(1), dim3(64), 0, s, count, npx, max_inconsistent, gate);
    return check_launch("vsc_flow_fb_gate");
}
