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
// FP16 inference weights of the fused kernel: 14 MFMA A operands (16 rows co
// x 32 K), each stored lane-major (lane = 16*lg + co holds K 8*lg .. 8*lg+7):
// layer 1 fragment f holds tap rows dy = 2f + lg/2 (dy 3: zero), lane group
// lg%2 taps dx 0 and 1 (4 channels each, channel 3 zero) or dx 2 and a zero
// tap; layer 2 fragments 0..2 hold taps (dy, dx = lg/2), channels
// 8*(lg%2) .. +7 (output channel co = row), fragments 3..6 tap dx 2 of two
// consecutive input rows (lane group lg/2 the second), at tap rows (0, 1),
// (2, -), (-, 0), (1, 2) (-: zero); layer 3 fragment st holds taps dx = 2*st
// + lg/2 (dx 3: zero) of all three tap rows, row m = 4*co + dy (co < 2, dy <
// 3; other rows zero)
constexpr int H_FR1 = 0, H_FR2 = 2, H_FR3 = 9, H_NFR = 11;
constexpr int H_N = H_NFR * 64 * 8;
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
        const int f = i / 512, lane = (i / 8) % 64, j = i % 8, co = lane % 16, lg = lane / 16;
        float v = 0.0f;
        if (f < H_FR2) {
            const int dy = 2 * f + (lg >> 1), dx = 2 * (lg & 1) + (j >> 2), c = j & 3;
            if (dy < 3 && dx < 3 && c < 3) v = p[U_W1 + (co * 9 + dy * 3 + dx) * U_C1 + c];
        } else if (f < H_FR3) {
            const int g = f - H_FR2, h = lg >> 1, ci = 8 * (lg & 1) + j;
            constexpr int pair_dy[4][2] = {{0, 1}, {2, -1}, {-1, 0}, {1, 2}};
            const int dy = g < 3 ? g : pair_dy[g - 3][h], dx = g < 3 ? h : 2;
            if (dy >= 0) v = p[U_W2 + (co * 9 + dy * 3 + dx) * 16 + ci];
        } else {
            const int m = co, oc = m >> 2, dy = m & 3, dx = 2 * (f - H_FR3) + (lg >> 1), ci = 8 * (lg & 1) + j;
            if (oc < 2 && dy < 3 && dx < 3) v = p[U_W3 + (oc * 9 + dy * 3 + dx) * 16 + ci];
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
// Fused FP16 inference: the three convolutions of one 20 x 28 output tile in
// one block, activations in LDS.  The input tile (26 x 34 pixels, a 3-pixel
// halo) holds (u, v, luma, 0) per pixel; layer 1 writes a 24 x 32 tile
// (2-pixel halo), layer 2 a 22 x 32 tile (1-pixel halo and two spare
// columns), layer 3 the 20 x 28 output.  Activations of pixels outside the
// frame are zero (each layer's zero padding).
//
// Each wave computes a strip of R rows x 16 columns of a layer as MFMA GEMMs
// D[co][px] = W[co][K] * X[K][px] (A = weights in registers, B = 16 pixels
// of one row) and reads the R + 2 input rows of its strip once: the
// fragments of input row rho feed output rows rho - 2 .. rho, each with the
// weights of its tap row.  Layer 1 (4 channels per pixel) takes one K-step
// per two tap rows (rows dy 0 and 1; row dy 2 beside a zero-weight row);
// layer 2 one K-step per tap row for taps dx 0 and 1 and one per two input
// rows for tap dx 2 (ups_strip16p); layer 3, whose 2 output channels would
// leave 14 of the MFMA's 16 rows idle, holds (channel, tap row) in the rows
// instead and rolls its accumulator from input row to input row
// (ups_strip_out).  Every LDS address is a per-lane base plus a compile-time
// offset, the accumulators start at the bias, and a row's epilogue is two
// packed FP16 conversions and a packed ReLU.
// ---------------------------------------------------------------------------
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 up_f16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float up_f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t up_u32x4;
typedef _Float16 up_h2 __attribute__((ext_vector_type(2)));

constexpr int UF_TH = 20, UF_TW = 28;                     // output tile
constexpr int UF_XP = UF_TW + 6, UF_XR = UF_TH + 8;       // input tile pitch (pixels); rows 26, 27 are zero
constexpr int UF_AP = 32;                                 // activation tile pitch (pixels)
constexpr int UF_AS = 24;                                 // halfs per activation pixel: 16 channels + pad (48 B)
constexpr int UF_ROWB = UF_AP * UF_AS * 2;                // bytes per activation row
constexpr int UF_1R = UF_TH + 4, UF_2R = UF_TH + 2;       // rows of the layer-1 and layer-2 tiles
// one zero row after each activation tile: the zero-weight column of a
// strip's last row reads two pixels into it
// (layer 1's also 128 bytes more: layer 2's paired tap dx 2 reads the first
// two pixels of the row after it)
constexpr int UF_S1B = (UF_1R + 1) * UF_ROWB + 128, UF_S2B = (UF_2R + 1) * UF_ROWB;
static_assert(UF_XR * UF_XP * 8 <= UF_2R * UF_ROWB, "the input tile lies under layer 2's rows");
static_assert(UF_TW + 4 == UF_AP, "layer 1 fills the activation pitch");

__device__ inline up_f32x4 up_mma(const up_u32x4& a, const up_u32x4& b, up_f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(up_f16x8, a), __builtin_bit_cast(up_f16x8, b), c,
                                                  0, 0, 0);
}

__device__ inline uint32_t pack_h2(float a, float b)
{
    return (uint32_t)f32_to_f16(a) | ((uint32_t)f32_to_f16(b) << 16);
}

// ReLU(x) of the lane's four channels as four FP16 values (round to nearest)
__device__ inline uint2 relu_h4(const up_f32x4& x)
{
    const up_h2 z = {(_Float16)0.0f, (_Float16)0.0f};
    up_h2 a, b;
    a.x = (_Float16)x[0];
    a.y = (_Float16)x[1];
    b.x = (_Float16)x[2];
    b.y = (_Float16)x[3];
    a = __builtin_elementwise_max(a, z);
    b = __builtin_elementwise_max(b, z);
    return make_uint2(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b));
}

// A strip of layer 2: rows 0 .. R-1 of the strip, put(r, acc) once row r
// has all its terms.  The fragment of taps dx 0 and 1 of input row rho (src:
// the lane's byte address of input row 0, pixel lg/2, channels 8*(lg%2)..)
// feeds output rows rho - 2 .. rho with the weights of tap rows 2 .. 0; tap
// dx 2 of input rows 2k and 2k+1 shares one K-step (src2: pixel 2 of input
// row lg/2) and feeds output rows 2k-2 .. 2k+1 with the tap-row pairs (2, -),
// (1, 2), (0, 1), (-, 0): five MFMAs per output row.
template <int R, typename Put>
__device__ inline void ups_strip16p(const uint8_t* src, const uint8_t* src2, const up_u32x4 (&A)[3],
                                    const up_u32x4 (&P)[4], const up_f32x4& bias, Put put)
{
    up_f32x4 acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = bias;
    up_u32x4 fa[2], fp[2];
    fa[0] = *reinterpret_cast<const up_u32x4*>(src);
    fp[0] = *reinterpret_cast<const up_u32x4*>(src2);
#pragma unroll
    for (int rho = 0; rho < R + 2; ++rho) {
        if (rho + 1 < R + 2) fa[(rho + 1) & 1] = *reinterpret_cast<const up_u32x4*>(src + (rho + 1) * UF_ROWB);
        if ((rho & 1) && rho + 1 < R + 2)
            fp[((rho + 1) >> 1) & 1] = *reinterpret_cast<const up_u32x4*>(src2 + (rho + 1) * UF_ROWB);
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
            const int r = rho - dy;
            if (r >= 0 && r < R) acc[r] = up_mma(A[dy], fa[rho & 1], acc[r]);
        }
        if (!(rho & 1)) {
            const up_u32x4& f = fp[(rho >> 1) & 1];
            if (rho - 2 >= 0 && rho - 2 < R) acc[rho - 2] = up_mma(P[1], f, acc[rho - 2]);
            if (rho - 1 >= 0 && rho - 1 < R) acc[rho - 1] = up_mma(P[3], f, acc[rho - 1]);
            if (rho < R) acc[rho] = up_mma(P[0], f, acc[rho]);
            if (rho + 1 < R) acc[rho + 1] = up_mma(P[2], f, acc[rho + 1]);
        }
        if (rho >= 2) put(rho - 2, acc[rho - 2]);
    }
}

// A strip of layer 3: D row m = 4*co + dy holds, after input row rho, the
// partial sum of output channel co of output row rho - dy, so each input row
// takes two MFMAs for all three output rows it feeds; the accumulator shifts
// by one element per input row and element 2 leaves complete.  put(r, v):
// output row r, channel lg of the lane's pixel (lanes lg < 2).
template <int R, typename Put>
__device__ inline void ups_strip_out(const uint8_t* src, const up_u32x4 (&A)[2], float bias, Put put)
{
    up_f32x4 acc = up_f32x4{bias, bias, bias, bias};
    up_u32x4 fa[2], fb[2];
    fa[0] = *reinterpret_cast<const up_u32x4*>(src);
    fb[0] = *reinterpret_cast<const up_u32x4*>(src + 2 * UF_AS * 2);
#pragma unroll
    for (int rho = 0; rho < R + 2; ++rho) {
        if (rho + 1 < R + 2) {
            fa[(rho + 1) & 1] = *reinterpret_cast<const up_u32x4*>(src + (rho + 1) * UF_ROWB);
            fb[(rho + 1) & 1] = *reinterpret_cast<const up_u32x4*>(src + (rho + 1) * UF_ROWB + 2 * UF_AS * 2);
        }
        const up_f32x4 c = up_f32x4{bias, acc[0], acc[1], acc[2]};
        acc = up_mma(A[0], fa[rho & 1], c);
        acc = up_mma(A[1], fb[rho & 1], acc);
        if (rho >= 2) put(rho - 2, acc[2]);
    }
}

// A strip of layer 1 (src: the lane's byte address of input-tile row lg/2,
// pixel 2*(lg%2) of the lane's column)
template <int R, typename Put>
__device__ inline void ups_strip4(const uint8_t* src, const up_u32x4 (&A)[2], const up_f32x4& bias, Put put)
{
    up_f32x4 acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = bias;
    auto load = [&](int rho) {
        const uint2 a = *reinterpret_cast<const uint2*>(src + rho * UF_XP * 8);
        const uint2 b = *reinterpret_cast<const uint2*>(src + rho * UF_XP * 8 + 8);
        return up_u32x4{a.x, a.y, b.x, b.y};
    };
    up_u32x4 f[2];
    f[0] = load(0);
#pragma unroll
    for (int rho = 0; rho < R + 2; ++rho) {
        if (rho + 1 < R + 2) f[(rho + 1) & 1] = load(rho + 1);
        if (rho < R) acc[rho] = up_mma(A[0], f[rho & 1], acc[rho]);
        if (rho >= 2) {
            acc[rho - 2] = up_mma(A[1], f[rho & 1], acc[rho - 2]);
            put(rho - 2, acc[rho - 2]);
        }
    }
}

// blockIdx.y selects the flow (up to two flows of one frame per launch)
struct UpsFlows { const float* lo[2]; float* out[2]; };

__global__ __launch_bounds__(512) void k_ups_fused(UpsFlows flows, int lw, int lh, int lch,
                                                   const uint8_t* __restrict__ rgba, int W, int H, float scale,
                                                   const uint16_t* __restrict__ wf, const float* __restrict__ prm,
                                                   const float* gate, int out_ch, int tiles_x, bool vec2)
{
    const float* __restrict__ lo = flows.lo[blockIdx.y];
    float* __restrict__ out = flows.out[blockIdx.y];
    constexpr int NT = 512;
    __shared__ __attribute__((aligned(16))) uint8_t s1[UF_S1B];
    __shared__ __attribute__((aligned(16))) uint8_t s2[UF_S2B];   // the input tile until layer 1 is done
    __shared__ float2 sbase[UF_TH * UF_TW];                        // FP32 bilinear flow of the output tile
    __shared__ int tx0[UF_XP], tx1[UF_XP], ty0[UF_XR], ty1[UF_XR];
    __shared__ float tfx[UF_XP], tfy[UF_XR];
    uint2* const sx = reinterpret_cast<uint2*>(s2);
    const int tid = threadIdx.x, q = tid / 64, lane = tid % 64, l16 = lane % 16, lg = lane / 16;
    const int y0 = (blockIdx.x / tiles_x) * UF_TH, x0 = (blockIdx.x % tiles_x) * UF_TW;
    // bilinear taps of the tile's columns and rows (kernel_bilinear mapping)
    if (tid < UF_XP) {
        const int ix = min(max(x0 - 3 + tid, 0), W - 1);
        const BiTap t = bilinear_tap(ix, 0, lw, lh, W, H);
        tx0[tid] = t.ix; tx1[tid] = t.ix1; tfx[tid] = t.fx;
    } else if (tid >= 64 && tid < 64 + UF_XR) {
        const int iy = min(max(y0 - 3 + tid - 64, 0), H - 1);
        const BiTap t = bilinear_tap(0, iy, lw, lh, W, H);
        ty0[tid - 64] = t.iy; ty1[tid - 64] = t.iy1; tfy[tid - 64] = t.fy;
    }
    // the zero rows after the activation tiles (never written again)
    if (tid < (UF_S1B - UF_1R * UF_ROWB) / 16)
        reinterpret_cast<uint4*>(s1 + UF_1R * UF_ROWB)[tid] = make_uint4(0u, 0u, 0u, 0u);
    if (tid < UF_ROWB / 16)
        reinterpret_cast<uint4*>(s2 + UF_2R * UF_ROWB)[tid] = make_uint4(0u, 0u, 0u, 0u);

    // weights: MFMA A fragments (rows co = l16), 8 halfs per lane, lane-major
    auto frag = [&](int f) { return *reinterpret_cast<const up_u32x4*>(wf + (f * 64 + lane) * 8); };
    up_u32x4 a1[2], a2[3], a2p[4], a3[2];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        a1[f] = frag(H_FR1 + f);
        a3[f] = frag(H_FR3 + f);
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) a2[f] = frag(H_FR2 + f);
#pragma unroll
    for (int f = 0; f < 4; ++f) a2p[f] = frag(H_FR2 + 3 + f);
    up_f32x4 b1, b2;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        b1[r] = prm[U_B1 + 4 * lg + r];
        b2[r] = prm[U_B2 + 4 * lg + r];
    }
    const float b3 = lg < 2 ? prm[U_B3 + lg] : 0.0f;
    const float gv = gate ? *gate : 1.0f;
    __syncthreads();

    // wave q: column group cg (16 columns), strip s of rows
    const int cg = q & 1, s = q >> 1, c = 16 * cg + l16;
    // Tiles whose input tile (with its 3-pixel halo) lies inside the frame
    // skip every frame test (block-uniform choice).
    const bool interior = y0 >= 3 && y0 + UF_TH + 3 <= H && x0 >= 3 && x0 + UF_TW + 3 <= W;
    auto run = [&](auto interior_c) {
        constexpr bool INTERIOR = decltype(interior_c)::value;
        // input tile: (u, v) = bilinear x scale, luma; zero outside the frame.
        // A thread's two pixels issue all their loads before either is
        // computed (addresses clamped into the frame, values dropped after).
        constexpr int NX = UF_XR * UF_XP, NE = (NX + NT - 1) / NT;
        float2 v00[NE], v10[NE], v01[NE], v11[NE];
        uint32_t pix[NE];
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const int e = min(tid + NT * k, NX - 1), r = e / UF_XP, cc = e % UF_XP;
            const int a0 = tx0[cc], a1_ = tx1[cc], r0 = ty0[r] * lw, r1 = ty1[r] * lw;
            if (vec2) {
                const float2* f2 = reinterpret_cast<const float2*>(lo);
                v00[k] = f2[r0 + a0]; v10[k] = f2[r0 + a1_]; v01[k] = f2[r1 + a0]; v11[k] = f2[r1 + a1_];
            } else {
                v00[k] = make_float2(lo[(r0 + a0) * lch], lo[(r0 + a0) * lch + 1]);
                v10[k] = make_float2(lo[(r0 + a1_) * lch], lo[(r0 + a1_) * lch + 1]);
                v01[k] = make_float2(lo[(r1 + a0) * lch], lo[(r1 + a0) * lch + 1]);
                v11[k] = make_float2(lo[(r1 + a1_) * lch], lo[(r1 + a1_) * lch + 1]);
            }
            const int iy = min(max(y0 - 3 + r, 0), H - 1), ix = min(max(x0 - 3 + cc, 0), W - 1);
            pix[k] = *reinterpret_cast<const uint32_t*>(rgba + ((long long)iy * W + ix) * 4);
        }
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const int e = tid + NT * k, r = e / UF_XP, cc = e % UF_XP;
            if (e >= NX) break;
            const int iy = y0 - 3 + r, ix = x0 - 3 + cc;
            uint2 v = make_uint2(0u, 0u);
            if (r < UF_TH + 6 && (INTERIOR || (iy >= 0 && iy < H && ix >= 0 && ix < W))) {
                // bilinear_sample's operation order
                const float fx = tfx[cc], fy = tfy[r], gx = 1.0f - fx, gy = 1.0f - fy;
                const float u = ((((v00[k].x * gx) * gy + (v10[k].x * fx) * gy) + (v01[k].x * gx) * fy) +
                                 (v11[k].x * fx) * fy) * scale;
                const float w = ((((v00[k].y * gx) * gy + (v10[k].y * fx) * gy) + (v01[k].y * gx) * fy) +
                                 (v11[k].y * fx) * fy) * scale;
                const float luma = (0.299f * (float)(pix[k] & 0xffu) + 0.587f * (float)((pix[k] >> 8) & 0xffu) +
                                    0.114f * (float)((pix[k] >> 16) & 0xffu)) * (1.0f / 255.0f);
                v = make_uint2(pack_h2(u, w), pack_h2(luma, 0.0f));
                if (r >= 3 && r < 3 + UF_TH && cc >= 3 && cc < 3 + UF_TW)
                    sbase[(r - 3) * UF_TW + cc - 3] = make_float2(u, w);
            }
            sx[e] = v;
        }
        __syncthreads();

        // layer 1: rows 6s .. 6s+5 of the 24 x 32 tile, ReLU, zero outside the frame
        {
            const int r0 = 6 * s;
            const uint8_t* src = s2 + ((r0 + (lg >> 1)) * UF_XP + c + 2 * (lg & 1)) * 8;
            uint8_t* dst = s1 + (r0 * UF_AP + c) * UF_AS * 2 + 8 * lg;
            const bool col_in = INTERIOR || (x0 - 2 + c >= 0 && x0 - 2 + c < W);
            ups_strip4<6>(src, a1, b1, [&](int r, const up_f32x4& acc) {
                uint2 o = relu_h4(acc);
                if (!INTERIOR) {
                    const int iy = y0 - 2 + r0 + r;
                    if (!(col_in && iy >= 0 && iy < H)) o = make_uint2(0u, 0u);
                }
                *reinterpret_cast<uint2*>(dst + r * UF_ROWB) = o;
            });
        }
        __syncthreads();

        // layer 2: strips of 6, 5, 6, 5 rows of the 22 x 32 tile
        {
            const int r0 = (s >> 1) * 11 + (s & 1) * 6;
            const uint8_t* src = s1 + ((r0 * UF_AP + c + (lg >> 1)) * UF_AS + 8 * (lg & 1)) * 2;
            const uint8_t* src2 = s1 + (((r0 + (lg >> 1)) * UF_AP + c + 2) * UF_AS + 8 * (lg & 1)) * 2;
            uint8_t* dst = s2 + (r0 * UF_AP + c) * UF_AS * 2 + 8 * lg;
            const bool col_in = INTERIOR || (x0 - 1 + c >= 0 && x0 - 1 + c < W);
            auto put = [&](int r, const up_f32x4& acc) {
                uint2 o = relu_h4(acc);
                if (!INTERIOR) {
                    const int iy = y0 - 1 + r0 + r;
                    if (!(col_in && iy >= 0 && iy < H)) o = make_uint2(0u, 0u);
                }
                *reinterpret_cast<uint2*>(dst + r * UF_ROWB) = o;
            };
            if (s & 1) ups_strip16p<5>(src, src2, a2, a2p, b2, put);
            else ups_strip16p<6>(src, src2, a2, a2p, b2, put);
        }
        __syncthreads();

        // layer 3: rows 5s .. 5s+4 of the output tile, added to the FP32
        // bilinear flow; lane group lg < 2 holds output channel lg
        {
            const int r0 = 5 * s;
            const uint8_t* src = s2 + ((r0 * UF_AP + c + (lg >> 1)) * UF_AS + 8 * (lg & 1)) * 2;
            const bool st = lg < 2 && c < UF_TW && (INTERIOR || x0 + c < W);
            float* o = out + ((long long)(y0 + r0) * W + x0 + c) * out_ch + lg;
            const float* bs = reinterpret_cast<const float*>(sbase) + (r0 * UF_TW + c) * 2 + lg;
            ups_strip_out<5>(src, a3, b3, [&](int r, float v) {
                if (st && (INTERIOR || y0 + r0 + r < H)) {
                    float* p = o + (long long)r * W * out_ch;
                    p[0] = bs[r * UF_TW * 2] + gv * v;
                    if (out_ch == 3 && lg == 1) p[1] = 0.0f;
                }
            });
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
    if (reinterpret_cast<uintptr_t>(rgba) % 4 != 0)
        return set_error(VSC_E_INVALID, "vsc_upsampler_run: rgba must be 4-byte aligned");
    UpsFlows f{};
    bool vec2 = lo_ch == 2;   // both flow channels of a pixel from one 8-byte load
    for (int i = 0; i < n; ++i) {
        f.lo[i] = flow_lo[i];
        f.out[i] = flow_out[i];
        vec2 = vec2 && reinterpret_cast<uintptr_t>(flow_lo[i]) % 8 == 0;
    }
    const int tx = (width + UF_TW - 1) / UF_TW, ty = (height + UF_TH - 1) / UF_TH;
    hipLaunchKernelGGL(k_ups_fused, dim3(tx * ty, n), dim3(512), 0, s, f, lo_w, lo_h, lo_ch, rgba, width, height,
                       scale, u->w16, u->p, gate, out_ch, tx, vec2);
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
