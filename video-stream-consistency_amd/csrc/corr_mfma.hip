// Reduced-precision Correlation on MFMA (gfx950): BF16 / FP16 / FP8-e4m3
// features, FP32 accumulation.  Same output layout and zero-padding semantics
// as the reference op (correlation.cc:147-275; legacy: correlation_cuda.cu:182-331).
//
// Matrix form.  For output row h, a segment of 16 pixels w0..w0+15 and one
// vertical displacement dy, the 16 x 24 product
//     G[i][j] = sum_c f1[c][h][w0+i] * f2[c][h+dy][w0-4+j]
// holds all horizontal displacements on its diagonals: out[dy][dx][h][w0+i] =
// G[i][i+dx+4].  G is computed as two 16 x 16 MFMA tiles (columns 0..15 and
// 8..23) with K = 32 channels per instruction (v_mfma_f32_16x16x32_{bf16,f16,
// fp8_fp8}); 9 of every 16 columns are used, which is fine because the op is
// bound by the 81 output planes it writes, not by the products.
//
// Block: 8 waves; wave q computes one 16-pixel segment (q & 1) of output row
// h0 + q/2 for all (2*MD+1)^2 displacements, 72 FP32 accumulators per lane.  The
// NCHW operands are first copied channels-last (channels padded to 32, zeros)
// by a bandwidth-bound conversion kernel; channels then stream through LDS in
// chunks of 32 as straight 8-byte copies: the f1 tile (4 rows x 32 px) and the
// f2 tile (12 rows x 40 px, MD-pixel halo, zeros outside the image), with a
// padded pixel stride, so that each lane's 8-channel MFMA operand is one
// 16-byte (8-byte for FP8) LDS read and 8 lanes cover all banks.
// After the channel loop, each wave moves the diagonals of its tiles through a
// small LDS stage and writes every (dy, dx) output row as 32 consecutive floats.
#include "vsc_internal.h"

#include <algorithm>
#include <type_traits>

namespace vsc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

static constexpr int CM_ROWS = 4;    // output rows per block (one per wave)
static constexpr int CM_PX = 32;     // output pixels per row per block
static constexpr int CM_HALO = 4;    // f2 halo (max displacement supported)
static constexpr int CM_KC = 32;     // channels per chunk (one MFMA K step)

// element type traits: bytes per element, LDS pixel stride (elements)
template <int EB> struct CmLayout;
template <> struct CmLayout<2> { static constexpr int STRIDE = 40; };   // 80 B: 20 dwords
template <> struct CmLayout<1> { static constexpr int STRIDE = 40; };   // 40 B: 10 dwords

enum CmKind { CM_BF16 = 0, CM_F16 = 1, CM_FP8 = 2 };

template <int KIND>
__device__ inline f32x4 cm_mfma(const void* a, const void* b, f32x4 acc)
{
    if constexpr (KIND == CM_BF16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(a),
                                                       *reinterpret_cast<const bf16x8*>(b), acc, 0, 0, 0);
    } else if constexpr (KIND == CM_F16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(*reinterpret_cast<const f16x8*>(a),
                                                      *reinterpret_cast<const f16x8*>(b), acc, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(*reinterpret_cast<const long*>(a),
                                                          *reinterpret_cast<const long*>(b), acc, 0, 0, 0);
    }
}

// One 8-byte LDS vector (GC = 8/EB channels of one pixel) for each of 4
// consecutive pixels, from GC channel-plane loads of those 4 pixels (NCHW):
// a register transpose with byte permutes.  raw[k] holds pixels 0..3 of
// channel k: 2 x 32 bits (16-bit elements) or 32 bits (FP8).
__device__ inline uint2 nchw_px16(const uint2 (&raw)[4], int j)
{
    const uint32_t sel = (j & 1) ? 0x07060302u : 0x05040100u;
    const bool hi = j >= 2;
    return make_uint2(__builtin_amdgcn_perm(hi ? raw[1].y : raw[1].x, hi ? raw[0].y : raw[0].x, sel),
                      __builtin_amdgcn_perm(hi ? raw[3].y : raw[3].x, hi ? raw[2].y : raw[2].x, sel));
}
__device__ inline uint2 nchw_px8(const uint32_t (&raw)[8], int j)
{
    const uint32_t sel = (uint32_t)j | ((uint32_t)(4 + j) << 8);
    const uint32_t t01 = __builtin_amdgcn_perm(raw[1], raw[0], sel), t23 = __builtin_amdgcn_perm(raw[3], raw[2], sel);
    const uint32_t t45 = __builtin_amdgcn_perm(raw[5], raw[4], sel), t67 = __builtin_amdgcn_perm(raw[7], raw[6], sel);
    return make_uint2(__builtin_amdgcn_perm(t23, t01, 0x05040100u), __builtin_amdgcn_perm(t67, t45, 0x05040100u));
}

// NCHW: the operands are read in their NCHW layout (width a multiple of 4,
// 8-byte (16-bit) or 4-byte (FP8) aligned): each staging unit loads 4 pixels
// of GC channel planes and writes them channels-last into LDS, so no
// conversion pass runs before the kernel.  Otherwise the operands are
// channels-last copies [n][y][x][CP].
template <int KIND, int MD, bool LEGACY, bool NCHW>
__device__ __forceinline__ void corr_mfma_tile(const void* __restrict__ f1v, const void* __restrict__ f2v,
                                               float* __restrict__ out, int C, int CP, int H, int W,
                                               const float* scale)
{
    constexpr int EB = (KIND == CM_FP8) ? 1 : 2;          // bytes per element
    constexpr int ST = CmLayout<EB>::STRIDE;              // LDS pixel stride in elements
    constexpr int P = 2 * MD + 1;
    constexpr int BR = CM_ROWS + 2 * CM_HALO, BC = CM_PX + 2 * CM_HALO;
    using E = typename std::conditional<EB == 2, uint16_t, uint8_t>::type;
    // LDS: the operand tiles during the channel loop, then the whole output
    // tile [row][plane][32 px] at a plane pitch of 36 floats (the operand
    // tiles are dead by then): the 16 lanes of a K group write 16 planes at
    // once, which a 32-float pitch puts on the same banks
    constexpr int SP = CM_PX + 4;
    constexpr int A_BYTES = CM_ROWS * CM_PX * ST * EB, B_BYTES = BR * BC * ST * EB;
    constexpr int OUT_BYTES = CM_ROWS * P * P * SP * 4;
    constexpr int LDS_BYTES = (A_BYTES + B_BYTES) > OUT_BYTES ? (A_BYTES + B_BYTES) : OUT_BYTES;
    __shared__ __attribute__((aligned(16))) unsigned char lds[LDS_BYTES];
    E (*sA)[CM_PX][ST] = reinterpret_cast<E (*)[CM_PX][ST]>(lds);
    E (*sB)[BC][ST] = reinterpret_cast<E (*)[BC][ST]>(lds + A_BYTES);
    float (*stage)[P * P][SP] = reinterpret_cast<float (*)[P * P][SP]>(lds);

    // operands are channels-last copies [n][y][x][CP] (CP: C padded to a
    // multiple of 32 with zeros), so a chunk of 32 channels of one pixel is
    // contiguous: staged with 8-byte vectors (VPP per pixel), no transpose
    constexpr int VPP = CM_KC * EB / 8;
    const E* f1 = static_cast<const E*>(f1v);
    const E* f2 = static_cast<const E*>(f2v);
    const int n = blockIdx.z;
    const int h0 = blockIdx.y * CM_ROWS, w0 = blockIdx.x * CM_PX;
    const long long HW = (long long)H * W;
    const E* a = f1 + (long long)n * HW * CP;
    const E* b = f2 + (long long)n * HW * CP;
    const int tid = threadIdx.x;
    const int q = tid / 64, lane = tid % 64;
    const int qr = q >> 1, seg = q & 1;         // output row in the block, 16-pixel segment
    const int lr = lane % 16, lg = lane / 16;   // operand row/column, K group

    f32x4 acc[P][2];
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[d][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // NCHW staging units: (channel group of GC, row, group of 4 pixels)
    constexpr int GC = 8 / EB, CG = CM_KC / GC;
    constexpr int NA4 = CG * CM_ROWS * (CM_PX / 4), NB4 = CG * BR * (BC / 4), NU4 = NA4 + NB4;
    constexpr int RU4 = (NU4 + 511) / 512;
    using Raw = typename std::conditional<EB == 2, uint2, uint32_t>::type;
    Raw raw[NCHW ? RU4 : 1][GC];
    auto fetch4 = [&](int c0) {
        const E* fa = f1 + (long long)n * C * HW;
        const E* fb = f2 + (long long)n * C * HW;
#pragma unroll
        for (int k = 0; k < RU4; ++k) {
            const int u = tid + 512 * k;
            const bool inA = u < NA4;
            const int ul = inA ? u : u - NA4;
            const int rows = inA ? CM_ROWS : BR, groups = inA ? CM_PX / 4 : BC / 4;
            const int g = ul % groups, r = (ul / groups) % rows, cg = ul / (groups * rows);
            const int gh = inA ? h0 + r : h0 - CM_HALO + r;
            const int gw = inA ? w0 + 4 * g : w0 - CM_HALO + 4 * g;
            const bool ok = u < NU4 && gh >= 0 && gh < H && gw >= 0 && gw < W;
            const E* src = (inA ? fa : fb) + (long long)gh * W + gw;
#pragma unroll
            for (int j = 0; j < GC; ++j) {
                const int c = c0 + cg * GC + j;
                Raw v{};
                if (ok && c < C) v = *reinterpret_cast<const Raw*>(src + (long long)c * HW);
                raw[k][j] = v;
            }
        }
    };
    auto stage4 = [&]() {
#pragma unroll
        for (int k = 0; k < RU4; ++k) {
            const int u = tid + 512 * k;
            if (u >= NU4) continue;
            const bool inA = u < NA4;
            const int ul = inA ? u : u - NA4;
            const int rows = inA ? CM_ROWS : BR, groups = inA ? CM_PX / 4 : BC / 4;
            const int g = ul % groups, r = (ul / groups) % rows, cg = ul / (groups * rows);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint2 v;
                if constexpr (EB == 2) v = nchw_px16(raw[k], j);
                else v = nchw_px8(raw[k], j);
                E* dst = inA ? &sA[r][4 * g + j][cg * GC] : &sB[r][4 * g + j][cg * GC];
                *reinterpret_cast<uint2*>(dst) = v;
            }
        }
    };

    constexpr int NA = CM_ROWS * CM_PX * VPP, NB = BR * BC * VPP, NU = NA + NB, RU = (NU + 511) / 512;
    uint2 rg[NCHW ? 1 : RU];
    auto fetch = [&](int c0) {
        if constexpr (NCHW) {
            fetch4(c0);
        } else {
#pragma unroll
            for (int k = 0; k < RU; ++k) {
                const int u = tid + 512 * k;
                uint2 v = make_uint2(0u, 0u);
                if (u < NA) {
                    const int p = u / VPP, cv = u % VPP, r = p / CM_PX, x = p % CM_PX;
                    const int gh = h0 + r, gw = w0 + x;
                    if (gh < H && gw < W)
                        v = *reinterpret_cast<const uint2*>(a + ((long long)gh * W + gw) * CP + c0 + cv * (8 / EB));
                } else if (u < NU) {
                    const int ub = u - NA, p = ub / VPP, cv = ub % VPP, r = p / BC, x = p % BC;
                    const int gh = h0 - CM_HALO + r, gw = w0 - CM_HALO + x;
                    if (gh >= 0 && gh < H && gw >= 0 && gw < W)
                        v = *reinterpret_cast<const uint2*>(b + ((long long)gh * W + gw) * CP + c0 + cv * (8 / EB));
                }
                rg[k] = v;
            }
        }
    };

    fetch(0);
    for (int c0 = 0; c0 < CP; c0 += CM_KC) {
        __syncthreads();
        if constexpr (NCHW) stage4();
#pragma unroll
        for (int k = 0; k < (NCHW ? 0 : RU); ++k) {
            const int u = tid + 512 * k;
            if (u < NA) {
                const int p = u / VPP, cv = u % VPP;
                *reinterpret_cast<uint2*>(&sA[p / CM_PX][p % CM_PX][cv * (8 / EB)]) = rg[k];
            } else if (u < NU) {
                const int ub = u - NA, p = ub / VPP, cv = ub % VPP;
                *reinterpret_cast<uint2*>(&sB[p / BC][p % BC][cv * (8 / EB)]) = rg[k];
            }
        }
        __syncthreads();
        if (c0 + CM_KC < CP) fetch(c0 + CM_KC);
        // A operand: pixel lr of this wave's segment, channels 8*lg .. 8*lg+7
        const E* pa = &sA[qr][16 * seg + lr][8 * lg];
#pragma unroll
        for (int d = 0; d < P; ++d) {
            const int br = qr + CM_HALO + d - MD;   // f2 row of displacement dy = d - MD
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const E* pb = &sB[br][16 * seg + 8 * t + lr][8 * lg];
                acc[d][t] = cm_mfma<KIND>(pa, pb, acc[d][t]);
            }
        }
    }

    // FP8: one dequantisation factor s = s1*s2 (per-channel factors are folded
    // into the operands by the equalisation, see vsc_quantize_fp8_pair)
    const float sc = scale ? *scale : 1.0f;
    // diagonals -> stage[row][plane][pixel]: lane holds G[4*lg + r][lr + 8*t]
    __syncthreads();
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 4 * lg + r, j = lr + 8 * t;
                const int dx = j - i - CM_HALO;
                if (dx >= -MD && dx <= MD && !(t == 1 && j < 16)) {
                    float v = acc[d][t][r];
                    if (KIND == CM_FP8) v = v * sc;
                    if (LEGACY) v = v / (float)C;
                    stage[qr][d * P + dx + MD][16 * seg + i] = v;
                }
            }
    __syncthreads();
    // each group of 8 lanes writes one 128-byte row segment of one plane, 16
    // bytes per lane where the 4 pixels are in the image and the output rows
    // are 16-byte aligned (a 16-byte aligned output, W a multiple of 4);
    // the cost volume is not re-read by this kernel: non-temporal stores
    const bool vec = (W & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    for (int e = tid; e < CM_ROWS * P * P * CM_PX / 4; e += 512) {
        const int x = 4 * (e % (CM_PX / 4)), pl = (e / (CM_PX / 4)) % (P * P), r = e / (CM_PX / 4 * P * P);
        const int h = h0 + r, w = w0 + x;
        if (h >= H) continue;
        float* o = out + ((long long)n * P * P + pl) * HW + (long long)h * W + w;
        if (vec && w + 3 < W) {
            __builtin_nontemporal_store(*reinterpret_cast<const f32x4*>(&stage[r][pl][x]), reinterpret_cast<f32x4*>(o));
        } else {
            for (int j = 0; j < 4; ++j)
                if (w + j < W) __builtin_nontemporal_store(stage[r][pl][x + j], o + j);
        }
    }
}

// The kernel as is (152 VGPRs for md = 4: one 512-thread block per CU), and
// the same tile capped at 128 VGPRs for two blocks per CU (13 of them spill
// to scratch at md = 4).  Grids of at least 256 blocks take the capped form
// (1920x1088 pyramid, BF16: 32.9 vs 38.4 us at 480x272, 19.7 vs 21.5 us at
// 240x136); smaller grids lose to the spills (120x68: 15.7 vs 13.9 us).
template <int KIND, int MD, bool LEGACY, bool NCHW>
__global__ __launch_bounds__(512) void k_corr_mfma(const void* __restrict__ f1v, const void* __restrict__ f2v,
                                                   float* __restrict__ out, int C, int CP, int H, int W,
                                                   const float* scale)
{
    corr_mfma_tile<KIND, MD, LEGACY, NCHW>(f1v, f2v, out, C, CP, H, W, scale);
}

template <int KIND, int MD, bool LEGACY, bool NCHW>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_corr_mfma2(
    const void* __restrict__ f1v, const void* __restrict__ f2v, float* __restrict__ out, int C, int CP, int H, int W,
    const float* scale)
{
    corr_mfma_tile<KIND, MD, LEGACY, NCHW>(f1v, f2v, out, C, CP, H, W, scale);
}

template <int KIND, int MD>
static void launch_mfma(const void* f1, const void* f2, float* out, int N, int C, int CP, int H, int W, bool legacy,
                        bool nchw, const float* scale, hipStream_t s)
{
    dim3 grid((W + CM_PX - 1) / CM_PX, (H + CM_ROWS - 1) / CM_ROWS, N);
    const bool two = (long long)grid.x * grid.y * grid.z >= 256;
    if (legacy && nchw)
        hipLaunchKernelGGL(two ? (k_corr_mfma2<KIND, MD, true, true>) : (k_corr_mfma<KIND, MD, true, true>), grid,
                           dim3(512), 0, s, f1, f2, out, C, CP, H, W, scale);
    else if (legacy)
        hipLaunchKernelGGL(two ? (k_corr_mfma2<KIND, MD, true, false>) : (k_corr_mfma<KIND, MD, true, false>), grid,
                           dim3(512), 0, s, f1, f2, out, C, CP, H, W, scale);
    else if (nchw)
        hipLaunchKernelGGL(two ? (k_corr_mfma2<KIND, MD, false, true>) : (k_corr_mfma<KIND, MD, false, true>), grid,
                           dim3(512), 0, s, f1, f2, out, C, CP, H, W, scale);
    else
        hipLaunchKernelGGL(two ? (k_corr_mfma2<KIND, MD, false, false>) : (k_corr_mfma<KIND, MD, false, false>), grid,
                           dim3(512), 0, s, f1, f2, out, C, CP, H, W, scale);
}

template <int KIND>
static int dispatch_md(const void* f1, const void* f2, float* out, int N, int C, int CP, int H, int W, int md,
                       bool legacy, bool nchw, const float* scale, hipStream_t s)
{
    switch (md) {
        case 0: launch_mfma<KIND, 0>(f1, f2, out, N, C, CP, H, W, legacy, nchw, scale, s); break;
        case 1: launch_mfma<KIND, 1>(f1, f2, out, N, C, CP, H, W, legacy, nchw, scale, s); break;
        case 2: launch_mfma<KIND, 2>(f1, f2, out, N, C, CP, H, W, legacy, nchw, scale, s); break;
        case 3: launch_mfma<KIND, 3>(f1, f2, out, N, C, CP, H, W, legacy, nchw, scale, s); break;
        case 4: launch_mfma<KIND, 4>(f1, f2, out, N, C, CP, H, W, legacy, nchw, scale, s); break;
        default:
            return set_error(VSC_E_UNSUPPORTED,
                             "vsc_correlation_forward: reduced-precision correlation supports max_displacement <= 4");
    }
    return check_launch("correlation mfma");
}

// NCHW -> channels-last [n][y][x][CP], channels >= C zero.  One thread per
// 4 consecutive pixels and 8-byte channel group (G = 4 BF16/F16 or 8 FP8
// channels): G vector loads of 4 pixels of one channel (coalesced along x
// across the wave), a register transpose, 4 vector stores of G channels.
// Needs H*W % 4 == 0 for the vector loads; k_to_nhwc1 is the general form.
template <typename E>
__global__ void k_to_nhwc4(const E* __restrict__ in, E* __restrict__ out, int C, int CP, long long HW, int N)
{
    constexpr int G = 8 / sizeof(E);
    using V4 = typename std::conditional<sizeof(E) == 2, uint2, uint32_t>::type;   // 4 pixels of one channel
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long quads = HW / 4;
    const int groups = CP / G;
    const long long p4 = e % quads;
    const long long rest = e / quads;
    if (rest >= (long long)groups * N) return;
    const int g = (int)(rest % groups), n = (int)(rest / groups);
    E v[G][4];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const int c = g * G + k;
        V4 x;
        if (c < C) x = *reinterpret_cast<const V4*>(in + ((long long)n * C + c) * HW + p4 * 4);
        else __builtin_memset(&x, 0, sizeof(x));
        __builtin_memcpy(v[k], &x, sizeof(x));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        E w[G];
#pragma unroll
        for (int k = 0; k < G; ++k) w[k] = v[k][j];
        uint2 pk;
        __builtin_memcpy(&pk, w, 8);
        *reinterpret_cast<uint2*>(out + ((long long)n * HW + p4 * 4 + j) * CP + g * G) = pk;
    }
}

template <typename E>
__global__ void k_to_nhwc1(const E* __restrict__ in, E* __restrict__ out, int C, int CP, long long HW, int N)
{
    constexpr int G = 8 / sizeof(E);
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int groups = CP / G;
    const long long p = e % HW;
    const long long rest = e / HW;
    if (rest >= (long long)groups * N) return;
    const int g = (int)(rest % groups), n = (int)(rest / groups);
    E v[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const int c = g * G + k;
        v[k] = c < C ? in[((long long)n * C + c) * HW + p] : (E)0;
    }
    uint2 pk;
    __builtin_memcpy(&pk, v, 8);
    *reinterpret_cast<uint2*>(out + ((long long)n * HW + p) * CP + g * G) = pk;
}

template <typename E>
static void to_nhwc(const void* in, void* out, int C, int CP, long long HW, int N, hipStream_t s)
{
    constexpr int G = 8 / sizeof(E);
    if (HW % 4 == 0) {
        const long long units = (HW / 4) * (CP / G) * N;
        hipLaunchKernelGGL(k_to_nhwc4<E>, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, s,
                           static_cast<const E*>(in), static_cast<E*>(out), C, CP, HW, N);
    } else {
        const long long units = HW * (CP / G) * N;
        hipLaunchKernelGGL(k_to_nhwc1<E>, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, s,
                           static_cast<const E*>(in), static_cast<E*>(out), C, CP, HW, N);
    }
}

size_t corr_mfma_workspace_bytes(int N, int C, int H, int W, int in_dtype)
{
    if (in_dtype != VSC_BF16 && in_dtype != VSC_F16 && in_dtype != VSC_FP8E4M3) return 0;
    const size_t eb = in_dtype == VSC_FP8E4M3 ? 1 : 2;
    const size_t CP = (size_t)(C + CM_KC - 1) / CM_KC * CM_KC;
    const size_t one = ((size_t)N * H * W * CP * eb + 255) / 256 * 256;
    return 2 * one;
}

// widths that are multiples of 4 with aligned operands: the kernel reads the
// NCHW operands itself, with no channels-last copy and no workspace
bool corr_mfma_reads_nchw(const void* f1, const void* f2, int W, int in_dtype)
{
    const uintptr_t align = in_dtype == VSC_FP8E4M3 ? 4 : 8;
    return W % 4 == 0 && reinterpret_cast<uintptr_t>(f1) % align == 0 && reinterpret_cast<uintptr_t>(f2) % align == 0;
}

// workspace: corr_mfma_workspace_bytes() bytes of device memory, or NULL to
// allocate it stream-ordered for this call (unused when the kernel reads the
// NCHW operands directly)
int corr_mfma_dispatch(const void* f1, const void* f2, float* out, int N, int C, int H, int W, int md, int legacy,
                       int in_dtype, const float* chan_scale, void* workspace, hipStream_t s)
{
    if (in_dtype != VSC_BF16 && in_dtype != VSC_F16 && in_dtype != VSC_FP8E4M3)
        return set_error(VSC_E_DTYPE, "vsc_correlation_forward: unsupported input type");
    if (md > 4)
        return set_error(VSC_E_UNSUPPORTED,
                         "vsc_correlation_forward: reduced-precision correlation supports max_displacement <= 4");
    const int eb = in_dtype == VSC_FP8E4M3 ? 1 : 2;
    const int CP = (C + CM_KC - 1) / CM_KC * CM_KC;
    const long long HW = (long long)H * W;
    if (corr_mfma_reads_nchw(f1, f2, W, in_dtype)) {
        if (in_dtype == VSC_BF16) return dispatch_md<CM_BF16>(f1, f2, out, N, C, CP, H, W, md, legacy, true, nullptr, s);
        if (in_dtype == VSC_F16) return dispatch_md<CM_F16>(f1, f2, out, N, C, CP, H, W, md, legacy, true, nullptr, s);
        return dispatch_md<CM_FP8>(f1, f2, out, N, C, CP, H, W, md, legacy, true, chan_scale, s);
    }
    const size_t bytes = corr_mfma_workspace_bytes(N, C, H, W, in_dtype);
    void* ws = workspace;
    int rc;
    if (!ws && (rc = check_hip(hipMallocAsync(&ws, bytes, s), "vsc_correlation_forward workspace"))) return rc;
    void* t1 = ws;
    void* t2 = static_cast<char*>(ws) + bytes / 2;
    if (eb == 2) {
        to_nhwc<uint16_t>(f1, t1, C, CP, HW, N, s);
        to_nhwc<uint16_t>(f2, t2, C, CP, HW, N, s);
    } else {
        to_nhwc<uint8_t>(f1, t1, C, CP, HW, N, s);
        to_nhwc<uint8_t>(f2, t2, C, CP, HW, N, s);
    }
    rc = check_launch("correlation to channels-last");
    if (!rc) {
        if (in_dtype == VSC_BF16) rc = dispatch_md<CM_BF16>(t1, t2, out, N, C, CP, H, W, md, legacy, false, nullptr, s);
        else if (in_dtype == VSC_F16) rc = dispatch_md<CM_F16>(t1, t2, out, N, C, CP, H, W, md, legacy, false, nullptr, s);
        else rc = dispatch_md<CM_FP8>(t1, t2, out, N, C, CP, H, W, md, legacy, false, chan_scale, s);
    }
    if (workspace) return rc;
    const int rc2 = check_hip(hipFreeAsync(ws, s), "vsc_correlation_forward workspace free");
    return rc ? rc : rc2;
}

// ---------------------------------------------------------------------------
// conversions
// ---------------------------------------------------------------------------
// binary32 -> OCP FP8 E4M3 (bias 7, max 448, no infinities), round to nearest
// even, saturating to +-448; NaN -> 0x7F.
__host__ __device__ inline uint8_t f32_to_e4m3(float x)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    const uint32_t sign = (u >> 24) & 0x80u;
    const uint32_t au = u & 0x7FFFFFFFu;
    if (au > 0x7F800000u) return 0x7F;
    const float a = __builtin_bit_cast(float, au);
    uint32_t code;
    if (a >= 448.0f) {
        code = 0x7E;
    } else if (a < 0.015625f) {                       // below 2^-6: subnormal grid of 2^-9
        code = (uint32_t)__builtin_rintf(a * 512.0f);  // exact scaling, RNE; 8 -> min normal 0x08
    } else {
        int e = (int)(au >> 23) - 127;                 // -6 .. 8
        uint32_t mant = au & 0x7FFFFFu;
        uint32_t keep = mant >> 20, rest = mant & 0xFFFFFu;
        if (rest > 0x80000u || (rest == 0x80000u && (keep & 1u))) ++keep;
        if (keep == 8u) { keep = 0; ++e; }
        if (e > 8 || (e == 8 && keep == 7u)) code = 0x7E;
        else code = ((uint32_t)(e + 7) << 3) | keep;
    }
    return (uint8_t)(sign | code);
}

__global__ void k_to_bf16(const float* in, uint16_t* out, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f32_to_bf16(in[i]);
}

__global__ void k_to_f16(const float* in, _Float16* out, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (_Float16)in[i];
}

__global__ void k_to_e4m3(const float* in, uint8_t* out, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f32_to_e4m3(in[i]);
}

// ---------------------------------------------------------------------------
// FP8 pair quantisation (SmoothQuant-style equalisation)
// ---------------------------------------------------------------------------
// per-channel max |f| over N, H, W, as ordered float bits (non-negative floats
// order like their bit patterns), accumulated with atomicMax
__global__ void k_chan_absmax(const float* f, uint32_t* mx, int N, int C, long long HW)
{
    const int c = blockIdx.y, n = blockIdx.z;
    const float* p = f + ((long long)n * C + c) * HW;
    float m = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < HW; i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(p[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&mx[c], __builtin_bit_cast(uint32_t, m));
}

__global__ void k_zero_u32(uint32_t* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// r_c = sqrt(max2_c / max1_c) (1 when either is 0); s1 = max_c(max1_c r_c)/448,
// s2 = max_c(max2_c / r_c)/448; writes r[c], inv scales and scale = s1*s2.
__global__ void k_fp8_scales(const uint32_t* mx1, const uint32_t* mx2, float* r, float* sc, int C)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float m1 = 0.0f, m2 = 0.0f;
    for (int c = 0; c < C; ++c) {
        const float a = __builtin_bit_cast(float, mx1[c]), b = __builtin_bit_cast(float, mx2[c]);
        const float rc = (a > 0.0f && b > 0.0f) ? sqrtf(b / a) : 1.0f;
        r[c] = rc;
        m1 = fmaxf(m1, a * rc);
        m2 = fmaxf(m2, b / rc);
    }
    const float s1 = m1 > 0.0f ? m1 / 448.0f : 1.0f;
    const float s2 = m2 > 0.0f ? m2 / 448.0f : 1.0f;
    sc[0] = s1 * s2;
    sc[1] = s1;
    sc[2] = s2;
}

__global__ void k_fp8_quant(const float* f, uint8_t* q, const float* r, const float* sc, int which, int C,
                            long long HW, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)((i / HW) % C);
    const float v = which == 0 ? (f[i] * r[c]) / sc[1] : (f[i] / r[c]) / sc[2];
    q[i] = f32_to_e4m3(v);
}

}  // namespace vsc

using namespace vsc;

extern "C" size_t vsc_quantize_fp8_pair_workspace_size(int C)
{
    return C < 1 ? 0 : (size_t)(3 * C + 4) * 4;
}

// workspace: vsc_quantize_fp8_pair_workspace_size(C) bytes, or NULL to
// allocate it stream-ordered for this call
static int quantize_fp8_pair(const float* f1, const float* f2, uint8_t* q1, uint8_t* q2, float* scale, int N, int C,
                             int H, int W, void* workspace, hipStream_t s)
{
    if (!f1 || !f2 || !q1 || !q2 || !scale) return set_error(VSC_E_INVALID, "vsc_quantize_fp8_pair: null pointer");
    if (N < 1 || C < 1 || H < 1 || W < 1) return set_error(VSC_E_INVALID, "vsc_quantize_fp8_pair: empty tensor");
    const long long HW = (long long)H * W, n = (long long)N * C * HW;
    // workspace: 2*C maxima + C equalisation factors + 3 scales
    void* ws = workspace;
    int rc;
    if (!ws && (rc = check_hip(hipMallocAsync(&ws, vsc_quantize_fp8_pair_workspace_size(C), s),
                               "vsc_quantize_fp8_pair workspace")))
        return rc;
    uint32_t* mx = static_cast<uint32_t*>(ws);
    float* r = reinterpret_cast<float*>(mx + 2 * C);
    float* sc = r + C;
    hipLaunchKernelGGL(k_zero_u32, dim3((2 * C + 255) / 256), dim3(256), 0, s, mx, 2 * C);
    const unsigned gx = (unsigned)std::min<long long>((HW + 255) / 256, 64);
    hipLaunchKernelGGL(k_chan_absmax, dim3(gx, C, N), dim3(256), 0, s, f1, mx, N, C, HW);
    hipLaunchKernelGGL(k_chan_absmax, dim3(gx, C, N), dim3(256), 0, s, f2, mx + C, N, C, HW);
    hipLaunchKernelGGL(k_fp8_scales, dim3(1), dim3(64), 0, s, mx, mx + C, r, sc, C);
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_fp8_quant, dim3(g), dim3(256), 0, s, f1, q1, r, sc, 0, C, HW, n);
    hipLaunchKernelGGL(k_fp8_quant, dim3(g), dim3(256), 0, s, f2, q2, r, sc, 1, C, HW, n);
    rc = check_launch("vsc_quantize_fp8_pair");
    if (!rc) rc = check_hip(hipMemcpyAsync(scale, sc, sizeof(float), hipMemcpyDeviceToDevice, s), "vsc_quantize_fp8_pair");
    if (workspace) return rc;
    const int rc2 = check_hip(hipFreeAsync(ws, s), "vsc_quantize_fp8_pair workspace free");
    return rc ? rc : rc2;
}

extern "C" int vsc_quantize_fp8_pair(const float* f1, const float* f2, uint8_t* q1, uint8_t* q2, float* scale, int N,
                                     int C, int H, int W, vsc_hip_stream stream)
{
    return quantize_fp8_pair(f1, f2, q1, q2, scale, N, C, H, W, nullptr, as_stream(stream));
}

extern "C" int vsc_quantize_fp8_pair_ws(const float* f1, const float* f2, uint8_t* q1, uint8_t* q2, float* scale,
                                        int N, int C, int H, int W, void* workspace, size_t workspace_bytes,
                                        vsc_hip_stream stream)
{
    if (!workspace || workspace_bytes < vsc_quantize_fp8_pair_workspace_size(C))
        return set_error(VSC_E_INVALID, "vsc_quantize_fp8_pair_ws: workspace missing or smaller than "
                                        "vsc_quantize_fp8_pair_workspace_size()");
    return quantize_fp8_pair(f1, f2, q1, q2, scale, N, C, H, W, workspace, as_stream(stream));
}

extern "C" int vsc_convert_f32(const float* in, void* out, size_t count, int out_dtype, vsc_hip_stream stream)
{
    if (!in || !out) return set_error(VSC_E_INVALID, "vsc_convert_f32: null pointer");
    if (count == 0) return VSC_OK;
    hipStream_t s = as_stream(stream);
    const long long n = (long long)count;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (out_dtype == VSC_BF16)
        hipLaunchKernelGGL(k_to_bf16, dim3(g), dim3(256), 0, s, in, static_cast<uint16_t*>(out), n);
    else if (out_dtype == VSC_F16)
        hipLaunchKernelGGL(k_to_f16, dim3(g), dim3(256), 0, s, in, static_cast<_Float16*>(out), n);
    else if (out_dtype == VSC_FP8E4M3)
        hipLaunchKernelGGL(k_to_e4m3, dim3(g), dim3(256), 0, s, in, static_cast<uint8_t*>(out), n);
    else
        return set_error(VSC_E_DTYPE, "vsc_convert_f32: out_dtype must be BF16, F16 or FP8E4M3");
    return check_launch("vsc_convert_f32");
}
