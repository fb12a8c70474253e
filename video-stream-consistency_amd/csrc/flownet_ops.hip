// Channels-last (NHWC) kernels of the flow network besides the convolutions:
// the cost volume, the backward warp (with the flow upsampling, the
// ConvTranspose2d moduleUpflow, computed inside it), input preprocessing and
// the final flow resize.  FP32 arithmetic throughout;
// tensors are stored as FP32 or BF16 (the element kind of the network form).
#include "vsc_internal.h"
#include "flownet_internal.h"

namespace vsc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 fo_bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float fo_f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t fo_u32x4;

// ---------------------------------------------------------------------------
// Cost volume (Correlation, md = 4, non-legacy: pwcnet.py:337-355 with the
// op of correlation.cc:203-275) on MFMA, fused with LeakyReLU(0.1)
// (pwcnet.py:568,577), written into the decoder's concatenation buffer as
// channels [0, 81) of each pixel, with channels [81, 88) zeroed; the f1
// features it stages are copied beside them (channel f1_c0 on).
//
// For output row h and 16 pixels w0..w0+15, G[i][j] = sum_c f1[h][w0+i][c] *
// f2[h+dy][w0-4+j][c] (j < 24, two 16x16 MFMA tiles) holds every horizontal
// displacement on its diagonals.  Block: 4 rows x 32 pixels, 8 waves (row,
// 16-pixel half); channel chunks of KC stream through LDS as straight 16-byte
// copies (the tensors are already channels-last).
// ---------------------------------------------------------------------------
template <int EK> struct CorrK;
template <> struct CorrK<EK_F32> { static constexpr int KC = 16; };
template <> struct CorrK<EK_BF16> { static constexpr int KC = 32; };

struct CorrArgs {
    const void* f1[FLOWNET_MAX_PAIRS];       // per batch item, channel 0 of pixel 0
    const void* f2[FLOWNET_MAX_PAIRS];
    int cs1, cs2, C, H, W;
    void* out; int ocs;                      // out: channel 0 of pixel 0 of batch 0
    int f1_c0;                               // >= 0: f1's C channels are also copied to out at this channel
};

template <int EK>
__device__ inline fo_f32x4 corr_mma(const fo_u32x4& a, const fo_u32x4& b, fo_f32x4 acc)
{
    if constexpr (EK == EK_BF16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(fo_bf16x8, a),
                                                       __builtin_bit_cast(fo_bf16x8, b), acc, 0, 0, 0);
    } else {
        const fo_f32x4 fa = __builtin_bit_cast(fo_f32x4, a), fb = __builtin_bit_cast(fo_f32x4, b);
#pragma unroll
        for (int t = 0; t < 4; ++t) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[t], fb[t], acc, 0, 0, 0);
        return acc;
    }
}

// BFMM (FP32 storage only): the FP32 features are rounded to BF16 as they are
// staged and multiplied on the BF16 MFMA (FP32 accumulation): the BF16 cost
// volume of an otherwise FP32 network.
template <int EK, bool BFMM>
__global__ __launch_bounds__(512) void k_corr_nhwc(CorrArgs a)
{
    static_assert(!BFMM || EK == EK_F32, "BF16 products of FP32 storage only");
    using E = Elem<EK>;
    using T = typename E::T;
    constexpr int MK = BFMM ? EK_BF16 : EK;             // element kind of the LDS operands and the MFMA
    using M = typename Elem<MK>::T;
    constexpr int VEC = E::VEC, KC = CorrK<MK>::KC, KV = KC / VEC;
    constexpr int MVEC = Elem<MK>::VEC, ROW = KC + MVEC;
    constexpr int RR = 4, RP = 32, HALO = 4, BR = RR + 2 * HALO, BP = RP + 2 * HALO;
    constexpr int NA = RR * RP * KV, NB = BR * BP * KV, NU = NA + NB, RU = (NU + 511) / 512;
    constexpr int OPX = 88 / VEC;                       // output vectors per pixel (81 + 7 zeros)
    constexpr int SROW = 89;                            // stage floats per pixel (odd: fewer conflicts)
    constexpr int OPER_BYTES = (RR * RP + BR * BP) * ROW * (int)sizeof(M);
    constexpr int STAGE_BYTES = RR * RP * SROW * 4;
    constexpr int LDS = OPER_BYTES > STAGE_BYTES ? OPER_BYTES : STAGE_BYTES;
    __shared__ __attribute__((aligned(16))) unsigned char lds[LDS];
    M* sA = reinterpret_cast<M*>(lds);
    M* sB = sA + RR * RP * ROW;
    float* stage = reinterpret_cast<float*>(lds);

    const int n = blockIdx.z, h0 = blockIdx.y * RR, w0 = blockIdx.x * RP;
    const T* f1 = static_cast<const T*>(a.f1[n]);
    const T* f2 = static_cast<const T*>(a.f2[n]);
    const int tid = threadIdx.x, q = tid / 64, lane = tid % 64;
    const int qr = q >> 1, seg = q & 1, l16 = lane % 16, lg = lane / 16;

    fo_f32x4 acc[9][2];
#pragma unroll
    for (int d = 0; d < 9; ++d)
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[d][t] = fo_f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    fo_u32x4 rg[RU];
    auto fetch = [&](int c0) {
#pragma unroll
        for (int k = 0; k < RU; ++k) {
            const int u = tid + 512 * k;
            fo_u32x4 v = fo_u32x4{0u, 0u, 0u, 0u};
            if (u < NA) {
                const int p = u / KV, cv = u % KV, r = p / RP, x = p % RP;
                const int gh = h0 + r, gw = w0 + x, ch = c0 + cv * VEC;
                if (gh < a.H && gw < a.W && ch < a.C) {
                    v = *reinterpret_cast<const fo_u32x4*>(f1 + ((long long)gh * a.W + gw) * a.cs1 + ch);
                    // the decoder's concatenation takes f1 beside the cost volume
                    // (pwcnet.py:578): every block stages its own pixels' f1 once
                    if (a.f1_c0 >= 0)
                        *reinterpret_cast<fo_u32x4*>(static_cast<T*>(a.out) +
                                                     (((long long)n * a.H + gh) * a.W + gw) * a.ocs + a.f1_c0 + ch) = v;
                }
            } else if (u < NU) {
                const int ub = u - NA, p = ub / KV, cv = ub % KV, r = p / BP, x = p % BP;
                const int gh = h0 - HALO + r, gw = w0 - HALO + x, ch = c0 + cv * VEC;
                if (gh >= 0 && gh < a.H && gw >= 0 && gw < a.W && ch < a.C)
                    v = *reinterpret_cast<const fo_u32x4*>(f2 + ((long long)gh * a.W + gw) * a.cs2 + ch);
            }
            rg[k] = v;
        }
    };

    fetch(0);
    for (int c0 = 0; c0 < a.C; c0 += KC) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RU; ++k) {
            const int u = tid + 512 * k;
            if (u >= NU) continue;
            M* dst = u < NA ? &sA[(u / KV) * ROW + (u % KV) * VEC] : &sB[((u - NA) / KV) * ROW + ((u - NA) % KV) * VEC];
            if constexpr (BFMM) {
                // 4 FP32 channels -> 4 BF16 (round to nearest even)
                const fo_f32x4 f = __builtin_bit_cast(fo_f32x4, rg[k]);
                using B = Elem<EK_BF16>;
                *reinterpret_cast<uint2*>(dst) =
                    make_uint2((uint32_t)B::st(f[0]) | ((uint32_t)B::st(f[1]) << 16),
                               (uint32_t)B::st(f[2]) | ((uint32_t)B::st(f[3]) << 16));
            } else {
                *reinterpret_cast<fo_u32x4*>(dst) = rg[k];
            }
        }
        __syncthreads();
        if (c0 + KC < a.C) fetch(c0 + KC);
        const fo_u32x4 fa = *reinterpret_cast<const fo_u32x4*>(&sA[(qr * RP + 16 * seg + l16) * ROW + lg * (KC / 4)]);
#pragma unroll
        for (int d = 0; d < 9; ++d) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const fo_u32x4 fb = *reinterpret_cast<const fo_u32x4*>(
                    &sB[((qr + d) * BP + 16 * seg + 8 * t + l16) * ROW + lg * (KC / 4)]);
                acc[d][t] = corr_mma<MK>(fa, fb, acc[d][t]);
            }
        }
    }

    // diagonals -> stage[row*32 + px][plane]
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 9; ++d)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 4 * lg + r, j = l16 + 8 * t, dx = j - i - HALO;
                if (dx >= -4 && dx <= 4 && !(t == 1 && j < 16))
                    stage[(qr * RP + 16 * seg + i) * SROW + d * 9 + dx + 4] = acc[d][t][r];
            }
    __syncthreads();
    T* out = static_cast<T*>(a.out) + (long long)n * a.H * a.W * a.ocs;
    for (int e = tid; e < RR * RP * OPX; e += 512) {
        const int p = e / OPX, v = e % OPX, r = p / RP, x = p % RP;
        const int h = h0 + r, w = w0 + x;
        if (h >= a.H || w >= a.W) continue;
        float o[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const int pl = v * VEC + k;
            const float g = pl < 81 ? stage[p * SROW + pl] : 0.0f;
            o[k] = g > 0.0f ? g : g * 0.1f;
        }
        T* dst = out + ((long long)h * a.W + w) * a.ocs + v * VEC;
        if constexpr (EK == EK_F32) {
            *reinterpret_cast<fo_f32x4*>(dst) = fo_f32x4{o[0], o[1], o[2], o[3]};
        } else {
            fo_u32x4 pk;
#pragma unroll
            for (int k = 0; k < 4; ++k) pk[k] = (uint32_t)E::st(o[2 * k]) | ((uint32_t)E::st(o[2 * k + 1]) << 16);
            *reinterpret_cast<fo_u32x4*>(dst) = pk;
        }
    }
}

int launch_corr_nhwc(int ek, const void* const* f1, const void* const* f2, int cs1, int cs2, int C, int N,
                     int H, int W, void* out, int ocs, int f1_c0, hipStream_t s, bool bf16_products)
{
    if (N < 1 || N > FLOWNET_MAX_PAIRS) return set_error(VSC_E_INVALID, "corr nhwc: batch of 1 to 8 items");
    CorrArgs a{};
    for (int i = 0; i < N; ++i) { a.f1[i] = f1[i]; a.f2[i] = f2[i]; }
    a.cs1 = cs1; a.cs2 = cs2; a.C = C; a.H = H; a.W = W; a.out = out; a.ocs = ocs;
    a.f1_c0 = f1_c0;
    if (f1_c0 >= 0 && f1_c0 + C > ocs) return set_error(VSC_E_INVALID, "corr nhwc: f1 copy outside the pixel");
    dim3 grid((W + 31) / 32, (H + 3) / 4, N);
    if (ek == EK_BF16) hipLaunchKernelGGL((k_corr_nhwc<EK_BF16, false>), grid, dim3(512), 0, s, a);
    else if (bf16_products) hipLaunchKernelGGL((k_corr_nhwc<EK_F32, true>), grid, dim3(512), 0, s, a);
    else hipLaunchKernelGGL((k_corr_nhwc<EK_F32, false>), grid, dim3(512), 0, s, a);
    return check_launch("corr nhwc");
}

// ---------------------------------------------------------------------------
// moduleUpflow: ConvTranspose2d(2, 2, kernel 4, stride 2, padding 1)
// (pwcnet.py:472): out[co][oy][ox] = b[co] + sum in[ci][iy][ix] * w[ci][co][ky][kx]
// over oy = 2*iy - 1 + ky, ox = 2*ix - 1 + kx, for one output pixel of item n
// (FP32 flow in, 2 channels).
// ---------------------------------------------------------------------------
__device__ inline void upflow_px(const float* __restrict__ in, int Hi, int Wi, const float* __restrict__ w,
                                 const float* __restrict__ b, int n, int oy, int ox, float acc[2])
{
    acc[0] = b[0];
    acc[1] = b[1];
    for (int ky = 0; ky < 4; ++ky) {
        const int ty = oy + 1 - ky;
        if (ty < 0 || (ty & 1)) continue;
        const int iy = ty >> 1;
        if (iy >= Hi) continue;
        for (int kx = 0; kx < 4; ++kx) {
            const int tx = ox + 1 - kx;
            if (tx < 0 || (tx & 1)) continue;
            const int ix = tx >> 1;
            if (ix >= Wi) continue;
            const float* s = in + (((long long)n * Hi + iy) * Wi + ix) * 2;
            for (int co = 0; co < 2; ++co)
                acc[co] = acc[co] + (s[0] * w[(0 * 2 + co) * 16 + ky * 4 + kx] + s[1] * w[(1 * 2 + co) * 16 + ky * 4 + kx]);
        }
    }
}

// ---------------------------------------------------------------------------
// Backward warp of f2 by flow * scale (Backward(): pwcnet.py:575-576 through
// the Warp op, warp.cc:71-134): bilinear taps, per-corner bounds masks, the
// sample kept only where the in-bounds weight mass passes 0.999 (the FP32
// rule of the op kernel), zero otherwise.  Thread = (pixel, 16-byte group).
// ---------------------------------------------------------------------------
struct WarpArgs {
    const void* src[FLOWNET_MAX_PAIRS];     // f2 features of each batch item
    int scs;
    const float* flow; int fcs;             // FP32 flow, channels (x, y) at flow + pix*fcs
    float scale;
    void* out; int ocs;                     // warped features
    int C, H, W, groups;
    // with up_in: the flow is the upflow of up_in (H/2 x W/2, computed here
    // per pixel), and channel group 0 also writes it to up_out (FP32, 2
    // channels) and, in the element kind, to cat at channel cc0
    const float* up_in; const float* up_w; const float* up_b;
    float* up_out; void* cat; int ccs, cc0;
};

template <int EK>
__global__ __launch_bounds__(256) void k_warp_nhwc(WarpArgs a)
{
    using E = Elem<EK>;
    using T = typename E::T;
    constexpr int VEC = E::VEC;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long npx = (long long)a.H * a.W;
    const int n = blockIdx.y;
    if (e >= npx * a.groups) return;
    const int g = (int)(e % a.groups);
    const long long p = e / a.groups;
    const int x = (int)(p % a.W), y = (int)(p / a.W);
    float fv[2];
    if (a.up_in) {
        upflow_px(a.up_in, a.H / 2, a.W / 2, a.up_w, a.up_b, n, y, x, fv);
        if (g == 0) {
            float* fo = a.up_out + ((long long)n * npx + p) * 2;
            fo[0] = fv[0];
            fo[1] = fv[1];
            T* c = static_cast<T*>(a.cat) + ((long long)n * npx + p) * a.ccs + a.cc0;
            c[0] = E::st(fv[0]);
            c[1] = E::st(fv[1]);
        }
    } else {
        const float* fl = a.flow + ((long long)n * npx + p) * a.fcs;
        fv[0] = fl[0];
        fv[1] = fl[1];
    }
    const float fx = fv[0] * a.scale, fy = fv[1] * a.scale;
    const float xf = (float)x + fx, yf = (float)y + fy;
    const float xL = floorf(xf), yT = floorf(yf);
    const float al = xf - xL, be = yf - yT;
    const float xR = xL + 1.0f, yB = yT + 1.0f;
    const bool mL = 0.0f <= xL && xL <= (float)(a.W - 1), mR = 0.0f <= xR && xR <= (float)(a.W - 1);
    const bool mT = 0.0f <= yT && yT <= (float)(a.H - 1), mB = 0.0f <= yB && yB <= (float)(a.H - 1);
    const float wTL = (1.0f - al) * (1.0f - be), wTR = al * (1.0f - be), wBL = (1.0f - al) * be, wBR = al * be;
    float mask = 0.0f;
    mask = mask + (mT && mL ? wTL : 0.0f);
    mask = mask + (mT && mR ? wTR : 0.0f);
    mask = mask + (mB && mL ? wBL : 0.0f);
    mask = mask + (mB && mR ? wBR : 0.0f);
    float v[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) v[k] = 0.0f;
    if (mask >= 0.999f) {
        const T* src = static_cast<const T*>(a.src[n]) + g * VEC;
        // the four corners are loaded before any is used, from addresses
        // clamped into the image (every load valid); a corner outside it is
        // still left out of the sum
        const int iL = min(max((int)xL, 0), a.W - 1), iR = min(max((int)xR, 0), a.W - 1);
        const int iT = min(max((int)yT, 0), a.H - 1), iB = min(max((int)yB, 0), a.H - 1);
        auto ld = [&](int ix, int iy) {
            return *reinterpret_cast<const fo_u32x4*>(src + ((long long)iy * a.W + ix) * a.scs);
        };
        const fo_u32x4 rTL = ld(iL, iT), rTR = ld(iR, iT), rBL = ld(iL, iB), rBR = ld(iR, iB);
        auto acc = [&](bool m, const fo_u32x4& raw, float w) {
            if (!m) return;
            if constexpr (EK == EK_F32) {
                const fo_f32x4 f = __builtin_bit_cast(fo_f32x4, raw);
#pragma unroll
                for (int k = 0; k < 4; ++k) v[k] = v[k] + w * f[k];
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    v[2 * k] = v[2 * k] + w * E::ld((T)(raw[k] & 0xFFFFu));
                    v[2 * k + 1] = v[2 * k + 1] + w * E::ld((T)(raw[k] >> 16));
                }
            }
        };
        acc(mT && mL, rTL, wTL);
        acc(mT && mR, rTR, wTR);
        acc(mB && mL, rBL, wBL);
        acc(mB && mR, rBR, wBR);
    }
    T* dst = static_cast<T*>(a.out) + ((long long)n * npx + p) * a.ocs + g * VEC;
    if constexpr (EK == EK_F32) {
        *reinterpret_cast<fo_f32x4*>(dst) = fo_f32x4{v[0], v[1], v[2], v[3]};
    } else {
        fo_u32x4 pk;
#pragma unroll
        for (int k = 0; k < 4; ++k) pk[k] = (uint32_t)E::st(v[2 * k]) | ((uint32_t)E::st(v[2 * k + 1]) << 16);
        *reinterpret_cast<fo_u32x4*>(dst) = pk;
    }
}

int launch_warp_nhwc(int ek, const void* const* src, int scs, const float* flow, int fcs, float scale, void* out,
                     int ocs, int C, int N, int H, int W, hipStream_t s, const UpflowArgs* up)
{
    if (N < 1 || N > FLOWNET_MAX_PAIRS) return set_error(VSC_E_INVALID, "warp nhwc: batch of 1 to 8 items");
    WarpArgs a{};
    for (int i = 0; i < N; ++i) a.src[i] = src[i];
    a.scs = scs; a.flow = flow; a.fcs = fcs; a.scale = scale; a.out = out; a.ocs = ocs;
    a.C = C; a.H = H; a.W = W;
    if (up) {
        if ((H & 1) || (W & 1)) return set_error(VSC_E_INVALID, "warp nhwc: upflow needs even sizes");
        a.up_in = up->in; a.up_w = up->w; a.up_b = up->b; a.up_out = up->out;
        a.cat = up->cat; a.ccs = up->ccs; a.cc0 = up->cc0;
    }
    const int vec = ek == EK_F32 ? 4 : 8;
    a.groups = (C + vec - 1) / vec;
    const long long units = (long long)H * W * a.groups;
    dim3 grid((unsigned)((units + 255) / 256), N);
    if (ek == EK_BF16) hipLaunchKernelGGL(k_warp_nhwc<EK_BF16>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_warp_nhwc<EK_F32>, grid, dim3(256), 0, s, a);
    return check_launch("warp nhwc");
}

// ---------------------------------------------------------------------------
// Input preprocessing (convert_onnx.py:63-71 with FLOWDOWNSCALE,
// flowmodel.cpp:126-133): RGBA8 frame -> optional nearest downscale to
// (Wd, Hd) (pixel centres: src = floor((dst + 0.5) * W / Wd)) -> /255 ->
// bilinear resize (align_corners = False) to (Wp, Hp), multiples of 64 ->
// NHWC with VEC channels (R, G, B, zeros).
// ---------------------------------------------------------------------------
__device__ inline void resize_tap(int dst, int in, int out, int& i0, int& i1, float& lam)
{
    const float scale = (float)in / (float)out;
    float src = ((float)dst + 0.5f) * scale - 0.5f;
    if (src < 0.0f) src = 0.0f;
    i0 = (int)src;
    if (i0 > in - 1) i0 = in - 1;
    i1 = i0 + (i0 < in - 1 ? 1 : 0);
    lam = src - (float)i0;
}

template <int EK>
__global__ void k_preprocess(const uint8_t* __restrict__ rgba, int W, int H, int Wd, int Hd, int Wp, int Hp,
                             void* out, int ocs)
{
    using E = Elem<EK>;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= Wp || y >= Hp) return;
    int x0, x1, y0, y1;
    float lx, ly;
    resize_tap(x, Wd, Wp, x0, x1, lx);
    resize_tap(y, Hd, Hp, y0, y1, ly);
    auto src = [&](int xd, int yd, int c) {
        const int xs = Wd == W ? xd : min(W - 1, (int)(((float)xd + 0.5f) * ((float)W / (float)Wd)));
        const int ys = Hd == H ? yd : min(H - 1, (int)(((float)yd + 0.5f) * ((float)H / (float)Hd)));
        return (float)rgba[((long long)ys * W + xs) * 4 + c] / 255.0f;
    };
    typename E::T* o = static_cast<typename E::T*>(out) + ((long long)y * Wp + x) * ocs;
    for (int c = 0; c < 3; ++c) {
        const float v = (1.0f - ly) * ((1.0f - lx) * src(x0, y0, c) + lx * src(x1, y0, c)) +
                        ly * ((1.0f - lx) * src(x0, y1, c) + lx * src(x1, y1, c));
        o[c] = E::st(v);
    }
    for (int c = 3; c < ocs; ++c) o[c] = E::st(0.0f);
}

int launch_preprocess(int ek, const uint8_t* rgba, int W, int H, int Wd, int Hd, int Wp, int Hp, void* out, int ocs,
                      hipStream_t s)
{
    dim3 grid((Wp + 255) / 256, Hp);
    if (ek == EK_BF16) hipLaunchKernelGGL(k_preprocess<EK_BF16>, grid, dim3(256), 0, s, rgba, W, H, Wd, Hd, Wp, Hp, out, ocs);
    else hipLaunchKernelGGL(k_preprocess<EK_F32>, grid, dim3(256), 0, s, rgba, W, H, Wd, Hd, Wp, Hp, out, ocs);
    return check_launch("preprocess");
}

// ---------------------------------------------------------------------------
// Final flow (pwcnet.py:1003 then convert_onnx.py:22-34,80-88): the level
// flow (h x w, FP32) times 20, bilinear x8 with align_corners = True to
// (Hp, Wp), then bilinear (align_corners = False) to the frame (Hd, Wd), with
// channel 0 scaled by Hd/Hp and channel 1 by Wd/Wp (the wrapper's order),
// written HWC with out_ch (2 or 3; the third is 0) channels.
// ---------------------------------------------------------------------------
struct OutPair { float* p[FLOWNET_MAX_PAIRS]; };

// blockIdx.z: the pair (flow item) of the batch
__global__ void k_flow_final(const float* __restrict__ flows, int h, int w, int Hp, int Wp, OutPair outs, int Wd,
                             int Hd, int out_ch, int swap_scale)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= Wd || y >= Hd) return;
    const float* fl = flows + (long long)blockIdx.z * h * w * 2;
    float* out = outs.p[blockIdx.z];
    int x0, x1, y0, y1;
    float lx, ly;
    resize_tap(x, Wp, Wd, x0, x1, lx);
    resize_tap(y, Hp, Hd, y0, y1, ly);
    // align_corners = True upsample of the level flow, evaluated at (X, Y) of the Hp x Wp grid
    const float sx = w > 1 ? (float)(w - 1) / (float)(Wp - 1) : 0.0f;
    const float sy = h > 1 ? (float)(h - 1) / (float)(Hp - 1) : 0.0f;
    auto up = [&](int X, int Y, int c) {
        const float fxs = (float)X * sx, fys = (float)Y * sy;
        const int i0 = min((int)fxs, w - 1), j0 = min((int)fys, h - 1);
        const int i1 = min(i0 + 1, w - 1), j1 = min(j0 + 1, h - 1);
        const float ax = fxs - (float)i0, ay = fys - (float)j0;
        const float v00 = fl[((long long)j0 * w + i0) * 2 + c] * 20.0f, v10 = fl[((long long)j0 * w + i1) * 2 + c] * 20.0f;
        const float v01 = fl[((long long)j1 * w + i0) * 2 + c] * 20.0f, v11 = fl[((long long)j1 * w + i1) * 2 + c] * 20.0f;
        return (1.0f - ay) * ((1.0f - ax) * v00 + ax * v10) + ay * ((1.0f - ax) * v01 + ax * v11);
    };
    const float su = swap_scale ? (float)Wd / (float)Wp : (float)Hd / (float)Hp;
    const float sv = swap_scale ? (float)Hd / (float)Hp : (float)Wd / (float)Wp;
    float* o = out + ((long long)y * Wd + x) * out_ch;
    for (int c = 0; c < 2; ++c) {
        const float v = (1.0f - ly) * ((1.0f - lx) * up(x0, y0, c) + lx * up(x1, y0, c)) +
                        ly * ((1.0f - lx) * up(x0, y1, c) + lx * up(x1, y1, c));
        o[c] = v * (c == 0 ? su : sv);
    }
    if (out_ch == 3) o[2] = 0.0f;
}

int launch_flow_final(const float* fl, int N, int h, int w, int Hp, int Wp, float* const* out, int Wd, int Hd,
                      int out_ch, int swap_scale, hipStream_t s)
{
    if (N < 1 || N > FLOWNET_MAX_PAIRS) return set_error(VSC_E_INVALID, "flow final: 1 to 8 flows");
    OutPair o{};
    for (int i = 0; i < N; ++i) o.p[i] = out[i];
    dim3 grid((Wd + 255) / 256, Hd, N);
    hipLaunchKernelGGL(k_flow_final, grid, dim3(256), 0, s, fl, h, w, Hp, Wp, o, Wd, Hd, out_ch, swap_scale);
    return check_launch("flow final");
}

}  // namespace vsc
