// 3x3 convolution over channels-last (NHWC) tensors on MFMA (gfx950): the
// building block of the flow network (Conv2d layers of pwcnet.py:186-241,
// 511-535, 614-644; padding = dilation, stride 1 or 2) and of the learned flow
// upsampler.  FP32 (v_mfma_f32_16x16x4_f32), BF16 and FP16
// (v_mfma_f32_16x16x32_{bf16,f16}) operands, FP32 accumulation, bias,
// LeakyReLU(0.1) / ReLU and a residual add fused into the epilogue.
//
// GEMM form: D[cout][pixel] = sum_{tap, cin} W[cout][tap][cin] * X[tap(pixel)][cin].
// A block of NW waves owns TH = NW*MR output rows x 16 output columns and NT
// output channels; wave q computes rows q*MR .. q*MR+MR-1, i.e. MR pixel tiles
// of 16 against NT/16 channel tiles.  Input channels stream through LDS in chunks of
// KC (32 for 16-bit, 16 for FP32; 8 or 16 for 16-bit layers that narrow, whose
// MFMA K-steps then hold 4 or 2 taps): the input patch of the block (TH+2 x 18
// pixels at stride 1, 2*TH+1 x 33 at stride 2) and the weights of the chunk
// for all nine taps; the nine taps then read shifted windows of the patch.
// Dilated layers (dilation d > 1) are tiled on the d x d sub-grids of equal
// (y mod d, x mod d): on a sub-grid the dilated conv is a dense 3x3 conv, so
// the patch stays TH+2 x 18 whatever d is.
// Operand roles: A = weights (row = output channel), B = pixels (column), so
// each lane ends with 4 consecutive output channels of one pixel, stored as one
// 8- or 16-byte vector in NHWC.
#include "vsc_internal.h"

#include <algorithm>

namespace vsc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 cv_bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 cv_f16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float cv_f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t cv_u32x4;

template <int EK> struct ConvK;
template <> struct ConvK<EK_F32> { static constexpr int KC = 16; };
template <> struct ConvK<EK_BF16> { static constexpr int KC = 32; };
template <> struct ConvK<EK_F16> { static constexpr int KC = 32; };

struct ConvArgs {
    const void* in; int in_cs, in_c0;
    int N, H, W, Cin;
    const void* wt; const float* bias;
    void* out; int out_cs, out_c0;
    void* out2; int out2_cs, out2_c0;   // optional second copy of the result, in the input's element kind (Cout <= 16)
    const void* add; int add_cs, add_c0;
    const float* gate;
    int Ho, Wo, Cout;
    int pad, pstep, act;
    int tiles_x, tiles_y;
    // split over input channels: blockIdx.y = part * ctiles + channel tile; part p
    // covers chunks [p*cpp, (p+1)*cpp) and writes raw FP32 sums to
    // partial + p*part_stride (k_conv_finalize adds them in part order)
    int ctiles, cpp;
    float* partial;
    long long part_stride;
};

template <int EK>
__device__ inline cv_f32x4 conv_mfma(const cv_u32x4& a, const cv_u32x4& b, cv_f32x4 acc)
{
    if constexpr (EK == EK_BF16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(cv_bf16x8, a),
                                                       __builtin_bit_cast(cv_bf16x8, b), acc, 0, 0, 0);
    } else if constexpr (EK == EK_F16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(cv_f16x8, a),
                                                      __builtin_bit_cast(cv_f16x8, b), acc, 0, 0, 0);
    } else {
        // lane group g holds channels 4g .. 4g+3 of the 16-channel chunk; step t
        // feeds channel 4g+t as the instruction's k = g (same map for A and B)
        const cv_f32x4 fa = __builtin_bit_cast(cv_f32x4, a), fb = __builtin_bit_cast(cv_f32x4, b);
#pragma unroll
        for (int t = 0; t < 4; ++t) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[t], fb[t], acc, 0, 0, 0);
        return acc;
    }
}

template <int EKO>
__device__ inline void conv_store4(void* out, long long off, const float v[4], int nvalid)
{
    using E = Elem<EKO>;
    typename E::T* p = static_cast<typename E::T*>(out) + off;
    if (nvalid >= 4) {
        if constexpr (EKO == EK_F32) {
            *reinterpret_cast<cv_f32x4*>(p) = cv_f32x4{v[0], v[1], v[2], v[3]};
        } else {
            const uint32_t lo = (uint32_t)E::st(v[0]) | ((uint32_t)E::st(v[1]) << 16);
            const uint32_t hi = (uint32_t)E::st(v[2]) | ((uint32_t)E::st(v[3]) << 16);
            *reinterpret_cast<uint2*>(p) = make_uint2(lo, hi);
        }
    } else {
        for (int r = 0; r < nvalid; ++r) p[r] = E::st(v[r]);
    }
}

template <int EKO>
__device__ inline void conv_load4(const void* src, long long off, float v[4], int nvalid)
{
    using E = Elem<EKO>;
    const typename E::T* p = static_cast<const typename E::T*>(src) + off;
    for (int r = 0; r < 4; ++r) v[r] += (r < nvalid) ? E::ld(p[r]) : 0.0f;
}

template <int EK, int EKO, int S, int NT, int MR, int NW, int KCX>
__global__ __launch_bounds__(64 * NW) void k_conv3x3(ConvArgs a)
{
    using E = Elem<EK>;
    using T = typename E::T;
    constexpr int NTH = 64 * NW;
    constexpr int VEC = E::VEC, KC = KCX, KV = KC / VEC;
    constexpr int TH = NW * MR, TW = 16;
    constexpr int PR = (TH - 1) * S + 3, PC = (TW - 1) * S + 3;
    // LDS pixel stride: the chunk alone for FP32 (64 bytes: the plain form
    // 1.3% faster than with a 16-byte pad), the chunk and two 16-byte pads for
    // 16-bit operands (96 bytes for 32-channel chunks: 5% faster than 80 on the
    // 1080p level-3 layers, 64 and 112 no better, 144 slower)
    constexpr int ROW = KC + (EK == EK_F32 ? 0 : 2) * VEC;
    constexpr int NP = PR * PC * KV, NWU = NT * 9 * KV;
    constexpr int RP = (NP + NTH - 1) / NTH, RW = (NWU + NTH - 1) / NTH;
    __shared__ __attribute__((aligned(16))) T sP[PR * PC * ROW];
    __shared__ __attribute__((aligned(16))) T sW[NT * 9 * ROW];

    const int tid = threadIdx.x, q = tid / 64, lane = tid % 64;
    const int l16 = lane % 16, lg = lane / 16;
    const int n = blockIdx.z, kp = blockIdx.y / a.ctiles, nt0 = (blockIdx.y % a.ctiles) * NT;
    const int cbeg = kp * a.cpp * KC, cend = min(a.Cin, (kp + 1) * a.cpp * KC);
    int t = blockIdx.x;
    const int tx = t % a.tiles_x; t /= a.tiles_x;
    const int ty = t % a.tiles_y; t /= a.tiles_y;
    const int ps = a.pstep, py = t / ps, px = t % ps;
    const int base_y = py - a.pad + ps * ty * TH * S, base_x = px - a.pad + ps * tx * TW * S;
    const T* in = static_cast<const T*>(a.in) + (long long)n * a.H * a.W * a.in_cs + a.in_c0;
    const T* wt = static_cast<const T*>(a.wt);

    cv_f32x4 acc[NT / 16][MR];
#pragma unroll
    for (int i = 0; i < NT / 16; ++i)
#pragma unroll
        for (int m = 0; m < MR; ++m) acc[i][m] = cv_f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // input channel chunks are staged through registers DEPTH chunks ahead of
    // the chunk in LDS, so the global loads of a chunk overlap the MFMAs of the
    // DEPTH chunks before it (deeper staging measured slower: 2 and 3 chunks
    // ahead cost the 1080p BF16 network 3% and 18%)
    constexpr int DEPTH = 1;
    cv_u32x4 rp[DEPTH][RP], rw[DEPTH][RW];
    auto fetch = [&](int c0, cv_u32x4 (&P)[RP], cv_u32x4 (&Wr)[RW]) {
#pragma unroll
        for (int k = 0; k < RP; ++k) {
            const int u = tid + NTH * k;
            const int p = u / KV, cv = u % KV;
            const int r = p / PC, c = p % PC;
            const int iy = base_y + ps * r, ix = base_x + ps * c, ch = c0 + cv * VEC;
            cv_u32x4 v = cv_u32x4{0u, 0u, 0u, 0u};
            if (u < NP && iy >= 0 && iy < a.H && ix >= 0 && ix < a.W && ch < a.Cin)
                v = *reinterpret_cast<const cv_u32x4*>(in + ((long long)iy * a.W + ix) * a.in_cs + ch);
            P[k] = v;
        }
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int u = tid + NTH * k;
            const int row = u / KV, cv = u % KV;           // row = co_local*9 + tap
            const int co = nt0 + row / 9, ch = c0 + cv * VEC;
            cv_u32x4 v = cv_u32x4{0u, 0u, 0u, 0u};
            if (u < NWU && co < a.Cout && ch < a.Cin)
                v = *reinterpret_cast<const cv_u32x4*>(wt + ((long long)co * 9 + row % 9) * a.Cin + ch);
            Wr[k] = v;
        }
    };
    auto stage = [&](const cv_u32x4 (&P)[RP], const cv_u32x4 (&Wr)[RW]) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RP; ++k) {
            const int u = tid + NTH * k;
            if (u < NP) *reinterpret_cast<cv_u32x4*>(&sP[(u / KV) * ROW + (u % KV) * VEC]) = P[k];
        }
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int u = tid + NTH * k;
            if (u < NWU) *reinterpret_cast<cv_u32x4*>(&sW[(u / KV) * ROW + (u % KV) * VEC]) = Wr[k];
        }
        __syncthreads();
    };
    // The MFMA's K (32 for 16-bit, 16 for FP32: four lane groups of CPL
    // channels) covers TPK taps of a chunk narrower than K: lane group lg takes
    // tap step*TPK + lg/LPT, channels (lg % LPT)*CPL.. of it (a 16-bit layer
    // with 8 or 16 input channels runs 3 or 5 K-steps instead of 9 mostly
    // zero ones).  Each step's fragments are read from LDS while the previous
    // step's MFMAs issue (two register sets).
    constexpr int KK = EK == EK_F32 ? 16 : 32;
    constexpr int TPK = KK / KC, LPT = 4 / TPK, CPL = KC / LPT;
    constexpr int NSTEP = (9 + TPK - 1) / TPK;
    static_assert(KK % KC == 0 && CPL * (int)sizeof(T) == 16, "a lane group reads 16 bytes of one tap");
    auto compute = [&]() {
        cv_u32x4 wf[2][NT / 16], xf[2][MR];
        auto frag = [&](int step, cv_u32x4 (&w)[NT / 16], cv_u32x4 (&x)[MR]) {
            const int tap = step * TPK + lg / LPT, coff = (lg % LPT) * CPL;
            if (TPK > 1 && tap >= 9) {   // padding of the last K-step
#pragma unroll
                for (int i = 0; i < NT / 16; ++i) w[i] = cv_u32x4{0u, 0u, 0u, 0u};
#pragma unroll
                for (int m = 0; m < MR; ++m) x[m] = cv_u32x4{0u, 0u, 0u, 0u};
                return;
            }
            const int ky = tap / 3, kx = tap % 3;
#pragma unroll
            for (int i = 0; i < NT / 16; ++i)
                w[i] = *reinterpret_cast<const cv_u32x4*>(&sW[((i * 16 + l16) * 9 + tap) * ROW + coff]);
#pragma unroll
            for (int m = 0; m < MR; ++m) {
                const int pr = (q * MR + m) * S + ky, pc = l16 * S + kx;
                x[m] = *reinterpret_cast<const cv_u32x4*>(&sP[(pr * PC + pc) * ROW + coff]);
            }
        };
        frag(0, wf[0], xf[0]);
#pragma unroll
        for (int step = 0; step < NSTEP; ++step) {
            if (step + 1 < NSTEP) frag(step + 1, wf[(step + 1) & 1], xf[(step + 1) & 1]);
            // keep the reads of the next step ahead of this step's MFMAs
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < NT / 16; ++i)
#pragma unroll
                for (int m = 0; m < MR; ++m) acc[i][m] = conv_mfma<EK>(wf[step & 1][i], xf[step & 1][m], acc[i][m]);
            __builtin_amdgcn_sched_barrier(0);
        }
    };

#pragma unroll
    for (int d = 0; d < DEPTH; ++d)
        if (cbeg + d * KC < cend) fetch(cbeg + d * KC, rp[d], rw[d]);
    for (int c0 = cbeg; c0 < cend; c0 += DEPTH * KC) {
#pragma unroll
        for (int d = 0; d < DEPTH; ++d) {
            const int cc = c0 + d * KC;
            if (cc >= cend) break;
            stage(rp[d], rw[d]);
            if (cc + DEPTH * KC < cend) fetch(cc + DEPTH * KC, rp[d], rw[d]);
            compute();
        }
    }

    // epilogue: lane holds output channels nt0 + 16i + 4lg + r of pixel l16 of row q*MR+m
    const float g = a.gate ? *a.gate : 1.0f;
#pragma unroll
    for (int m = 0; m < MR; ++m) {
        const int oy = py + ps * (ty * TH + q * MR + m), ox = px + ps * (tx * TW + l16);
        if (oy >= a.Ho || ox >= a.Wo) continue;
        const long long pix = ((long long)n * a.Ho + oy) * a.Wo + ox;
#pragma unroll
        for (int i = 0; i < NT / 16; ++i) {
            const int co = nt0 + i * 16 + 4 * lg;
            const int nv = min(4, a.Cout - co);
            if (nv <= 0) continue;
            if (a.partial) {
                float* d = a.partial + kp * a.part_stride + pix * a.Cout + co;
                for (int r = 0; r < nv; ++r) d[r] = acc[i][m][r];
                continue;
            }
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float x = acc[i][m][r];
                if (a.bias && r < nv) x += a.bias[co + r];
                if (a.act == 1) x = x > 0.0f ? x : x * 0.1f;
                else if (a.act == 2) x = x > 0.0f ? x : 0.0f;
                v[r] = a.gate ? x * g : x;
            }
            if (a.add) conv_load4<EKO>(a.add, pix * a.add_cs + a.add_c0 + co, v, nv);
            conv_store4<EKO>(a.out, pix * a.out_cs + a.out_c0 + co, v, nv);
            // (compiled into the 16-channel tiles only: the flow heads use it)
            if constexpr (NT == 16)
                if (a.out2) conv_store4<EK>(a.out2, pix * a.out2_cs + a.out2_c0 + co, v, nv);
        }
    }
}

// sum of the split partials in part order, then the epilogue of k_conv3x3
template <int EK, int EKO>
__global__ void k_conv_finalize(ConvArgs a, int parts)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)a.N * a.Ho * a.Wo * a.Cout;
    if (e >= total) return;
    const int co = (int)(e % a.Cout);
    const long long pix = e / a.Cout;
    float x = 0.0f;
    for (int k = 0; k < parts; ++k) x += a.partial[k * a.part_stride + e];
    if (a.bias) x += a.bias[co];
    if (a.act == 1) x = x > 0.0f ? x : x * 0.1f;
    else if (a.act == 2) x = x > 0.0f ? x : 0.0f;
    if (a.gate) x *= *a.gate;
    using E = Elem<EKO>;
    if (a.add) x += E::ld(static_cast<const typename E::T*>(a.add)[pix * a.add_cs + a.add_c0 + co]);
    static_cast<typename E::T*>(a.out)[pix * a.out_cs + a.out_c0 + co] = E::st(x);
    if (a.out2) static_cast<typename Elem<EK>::T*>(a.out2)[pix * a.out2_cs + a.out2_c0 + co] = Elem<EK>::st(x);
}

template <int EK, int EKO, int S, int NT, int MR, int NW, int KCX>
static void conv_go(const ConvArgs& a, hipStream_t s)
{
    constexpr int KC = KCX;
    const int ps = a.pstep;
    ConvArgs b = a;
    const int hs = (a.Ho + ps - 1) / ps, ws = (a.Wo + ps - 1) / ps;
    b.tiles_y = (hs + NW * MR - 1) / (NW * MR);
    b.tiles_x = (ws + 15) / 16;
    b.ctiles = (a.Cout + NT - 1) / NT;
    const int chunks = (a.Cin + KC - 1) / KC;
    // split the channel loop when the grid leaves most CUs idle (FP32: under
    // 128 blocks per item; 16-bit: under 32, where the partial-sum traffic
    // and the finalize launch cost more than idle CUs down to that size, as
    // measured on 1080p frames) and the caller provided room for the partial
    // sums.  The split depends on one batch item's geometry and the room per
    // item only, so an item's result has the same bits alone and in a batch
    // of any size.
    const long long item_blocks = (long long)b.tiles_x * b.tiles_y * ps * ps * b.ctiles;
    const long long item_stride = (long long)a.Ho * a.Wo * a.Cout;
    const long long stride = item_stride * a.N;
    int parts = 1;
    if (a.partial && chunks > 1 && item_blocks < (EK == EK_F32 ? 128 : 32)) {
        parts = (int)std::min<long long>(chunks, (256 + item_blocks - 1) / item_blocks);
        parts = (int)std::min<long long>(parts, a.part_stride / std::max<long long>(item_stride, 1));
        if (parts < 2) parts = 1;
    }
    b.cpp = (chunks + parts - 1) / parts;
    parts = (chunks + b.cpp - 1) / b.cpp;
    b.partial = parts > 1 ? a.partial : nullptr;
    b.part_stride = stride;
    dim3 grid(b.tiles_x * b.tiles_y * ps * ps, b.ctiles * parts, a.N);
    hipLaunchKernelGGL((k_conv3x3<EK, EKO, S, NT, MR, NW, KCX>), grid, dim3(64 * NW), 0, s, b);
    if (parts > 1)
        hipLaunchKernelGGL((k_conv_finalize<EK, EKO>), dim3((unsigned)((stride + 255) / 256)), dim3(256), 0, s, b, parts);
}

template <int EK, int EKO, int S, int NT, int KCX>
static void conv_mr(const ConvArgs& a, int mr, hipStream_t s)
{
    if (mr == 4) conv_go<EK, EKO, S, NT, (S == 1 ? 4 : 2), 4, KCX>(a, s);
    else if (mr == 2) conv_go<EK, EKO, S, NT, 2, 4, KCX>(a, s);
    else conv_go<EK, EKO, S, NT, 1, 4, KCX>(a, s);
}

// Tile choice.  Channel tile: FP32 layers take the widest tile they fill (64,
// 32 or 16 output channels); 16-bit layers take 32 (16 when the layer has at
// most 16): twice the blocks of 64-channel tiles on the small grids of the
// coarse levels and half the weights staged per block (1080p frames: 556 vs
// 521 frames/s on one stream, 670 vs 656 frame-pipelined).  Pixel tile:
// 256 pixels from two blocks per CU; 128 pixels from two blocks per CU (FP32)
// or from half a block per CU (16-bit: its MFMAs are 16x faster, so the weights
// every block re-reads dominate, and halving that traffic pays for the idle
// CUs of a smaller grid); else 64 pixels.
template <int EK, int EKO, int S>
static void conv_pick(const ConvArgs& a, hipStream_t s)
{
    const int ps = a.pstep;
    const long long px = (long long)((a.Ho + ps - 1) / ps) * ps * ((a.Wo + ps - 1) / ps) * ps * a.N;
    const int NTsel = EK == EK_F32 ? (a.Cout > 32 ? 64 : (a.Cout > 16 ? 32 : 16)) : (a.Cout > 16 ? 32 : 16);
    const long long ntiles = (a.Cout + NTsel - 1) / NTsel;
    auto blocks = [&](int mr) { return (px / (64LL * mr)) * ntiles; };
    int mr = (S == 1 && blocks(4) >= 512) ? 4 : (blocks(2) >= (EK == EK_F32 ? 512 : 128) ? 2 : 1);
    if (S == 2 && mr > 2) mr = 2;
    constexpr int KC = ConvK<EK>::KC;
    if constexpr (EK == EK_F32) {
        if (NTsel == 64) conv_mr<EK, EKO, S, 64, KC>(a, mr, s);
        else if (NTsel == 32) conv_mr<EK, EKO, S, 32, KC>(a, mr, s);
        else conv_mr<EK, EKO, S, 16, KC>(a, mr, s);
    } else {
        // layers with 8 or 16 input channels (the image and the first feature
        // level) stage chunks of that width
        if (NTsel == 32) {
            if (a.Cin <= 8) conv_mr<EK, EKO, S, 32, 8>(a, mr, s);
            else if (a.Cin <= 16) conv_mr<EK, EKO, S, 32, 16>(a, mr, s);
            else conv_mr<EK, EKO, S, 32, KC>(a, mr, s);
        } else {
            if (a.Cin <= 8) conv_mr<EK, EKO, S, 16, 8>(a, mr, s);
            else if (a.Cin <= 16) conv_mr<EK, EKO, S, 16, 16>(a, mr, s);
            else conv_mr<EK, EKO, S, 16, KC>(a, mr, s);
        }
    }
}

template <int EK, int EKO>
static void conv_stride(const ConvArgs& a, int stride, hipStream_t s)
{
    if (stride == 2) conv_pick<EK, EKO, 2>(a, s);
    else conv_pick<EK, EKO, 1>(a, s);
}

int launch_conv3x3(const ConvDesc& d, hipStream_t s)
{
    const int ek = d.ek_in, eko = d.ek_out;
    if (ek < 0 || eko < 0) return set_error(VSC_E_DTYPE, "conv3x3: unsupported element type");
    if (eko != ek && eko != EK_F32) return set_error(VSC_E_DTYPE, "conv3x3: output must be FP32 or the input type");
    const int vec = ek == EK_F32 ? 4 : 8;
    if (d.stride != 1 && d.stride != 2) return set_error(VSC_E_INVALID, "conv3x3: stride must be 1 or 2");
    if (d.dil < 1 || (d.stride == 2 && d.dil != 1)) return set_error(VSC_E_INVALID, "conv3x3: bad dilation");
    if (d.in.c <= 0 || d.in.c % vec || d.in.cs % vec || d.in.c0 % vec || d.in.c0 + d.in.c > d.in.cs)
        return set_error(VSC_E_INVALID, "conv3x3: input slice must be whole 16-byte vectors inside the pixel");
    const int Cout = d.out.c;
    if (Cout <= 0 || d.out.c0 + Cout > d.out.cs) return set_error(VSC_E_INVALID, "conv3x3: bad output slice");
    if ((d.out.c0 % 4 || d.out.cs % 4) && Cout >= 4)
        return set_error(VSC_E_INVALID, "conv3x3: output slice must be 4-channel aligned");
    if (d.add.p && (d.add.c0 % 4 || d.add.cs % 4) && Cout >= 4)
        return set_error(VSC_E_INVALID, "conv3x3: residual slice must be 4-channel aligned");
    const int Ho = (d.in.h + 2 * d.pad - 2 * d.dil - 1) / d.stride + 1;
    const int Wo = (d.in.w + 2 * d.pad - 2 * d.dil - 1) / d.stride + 1;
    if (Ho != d.out.h || Wo != d.out.w || d.out.n != d.in.n || (d.add.p && (d.add.h != Ho || d.add.w != Wo)))
        return set_error(VSC_E_INVALID, "conv3x3: output size does not match");
    if (d.out2.p && (Cout > 16 || d.out2.c != Cout || d.out2.c0 + Cout > d.out2.cs || d.out2.h != Ho ||
                     d.out2.w != Wo || d.out2.n != d.in.n || ((d.out2.c0 % 4 || d.out2.cs % 4) && Cout >= 4)))
        return set_error(VSC_E_INVALID, "conv3x3: bad second output slice (at most 16 channels)");
    if ((long long)d.in.n * d.in.h * d.in.w * d.in.cs >= (1LL << 31) ||
        (long long)d.out.n * Ho * Wo * d.out.cs >= (1LL << 31))
        return set_error(VSC_E_INVALID, "conv3x3: tensor too large");
    if (Ho <= 0 || Wo <= 0) return VSC_OK;
    ConvArgs a{};
    a.in = d.in.p; a.in_cs = d.in.cs; a.in_c0 = d.in.c0;
    a.N = d.in.n; a.H = d.in.h; a.W = d.in.w; a.Cin = d.in.c;
    a.wt = d.wt; a.bias = d.bias;
    a.out = d.out.p; a.out_cs = d.out.cs; a.out_c0 = d.out.c0;
    a.out2 = d.out2.p; a.out2_cs = d.out2.cs; a.out2_c0 = d.out2.c0;
    a.add = d.add.p; a.add_cs = d.add.cs; a.add_c0 = d.add.c0;
    a.gate = d.gate;
    a.partial = d.partial;
    a.part_stride = (long long)d.partial_cap;   // capacity per item in floats; conv_go turns it into the part stride
    a.Ho = Ho; a.Wo = Wo; a.Cout = Cout;
    a.pad = d.pad; a.pstep = d.dil; a.act = d.act;
    if (ek == EK_F32) conv_stride<EK_F32, EK_F32>(a, d.stride, s);
    else if (ek == EK_BF16 && eko == EK_BF16) conv_stride<EK_BF16, EK_BF16>(a, d.stride, s);
    else if (ek == EK_BF16) conv_stride<EK_BF16, EK_F32>(a, d.stride, s);
    else if (eko == EK_F16) conv_stride<EK_F16, EK_F16>(a, d.stride, s);
    else conv_stride<EK_F16, EK_F32>(a, d.stride, s);
    return check_launch("conv3x3");
}

}  // namespace vsc

extern "C" int vsc_conv3x3_nhwc(const void* in, int in_cs, int in_c0, int N, int H, int W, int Cin, const void* wt,
                                const float* bias, void* out, int out_cs, int out_c0, int Cout, int stride,
                                int dilation, int pad, int act, const void* add, int add_cs, int add_c0, int dtype,
                                int out_dtype, vsc_hip_stream stream)
{
    using namespace vsc;
    if (!in || !wt || !out || N <= 0 || H <= 0 || W <= 0)
        return set_error(VSC_E_INVALID, "vsc_conv3x3_nhwc: null pointer or empty shape");
    if (stride != 1 && stride != 2) return set_error(VSC_E_INVALID, "vsc_conv3x3_nhwc: stride must be 1 or 2");
    if (act < 0 || act > 2) return set_error(VSC_E_INVALID, "vsc_conv3x3_nhwc: act must be 0, 1 or 2");
    ConvDesc d{};
    const int Ho = (H + 2 * pad - 2 * dilation - 1) / stride + 1, Wo = (W + 2 * pad - 2 * dilation - 1) / stride + 1;
    d.in = Nhwc{const_cast<void*>(in), N, H, W, Cin, in_cs, in_c0};
    d.out = Nhwc{out, N, Ho, Wo, Cout, out_cs, out_c0};
    d.add = Nhwc{const_cast<void*>(add), N, Ho, Wo, Cout, add_cs, add_c0};
    d.wt = wt; d.bias = bias;
    d.stride = stride; d.dil = dilation; d.pad = pad; d.act = act;
    d.ek_in = elem_kind(dtype); d.ek_out = elem_kind(out_dtype);
    return launch_conv3x3(d, as_stream(stream));
}
