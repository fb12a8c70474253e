// Internal declarations shared by the HIP translation units of libvsc.
// Parity-critical arithmetic: every TU is compiled with -ffp-contract=off and
// additionally pins the contraction policy with the pragma below, so that no
// a*b+c is fused into an FMA (the CPU oracle follows the same policy).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/vsc.h"

#pragma clang fp contract(off)

namespace vsc {

// ---------------------------------------------------------------------------
// error state (thread-local message behind vsc_last_error)
// ---------------------------------------------------------------------------
int set_error(int code, const std::string& msg);
int check_launch(const char* what);
int check_hip(hipError_t e, const char* what);

inline hipStream_t as_stream(vsc_hip_stream s) { return reinterpret_cast<hipStream_t>(s); }

// kernel timing (vsc_profile_begin/end): returns a slot index or -1 when off
enum ProfKernel { PROF_CONSIST_BLOCKED = 0, PROF_FRONT_PASS = 1, PROF_NKERNELS = 2 };
int prof_start(ProfKernel k, double algorithmic_bytes, hipStream_t s);
void prof_stop(int slot, hipStream_t s);

// ---------------------------------------------------------------------------
// strided float image view: element (x, y, c) at d[c*cs + (y*w + x)*ps].
//   HWC interleaved (the reference GPUImage layout): cs = 1, ps = channels
//   planar (one plane per channel, used for solver state): cs = w*h, ps = 1
// ---------------------------------------------------------------------------
struct View {
    float* d;
    int w, h, ch;
    long long cs, ps;
    __host__ __device__ inline long long idx(int x, int y, int c) const {
        return (long long)c * cs + ((long long)y * w + x) * ps;
    }
    __device__ inline float rd(int x, int y, int c) const { return d[idx(x, y, c)]; }
    __device__ inline void wr(int x, int y, int c, float v) const { d[idx(x, y, c)] = v; }
};

// The same view with 32-bit offsets (every element offset < 2^31, checked by
// the launcher): cheaper addressing in the per-pixel kernels.
struct View32 {
    float* d;
    int w, cs, ps;
    __device__ inline int idx(int x, int y, int c) const { return c * cs + (y * w + x) * ps; }
    __device__ inline float rd(int x, int y, int c) const { return d[idx(x, y, c)]; }
    __device__ inline void wr(int x, int y, int c, float v) const { d[idx(x, y, c)] = v; }
};
__host__ __device__ inline View32 v32(const View& v) { return View32{v.d, v.w, (int)v.cs, (int)v.ps}; }
inline bool fits32(const View& v) { return (long long)(v.ch - 1) * v.cs + ((long long)v.h * v.w - 1) * v.ps < (1LL << 31); }

inline View hwc(const vsc_image& im) {
    return View{im.data, im.width, im.height, im.channels, 1, (long long)im.channels};
}
__host__ __device__ inline View planar(float* d, int w, int h, int ch) {
    return View{d, w, h, ch, (long long)w * h, 1};
}

// Byte to float as kernel_to_float_image (gpuimage.cu:49): (float)((double)v/255.0).
// 1/255 = 2^-8 + 2^-16 + 2^-24 + ...: q0 = v * (65793 * 2^-24) is exact
// (v * 65793 < 2^24) and v/255 = q0 * (1 + 2^-24 + ...), whose tail lies
// between half an ulp and an ulp of q0, so the correctly rounded quotient is
// fma(q0, 2^-24, q0) (q0 = 0 stays 0).  Equal to the double quotient rounded
// to float for all 256 values (exhaustive device test).
__host__ __device__ inline float u8_to_f32(uint32_t v)
{
    const float q0 = (float)v * (65793.0f / 16777216.0f);
    return __builtin_fmaf(q0, 1.0f / 16777216.0f, q0);
}

// RGBA8 source image: channel c of pixel (x, y) at p[(y*w + x)*4 + c].
struct Rgba {
    const uint8_t* p;
    int w, h;
    __device__ inline float rd(int x, int y, int c) const { return u8_to_f32(p[((long long)y * w + x) * 4 + c]); }
};

// ---------------------------------------------------------------------------
// exp: Cody & Waite (1980) reduction x = k*ln2 + r with a two-part ln2, then a
// degree-7 Taylor polynomial for e^r (|r| <= ~0.35), Horner form with fused
// multiply-adds, and the scale 2^k through the exponent field.  Only
// correctly rounded binary32 operations: identical results on every IEEE
// device (the oracle's std::fma on the host).
// ---------------------------------------------------------------------------
__host__ __device__ inline float vexp(float x)
{
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97208404541015625f) return 0.0f;
    const float log2e = 1.44269502162933349609375f;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.428606765330187045e-06f;
    float kf = __builtin_rintf(x * log2e);
    int k = (int)kf;
    float r = __builtin_fmaf(-kf, ln2_hi, x);
    r = __builtin_fmaf(-kf, ln2_lo, r);
    float p = 1.0f / 5040.0f;
    p = __builtin_fmaf(p, r, 1.0f / 720.0f);
    p = __builtin_fmaf(p, r, 1.0f / 120.0f);
    p = __builtin_fmaf(p, r, 1.0f / 24.0f);
    p = __builtin_fmaf(p, r, 1.0f / 6.0f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    if (k > 127) { p = p * 2.0f; k -= 1; }
    if (k < -126) { p = p * 5.42101086242752217003726400434970855712890625e-20f; k += 64; }
    return p * __builtin_bit_cast(float, (uint32_t)(k + 127) << 23);
}

// ---------------------------------------------------------------------------
// Stage formulas, one pixel at a time, in the reference's operation order.
// ---------------------------------------------------------------------------

// kernel_warp (flowconsistency.cu:85-102): sample position, clamped to
// [0, W-3] x [0, H-3], split into integer corner and fraction.
struct WarpTap {
    int ix, iy;
    float fx, fy;
};
__device__ inline WarpTap warp_tap(int x, int y, float flo_x, float flo_y, int W, int H)
{
    const float mfx = fmaxf(0.0f, fminf((float)x + flo_x, (float)(W - 3)));
    const float mfy = fmaxf(0.0f, fminf((float)y + flo_y, (float)(H - 3)));
    WarpTap t;
    t.ix = (int)floorf(mfx);
    t.iy = (int)floorf(mfy);
    t.fx = mfx - (float)t.ix;
    t.fy = mfy - (float)t.iy;
    return t;
}

// kernel_warp (flowconsistency.cu:104-113): bilinear from the four corners.
template <class Img>
__device__ inline float warp_sample(const Img& im, const WarpTap& t, int c)
{
    const float t1 = im.rd(t.ix, t.iy, c) * (1.0f - t.fx) + im.rd(t.ix + 1, t.iy, c) * t.fx;
    const float t2 = im.rd(t.ix, t.iy + 1, c) * (1.0f - t.fx) + im.rd(t.ix + 1, t.iy + 1, c) * t.fx;
    return t1 * (1.0f - t.fy) + t2 * t.fy;
}

// kernel_adap_comb (flowconsistency.cu:145-161) for one channel value.
__device__ inline void adap_comb(float ci, float cp, float pi, float pp, float ni, float np, float ls, float alpha,
                                 float& adIn, float& adPr)
{
    float wp = vexp(((-alpha) * (ci - pi)) * (ci - pi));
    float wn = vexp(((-alpha) * (ci - ni)) * (ci - ni));
    if (wp > 0.45f) wp = 0.45f;
    if (wn > 0.3f) wn = 0.3f;
    if (wp < 0.001f) wp = 0.0f;
    if (wn < 0.001f) wn = 0.0f;
    adIn = (wp * pi + wn * ni) + (1.0f - (wp + wn)) * ci;
    float ap = (wp * pp + wn * np) + (1.0f - (wp + wn)) * cp;
    adPr = wp * ls + (1.0f - wp) * ap;
}

// kernel_consist_wt (flowconsistency.cu:177-188).
__device__ inline float consist_wt(float ci, float adIn, float beta, float gamma)
{
    float w = gamma * vexp(((-beta) * (ci - adIn)) * (ci - adIn));
    if (w < 0.001f) w = 0.0f;
    return w;
}

// kernel_bilinear (flowconsistency.cu:58-74): source coordinate and fraction.
struct BiTap {
    int ix, iy, ix1, iy1;
    float fx, fy;
};
__device__ inline BiTap bilinear_tap(int ox, int oy, int inW, int inH, int outW, int outH)
{
    const float xx = ((float)ox * (float)inW) / (float)outW;
    const float yy = ((float)oy * (float)inH) / (float)outH;
    BiTap t;
    t.ix = (int)floorf(xx);
    t.iy = (int)floorf(yy);
    t.fx = xx - (float)t.ix;
    t.fy = yy - (float)t.iy;
    t.ix1 = min(t.ix + 1, inW - 1);
    t.iy1 = min(t.iy + 1, inH - 1);
    return t;
}
// the four taps' blend, in the reference's operation order
__device__ inline float bilinear_mix(float v00, float v10, float v01, float v11, float fx, float fy)
{
    return (((v00 * (1.0f - fx)) * (1.0f - fy) + (v10 * fx) * (1.0f - fy)) + (v01 * (1.0f - fx)) * fy) +
           (v11 * fx) * fy;
}
template <class Img>
__device__ inline float bilinear_sample(const Img& im, const BiTap& t, int c)
{
    const float v00 = im.rd(t.ix, t.iy, c);
    const float v10 = im.rd(t.ix1, t.iy, c);
    const float v01 = im.rd(t.ix, t.iy1, c);
    const float v11 = im.rd(t.ix1, t.iy1, c);
    return bilinear_mix(v00, v10, v01, v11, t.fx, t.fy);
}

// kernel_to_char_image (gpuimage.cu:62-65): floor(|v|*255) to uint32 with
// saturation (NaN -> 0), then the unsigned char cast keeps the low byte.
__device__ inline uint8_t to_u8(float v)
{
    const float f = fabsf(v) * 255.0f;
    uint32_t u;
    if (!(f == f)) u = 0u;
    else if (f >= 4294967296.0f) u = 0xFFFFFFFFu;
    else u = (uint32_t)floorf(f);
    return (uint8_t)(u & 0xFFu);
}

// ---------------------------------------------------------------------------
// launch helpers implemented in stab.hip and used by the stream driver
// ---------------------------------------------------------------------------
int launch_stab_warp(View in_f, const uint8_t* in_rgba, View flow, View out, hipStream_t s);
int launch_bilinear(View in, View out, hipStream_t s);
int launch_consist_sweep(View P, View A, View Wt, View Oin, View Oout, View U, float step, float mom, int use_mom,
                         hipStream_t s);
int launch_lap(const float* P, float* L, int W, int H, hipStream_t s);
// Temporally blocked solver: `halo` = the most sweeps one launch may do,
// `strips` = 4 (full-height tiles) or 2 (half-height tiles, two blocks per CU).
struct BlockedPlan {
    int halo;
    int strips;
};
BlockedPlan plan_consist_blocked(int W, int H, int iters, int fma);
// fma == 0: lapP, A, Wt are the level's lapP, A, w (reference op order);
// fma == 1: lapP holds k = step*fma(-w, A, lapP), Wt holds q = cnt + w, A unused.
// store_u = 0 skips writing the momentum (last launch of a level: the next
// level starts with U = 0).
int launch_consist_blocked(const float* lapP, const float* A, const float* Wt, const float* Oin, float* Oout,
                           const float* Uin, float* Uout, int W, int H, int sweeps, int first_sweep, float step,
                           float mom, int fma, const BlockedPlan& plan, hipStream_t s, int store_u = 1);
// in place: Wt <- cnt + Wt, L <- step*fma(-Wt, A, L) (the FMA-form constants of a planar level)
int launch_fma_consts(float* L, const float* A, float* Wt, int W, int H, float step, hipStream_t s);
int launch_to_rgba(View in, uint8_t* out, hipStream_t s);
int launch_rgba_to_view(const uint8_t* in, View out, hipStream_t s);
int launch_copy_view(View in, View out, hipStream_t s);

struct FrontArgs {
    Rgba oPrev, oCur, oNext, pPrev, pCur, pNext;  // RGBA8 frames (4-byte aligned rows)
    View last;                // previous stabilized frame (planar)
    View fwd, bwd;            // flows (HWC, 2|3 channels) at frame size
    View L0, A0, W0, P0;      // level-0 outputs (planar): lapP, A, w and (write_p) P
    View L1, A1, W1, O1, P1;  // level-1 outputs (planar), written when decimate != 0
    float alpha, beta, gamma;
    int decimate;
    int write_p;              // also write P0 (and P1): the pyramid needs them
    int fma;                  // write the FMA-form constants: L <- step*fma(-w, A, lapP), W <- cnt + w (A unused)
    float step;
};
int launch_front_pass(const FrontArgs& a, hipStream_t s);

// ---------------------------------------------------------------------------
// Batched frames (the multi-stream server): the fused frame of vsc_stab_frame
// (front pass, 1- or 2-level pyramid solve, RGBA8 out) for many streams at
// once, in persistent launches that walk a work list in device memory, so the
// launch sequence is the same for every batch and can be replayed from one
// captured graph.  Every solver launch of a batch runs BATCH_HALO sweeps per
// frame (fewer at a level's end); the results equal the per-stream path's,
// whose launch plan differs (the blocked solver is exact Jacobi for any plan).
// ---------------------------------------------------------------------------
constexpr int BATCH_HALO = 19;
// a 2-level batch item needs level 1 to be level 0's exact decimation and the
// upsampling back to be the exact 2x resample (even sizes, and the resampler's
// x*w/(2w) exact in float: 2*w*w < 2^24)
inline bool batch_exact_2x(int W, int H)
{
    const long long w = W / 2, h = H / 2;
    return (W % 2) == 0 && (H % 2) == 0 && 2 * w * w < (1LL << 24) && 2 * h * h < (1LL << 24);
}
constexpr int BATCH_MAX_LAUNCH = 16;   // solver launches per level: num_iter <= 16 * 19
struct BatchLevel {
    const float* L;       // lapP (reference order) or k (FMA form)
    const float* A;       // adapted processed frame (reference order)
    const float* Wt;      // w (reference order) or q (FMA form)
    float* O[2];
    float* U[2];
    int w, h, iters;
    int start;            // O[start] holds the initial solution
    int tiles_x, tiles_y; // tiles per channel at BATCH_HALO
};
struct BatchFrame {
    FrontArgs front;
    BatchLevel lev[2];    // level 0, level 1 (levels == 2)
    int levels;
    int up_idx;           // level-1 buffer holding its result
    int out_idx;          // level-0 buffer holding the result
    int rgba_vec;         // RGBA8 output with 4 pixels per thread
    float step, mom;
    uint8_t* out_rgba;
};
// work list: header, frames [cap], then int tables of cap+1 entries each:
// front blocks, up-sample blocks, RGBA8 blocks, and the solver tiles of each
// (level, launch)
struct BatchHeader {
    int n_frames;
    int n_launch[2];      // solver launches at level 0 and level 1 (max over the frames)
    int cap;
};
size_t batch_worklist_bytes(int cap);
// host-side fill of a work list (in host memory) from n frames; returns the
// solver launches per level in n_launch
void batch_build(unsigned char* host, int cap, const BatchFrame* frames, int n, int fma);
// the batch's launches on s: work list at `dev` (device copy of batch_build's
// output), n_launch solver launches per level, `grid` blocks per persistent launch
int launch_batch(const unsigned char* dev, const int n_launch[2], int fma, int grid, hipStream_t s);
// tiles per channel of a level at BATCH_HALO
void batch_tiles(int w, int h, int fma, int* tiles_x, int* tiles_y);

// ---------------------------------------------------------------------------
// 16-bit storage formats of the flow network (FP32 arithmetic in registers)
// ---------------------------------------------------------------------------
// binary32 -> bfloat16, round to nearest even (NaN stays NaN)
__host__ __device__ inline uint16_t f32_to_bf16(float x)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }
__host__ __device__ inline uint16_t f32_to_f16(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
__host__ __device__ inline float f16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// element kinds of the channels-last (NHWC) network tensors
enum ElemKind { EK_F32 = 0, EK_BF16 = 1, EK_F16 = 2 };
template <int EK> struct Elem;
template <> struct Elem<EK_F32> {
    using T = float;
    static constexpr int VEC = 4;    // elements per 16-byte load
    __device__ static inline float ld(T v) { return v; }
    __device__ static inline T st(float v) { return v; }
};
template <> struct Elem<EK_BF16> {
    using T = uint16_t;
    static constexpr int VEC = 8;
    __device__ static inline float ld(T v) { return bf16_to_f32(v); }
    __device__ static inline T st(float v) { return f32_to_bf16(v); }
};
template <> struct Elem<EK_F16> {
    using T = uint16_t;
    static constexpr int VEC = 8;
    __device__ static inline float ld(T v) { return f16_to_f32(v); }
    __device__ static inline T st(float v) { return f32_to_f16(v); }
};
inline int elem_kind(int vsc_dtype)
{
    return vsc_dtype == VSC_F32 ? EK_F32 : vsc_dtype == VSC_BF16 ? EK_BF16 : vsc_dtype == VSC_F16 ? EK_F16 : -1;
}
inline int elem_bytes(int ek) { return ek == EK_F32 ? 4 : 2; }

// NHWC tensor slice: channel c of pixel (n, y, x) at p[((n*h + y)*w + x)*cs + c0 + c]
struct Nhwc {
    void* p;
    int n, h, w, c;      // c: channels of the slice
    int cs, c0;          // pixel stride and channel offset (elements)
};

// 3x3 convolution on MFMA (conv.hip).  act: 0 none, 1 LeakyReLU(0.1), 2 ReLU;
// add (optional, same element kind as out) is added after the activation.
struct ConvDesc {
    Nhwc in, out, add;   // add.p == nullptr: no residual
    Nhwc out2;           // out2.p == nullptr: none; else a second copy of the result in the input's element kind (Cout <= 16)
    const void* wt;      // [Cout][3][3][in.c], element kind of in
    const float* bias;   // [Cout] or nullptr
    const float* gate;   // nullptr, or a device scalar multiplying the activated result before `add`
    int stride, dil, pad, act;
    int ek_in, ek_out;
    float* partial;      // optional FP32 scratch: small layers split their channel loop into it
    size_t partial_cap;  // scratch capacity in floats per batch item (the scratch holds in.n items' worth)
};
int launch_conv3x3(const ConvDesc& d, hipStream_t s);

// the flow network with room for `pairs` (1 .. 8) (first, second) pairs per
// decode (vsc_flownet_create: 2)
size_t flownet_footprint_pairs(int width, int height, int downscale, int dtype, int slots, int pairs);
int flownet_create_pairs(int width, int height, int downscale, int dtype, unsigned long long seed, int slots,
                         int pairs, vsc_flownet** out);
// features of m (<= ceil(pairs / 2)) frames into slots slot .. slot+m-1, one batch
int flownet_extract_frames(vsc_flownet* n, const uint8_t* const* rgba, int m, int slot, hipStream_t s);

// the fused learned upsampler on n (1 or 2) flows of one frame in one launch
// (upsampler.hip; vsc_upsampler_run is the one-flow case)
int upsampler_run_flows(vsc_upsampler* u, int n, const float* const flow_lo[2], int lo_w, int lo_h, int lo_ch,
                        const uint8_t* rgba, int width, int height, float scale, float* const flow_out[2], int out_ch,
                        const float* gate, hipStream_t s);

// host-state step of a pipeline frame replayed from a graph (pipeline.hip)
int pipeline_advance(vsc_pipeline* p, long long frame);
// Batched frames of a pipeline (the multi-stream server): whether the frame
// can join a batch (checks of vsc_pipeline_frame, frame-size flows, the
// batchable pyramid forms); the flows part of vsc_pipeline_frame alone on s
// (host state restored on failure); and the frame's batch item after its
// flows part was enqueued (host only; advances the stream's `last`).
int pipeline_batch_check(vsc_pipeline* p, const vsc_params* params);
int pipeline_flows_part(vsc_pipeline* p, const vsc_window* w, long long frame, hipStream_t s);
int pipeline_batch_item(vsc_pipeline* p, const vsc_window* w, long long frame, const vsc_params* params,
                        uint8_t* out_rgba, BatchFrame* bf);
// parameter and state checks of vsc_stab_frame / vsc_pipeline_frame, made
// before any state change (NULL params = defaults)
int stream_params_check(const vsc_stream* s, const vsc_params* params);
int stream_frame_check(const vsc_stream* s, const vsc_params* params);
// One frame of vsc_stab_frame as a batch item: the host-side part of the
// fused frame (flow upsampling, when the flows arrive downscaled, is launched
// on `prep`).  VSC_E_UNSUPPORTED when the frame needs the per-stream path
// (more than 2 levels, odd sizes at 2 levels, more sweeps than a batch runs).
// Advances the stream's `last` as the frame would.
int stream_batch_frame(vsc_stream* s, const vsc_frame_in* in, const vsc_params* params, uint8_t* out_rgba, int fma,
                       hipStream_t prep, BatchFrame* out);

// reduced-precision (BF16 / F16 / FP8 e4m3) correlation on MFMA, corr_mfma.hip
size_t corr_mfma_workspace_bytes(int N, int C, int H, int W, int in_dtype);
bool corr_mfma_reads_nchw(const void* f1, const void* f2, int W, int in_dtype);
int corr_mfma_dispatch(const void* f1, const void* f2, float* out, int N, int C, int H, int W, int md, int legacy,
                       int in_dtype, const float* chan_scale, void* workspace, hipStream_t s);

}  // namespace vsc
