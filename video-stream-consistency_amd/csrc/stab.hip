// Stabilization kernels for gfx950: the per-stage entry points of
// src/stabilization/flowconsistency.cuh and gpuimage.cu, the fused front pass
// and the consistency solver (single-sweep and temporally blocked forms).
#include "vsc_internal.h"

#include <algorithm>
#include <type_traits>

namespace vsc {

static constexpr int BX = 64, BY = 4;  // 256-thread 2D blocks, one wave per image row segment

static inline dim3 grid2d(int w, int h) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY, 1); }

// ---------------------------------------------------------------------------
// warp (get_warp_result): float view or RGBA8 source
// ---------------------------------------------------------------------------
template <class Img>
__global__ __launch_bounds__(256) void k_stab_warp(Img in, View flow, View out)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    const WarpTap t = warp_tap(x, y, flow.rd(x, y, 0), flow.rd(x, y, 1), out.w, out.h);
#pragma unroll
    for (int c = 0; c < 3; ++c) out.wr(x, y, c, warp_sample(in, t, c));
}

int launch_stab_warp(View in_f, const uint8_t* in_rgba, View flow, View out, hipStream_t s)
{
    if (in_rgba) {
        Rgba src{in_rgba, out.w, out.h};
        hipLaunchKernelGGL(k_stab_warp<Rgba>, grid2d(out.w, out.h), dim3(BX, BY), 0, s, src, flow, out);
    } else {
        hipLaunchKernelGGL(k_stab_warp<View>, grid2d(out.w, out.h), dim3(BX, BY), 0, s, in_f, flow, out);
    }
    return check_launch("stab_warp");
}

// ---------------------------------------------------------------------------
// adaptive combination and consistency weights (get_adap_comb, get_consist_wt)
// ---------------------------------------------------------------------------
struct AdapArgs {
    View cIn, cPr, pIn, pPr, nIn, nPr, adIn, adPr, last;
    float alpha;
};
__global__ __launch_bounds__(256) void k_adap_comb(AdapArgs a)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= a.cIn.w || y >= a.cIn.h) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float adIn, adPr;
        adap_comb(a.cIn.rd(x, y, c), a.cPr.rd(x, y, c), a.pIn.rd(x, y, c), a.pPr.rd(x, y, c), a.nIn.rd(x, y, c),
                  a.nPr.rd(x, y, c), a.last.rd(x, y, c), a.alpha, adIn, adPr);
        a.adIn.wr(x, y, c, adIn);
        a.adPr.wr(x, y, c, adPr);
    }
}

__global__ __launch_bounds__(256) void k_consist_wt(View cIn, View adIn, View wt, float beta, float gamma)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= cIn.w || y >= cIn.h) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) wt.wr(x, y, c, consist_wt(cIn.rd(x, y, c), adIn.rd(x, y, c), beta, gamma));
}

// ---------------------------------------------------------------------------
// bilinear resample (get_bilinear); channel count of the output
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bilinear(View in, View out)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    const BiTap t = bilinear_tap(x, y, in.w, in.h, out.w, out.h);
    for (int c = 0; c < out.ch; ++c) out.wr(x, y, c, bilinear_sample(in, t, c));
}

// planar frame with rows of a multiple of 4 pixels at a 16-byte aligned base
static bool planar4(const View& v)
{
    return v.ps == 1 && v.cs == (long long)v.w * v.h && v.w % 4 == 0 && reinterpret_cast<uintptr_t>(v.d) % 16 == 0;
}

// Exact 2x upsampling (out = 2*in in both axes, 2*in.w*in.w and 2*in.h*in.h
// below 2^24): bilinear_tap then computes x*in.w/out.w = x/2 exactly, so the
// output quad (2x..2x+1, 2y..2y+1) takes taps (x, y), (x+1, y), (x, y+1),
// (x+1, y+1) (clamped) with weights fx, fy in {0, 1/2}.  One thread per quad
// loads the 4 taps once per channel and blends them as bilinear_sample does;
// VEC writes each output row pair with one 8-byte store (planar output).
template <bool VEC>
__device__ __forceinline__ void up2_px(const View& in, const View& out, int x, int y)
{
    const int x1 = min(x + 1, in.w - 1), y1 = min(y + 1, in.h - 1);
    for (int c = 0; c < out.ch; ++c) {
        const float v00 = in.rd(x, y, c), v10 = in.rd(x1, y, c);
        const float v01 = in.rd(x, y1, c), v11 = in.rd(x1, y1, c);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const float fy = b ? 0.5f : 0.0f;
            const float p0 = bilinear_mix(v00, v10, v01, v11, 0.0f, fy);
            const float p1 = bilinear_mix(v00, v10, v01, v11, 0.5f, fy);
            if (VEC) {
                *reinterpret_cast<float2*>(out.d + out.idx(2 * x, 2 * y + b, c)) = make_float2(p0, p1);
            } else {
                out.wr(2 * x, 2 * y + b, c, p0);
                out.wr(2 * x + 1, 2 * y + b, c, p1);
            }
        }
    }
}

template <bool VEC>
__global__ __launch_bounds__(256) void k_bilinear_up2(View in, View out)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    up2_px<VEC>(in, out, x, y);
}

int launch_bilinear(View in, View out, hipStream_t s)
{
    const bool up2 = out.w == 2 * in.w && out.h == 2 * in.h && 2LL * in.w * in.w < (1LL << 24) &&
                     2LL * in.h * in.h < (1LL << 24);
    if (up2 && out.ps == 1 && out.cs % 2 == 0 && reinterpret_cast<uintptr_t>(out.d) % 8 == 0)
        hipLaunchKernelGGL(k_bilinear_up2<true>, grid2d(in.w, in.h), dim3(BX, BY), 0, s, in, out);
    else if (up2)
        hipLaunchKernelGGL(k_bilinear_up2<false>, grid2d(in.w, in.h), dim3(BX, BY), 0, s, in, out);
    else
        hipLaunchKernelGGL(k_bilinear, grid2d(out.w, out.h), dim3(BX, BY), 0, s, in, out);
    return check_launch("bilinear");
}

__global__ __launch_bounds__(256) void k_copy_view(View in, View out)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    for (int c = 0; c < out.ch; ++c) out.wr(x, y, c, in.rd(x, y, c));
}

int launch_copy_view(View in, View out, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_view, grid2d(out.w, out.h), dim3(BX, BY), 0, s, in, out);
    return check_launch("copy_view");
}

__global__ __launch_bounds__(256) void k_average(View a, View b, View out)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    // kernel_average_frames (flowconsistency.cu:43-47) evaluates 0.5*v1 + 0.5*v2
    // in double and stores a float.  Both halves are exact in float and their
    // double sum is exact whenever the exponents differ by less than 29, so
    // the single float addition rounds to the same value.
    for (int c = 0; c < 3; ++c) out.wr(x, y, c, 0.5f * a.rd(x, y, c) + 0.5f * b.rd(x, y, c));
}

// ---------------------------------------------------------------------------
// u8 <-> float (kernel_to_float_image, kernel_to_char_image)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rgba_to_view(const uint8_t* in, View out)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    const Rgba src{in, out.w, out.h};
#pragma unroll
    for (int c = 0; c < 3; ++c) out.wr(x, y, c, src.rd(x, y, c));
}

int launch_rgba_to_view(const uint8_t* in, View out, hipStream_t s)
{
    hipLaunchKernelGGL(k_rgba_to_view, grid2d(out.w, out.h), dim3(BX, BY), 0, s, in, out);
    return check_launch("rgba_to_float");
}

__device__ __forceinline__ void to_rgba_px(const View& in, uint8_t* out, int x, int y)
{
    uchar4 v;
    v.x = to_u8(in.rd(x, y, 0));
    v.y = to_u8(in.rd(x, y, 1));
    v.z = to_u8(in.rd(x, y, 2));
    v.w = 1;
    reinterpret_cast<uchar4*>(out)[(long long)y * in.w + x] = v;
}

__global__ __launch_bounds__(256) void k_to_rgba(View in, uint8_t* out)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    to_rgba_px(in, out, x, y);
}

// Planar frames whose rows are multiples of 4 pixels: 4 pixels per thread,
// 16-byte loads and stores (the same conversion as k_to_rgba).
__device__ __forceinline__ void to_rgba4_px(const View& in, uint8_t* out, int x0, int y)
{
    const float4 r = *reinterpret_cast<const float4*>(in.d + in.idx(x0, y, 0));
    const float4 g = *reinterpret_cast<const float4*>(in.d + in.idx(x0, y, 1));
    const float4 b = *reinterpret_cast<const float4*>(in.d + in.idx(x0, y, 2));
    auto px = [](float cr, float cg, float cb) {
        return (uint32_t)to_u8(cr) | ((uint32_t)to_u8(cg) << 8) | ((uint32_t)to_u8(cb) << 16) | (1u << 24);
    };
    const uint4 v = make_uint4(px(r.x, g.x, b.x), px(r.y, g.y, b.y), px(r.z, g.z, b.z), px(r.w, g.w, b.w));
    *reinterpret_cast<uint4*>(out + ((long long)y * in.w + x0) * 4) = v;
}

__global__ __launch_bounds__(256) void k_to_rgba_planar4(View in, uint8_t* out)
{
    const int x0 = 4 * (blockIdx.x * 64 + threadIdx.x % 64);
    const int y = blockIdx.y * 4 + threadIdx.x / 64;
    if (x0 >= in.w || y >= in.h) return;
    to_rgba4_px(in, out, x0, y);
}

int launch_to_rgba(View in, uint8_t* out, hipStream_t s)
{
    if (planar4(in) && reinterpret_cast<uintptr_t>(out) % 16 == 0)
        hipLaunchKernelGGL(k_to_rgba_planar4, dim3((in.w / 4 + 63) / 64, (in.h + 3) / 4), dim3(256), 0, s, in, out);
    else
        hipLaunchKernelGGL(k_to_rgba, grid2d(in.w, in.h), dim3(BX, BY), 0, s, in, out);
    return check_launch("float_to_rgba");
}

// ---------------------------------------------------------------------------
// consistency solver, one sweep per launch (Jacobi: Oin -> Oout; U in place,
// since U at a pixel is read and written only by that pixel's thread).
// kernel_consist_out, flowconsistency.cu:208-257.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_consist_sweep(View P, View A, View Wt, View Oin, View Oout, View U,
                                                       float step, float mom, int use_mom)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    const int W = Oin.w, H = Oin.h;
    if (x >= W || y >= H) return;
    const bool hr = (x + 1) < (W - 1), hl = (x - 1) >= 0, hd = (y + 1) < (H - 1), hu = (y - 1) >= 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int cnt = 0;
        float lp = 0.0f, lo = 0.0f;
        if (hr) { lp += P.rd(x + 1, y, c); lo += Oin.rd(x + 1, y, c); cnt += 1; }
        if (hl) { lp += P.rd(x - 1, y, c); lo += Oin.rd(x - 1, y, c); cnt += 1; }
        if (hd) { lp += P.rd(x, y + 1, c); lo += Oin.rd(x, y + 1, c); cnt += 1; }
        if (hu) { lp += P.rd(x, y - 1, c); lo += Oin.rd(x, y - 1, c); cnt += 1; }
        const float o = Oin.rd(x, y, c);
        lp -= (float)cnt * P.rd(x, y, c);
        lo -= (float)cnt * o;
        const float t1 = Wt.rd(x, y, c) * (o - A.rd(x, y, c));
        const float t2 = lo - lp;
        const float g = t2 - t1;
        const float sg = step * g;
        const float on = use_mom ? (o + sg) + mom * U.rd(x, y, c) : o + sg;
        U.wr(x, y, c, sg);
        Oout.wr(x, y, c, on);
    }
}

int launch_consist_sweep(View P, View A, View Wt, View Oin, View Oout, View U, float step, float mom, int use_mom,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_consist_sweep, grid2d(Oin.w, Oin.h), dim3(BX, BY), 0, s, P, A, Wt, Oin, Oout, U, step, mom,
                       use_mom);
    return check_launch("consist_sweep");
}

// ---------------------------------------------------------------------------
// consistency solver, temporally blocked (planar state, one channel per block)
//
// A block of 128*STRIPS threads owns a TW x TH tile of one channel plane
// (TW = 128 columns, TH = STRIPS*R rows) that overlaps its neighbours by a halo of `halo`
// pixels.  Thread (tx, strip) keeps the R pixels of column tx in rows
// strip*R .. strip*R+R-1 in registers: O, U, lapP, A, w.  O and U are read
// with halo from the *in* planes and written to the *out* planes (neighbouring
// blocks read the pixels a block writes).  After `sweeps` <= halo sweeps the
// inner (TW-2*halo) x (TH-2*halo) pixels are exact and written back; halo
// pixels whose out-of-tile neighbours are missing never reach the output.
//
// Neighbour exchange goes through LDS, stored column-major: the R values of
// (strip, column) are contiguous with a padded stride RS = R+4 (RS/4 odd, so a
// 16-byte access of 8 consecutive lanes covers all 32 banks once).  A sweep
// reads the left and right columns of one plane with 16-byte loads issued one
// 4-row chunk ahead of their use and writes each chunk's new values into the
// other plane with a 16-byte store as soon as they are computed; the planes
// alternate, one barrier per sweep.  The up/down neighbours come from the
// thread's own registers (LDS only at the strip ends).
//
// Arithmetic per pixel and sweep is that of k_consist_sweep; lapP (the
// Laplacian of the processed frame, a constant of the level) is an input,
// computed once per level with the same operations, so the result is identical.
// The reference's neighbour conditions (flowconsistency.cu:215-236) cost no
// per-pixel work where they are uniform: a column whose left (right) neighbour
// is excluded reads a zero column from LDS instead (the neighbour sum gains a
// +0 term, which changes no nonzero sum), so only waves whose strip meets the
// image's top or bottom rows test rows per pixel, and waves wholly outside the
// image (never stored, never read by an image pixel) skip the arithmetic.
// ---------------------------------------------------------------------------
static constexpr int SB_TW = 128;
static constexpr int SB_R = 28;      // rows per thread, reference op order (5 values per pixel)
static constexpr int SB_R_FMA = 32;  // FMA form: 4 values per pixel in registers, taller strips
static constexpr int SB_STRIPS = 4;

// Tile placement along one axis of length L: tiles of T pixels whose outer
// `h` pixels on a side facing another tile are halo.  When L >= T the first
// tile starts at 0 and the last ends at L (flush with the image: an image
// edge needs no halo, its missing neighbours are the reference's boundary),
// the tiles between output T - 2h pixels each; otherwise tile i starts at
// i*(T - 2h) - h and pixels outside the image are padding.  A tile computes
// [start, start + T) and stores [lo, hi); the output ranges partition [0, L).
struct TileSpan {
    int start, lo, hi;
};

__host__ __device__ inline int tile_count(int L, int T, int h)
{
    const int O = T - 2 * h;
    if (L < T) return (L + O - 1) / O;
    if (L == T) return 1;
    const int rest = L - 2 * (T - h);
    return 2 + (rest > 0 ? (rest + O - 1) / O : 0);
}

__host__ __device__ inline TileSpan tile_span(int i, int n, int L, int T, int h)
{
    const int O = T - 2 * h;
    if (L < T) {
        const int s = i * O - h;
        return TileSpan{s, s + h, s + T - h < L ? s + T - h : L};
    }
    if (L == T) return TileSpan{0, 0, L};
    if (i == 0) return TileSpan{0, 0, T - h};
    const int lo = T - h + (i - 1) * O;
    if (i == n - 1) return TileSpan{L - T, lo, L};
    return TileSpan{lo - h, lo, lo + O};
}

// Rows of a strip are visited in ascending (REV false) or descending (REV
// true) order.  A row's old value is still needed by the next row visited
// after the row itself is updated, so a sweep leaves its O and U shifted by
// one register; the kernel alternates the two orders, and a descending sweep
// shifts them back, so a pair of sweeps needs no register copies.  Each
// pixel's arithmetic (and its operation order) is the same in both orders.
// colL / colR hold the left / right neighbours, or zeros where the column has
// none (cntc = number of horizontal neighbours).  EDGE says which rows lack
// a vertical neighbour: SE_NONE none, SE_TOP row 0 (the image's top row),
// SE_BOTTOM rows R-2 and R-1 (the image's last two rows), SE_ROWS any row,
// tested per pixel against rU / rD.
enum StripEdge { SE_NONE = 0, SE_TOP = 1, SE_BOTTOM = 2, SE_ROWS = 3 };

template <int R, bool MOM, int EDGE, bool REV>
__device__ __forceinline__ void sweep_strip(float (&o)[R], float (&u)[R], const float (&a)[R], const float (&w)[R],
                                            const float (&lp)[R], const float* colL, const float* colR,
                                            float* nxt, float up_edge, float dn_edge, float step, float mom, int cntc,
                                            int rU, int rD)
{
    constexpr int NK = R / 4;
    float4 l4 = *reinterpret_cast<const float4*>(colL + (REV ? R - 4 : 0));
    float4 r4 = *reinterpret_cast<const float4*>(colR + (REV ? R - 4 : 0));
    float prev = REV ? dn_edge : up_edge;  // old value of the row visited before
#pragma unroll
    for (int i = 0; i < NK; ++i) {
        const int k = REV ? NK - 1 - i : i;
        float4 ln = l4, rn = r4;
        if (i + 1 < NK) {
            const int kn = REV ? k - 1 : k + 1;
            ln = *reinterpret_cast<const float4*>(colL + 4 * kn);
            rn = *reinterpret_cast<const float4*>(colR + 4 * kn);
        }
        const float lv[4] = {l4.x, l4.y, l4.z, l4.w};
        const float rv[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int j = REV ? 3 - jj : jj;
            const int r = 4 * k + j;
            const float cur = o[r];
            const float up = REV ? (r > 0 ? o[r - 1] : up_edge) : prev;
            const float dn = REV ? prev : (r < R - 1 ? o[r + 1] : dn_edge);
            float lo = rv[j] + lv[j];
            int cnt = cntc;
            if (EDGE == SE_ROWS) {
                if (r <= rD) { lo += dn; cnt += 1; }
                if (r >= rU) { lo += up; cnt += 1; }
            } else {
                if (!(EDGE == SE_BOTTOM && r >= R - 2)) { lo += dn; cnt += 1; }
                if (!(EDGE == SE_TOP && r == 0)) { lo += up; cnt += 1; }
            }
            lo -= (float)cnt * cur;
            const float t1 = w[r] * (cur - a[r]);
            const float t2 = lo - lp[r];
            const float g = t2 - t1;
            const float sg = step * g;
            o[r] = MOM ? (cur + sg) + mom * u[r] : cur + sg;
            u[r] = sg;
            prev = cur;
        }
        // publish the chunk's new values for the next sweep
        *reinterpret_cast<float4*>(nxt + 4 * k) = make_float4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
        l4 = ln;
        r4 = rn;
    }
}

// The same sweep in FMA form: q = cnt + w and kk = step*fma(-w, A, lapP) are
// per-pixel constants, and per pixel
//   X = fma(-q, O, sum of neighbours); U' = fma(step, X, -kk); O' = O + U' (+ mom*U)
// (7 VALU operations in the interior instead of 13).
template <int R, bool MOM, int EDGE, bool REV>
__device__ __forceinline__ void sweep_strip_fma(float (&o)[R], float (&u)[R], const float (&q)[R],
                                                const float (&kk)[R], const float* colL, const float* colR,
                                                float* nxt, float up_edge, float dn_edge, float step, float mom, int rU,
                                                int rD)
{
    constexpr int NK = R / 4;
    float4 l4 = *reinterpret_cast<const float4*>(colL + (REV ? R - 4 : 0));
    float4 r4 = *reinterpret_cast<const float4*>(colR + (REV ? R - 4 : 0));
    float prev = REV ? dn_edge : up_edge;  // old value of the row visited before
#pragma unroll
    for (int i = 0; i < NK; ++i) {
        const int k = REV ? NK - 1 - i : i;
        float4 ln = l4, rn = r4;
        if (i + 1 < NK) {
            const int kn = REV ? k - 1 : k + 1;
            ln = *reinterpret_cast<const float4*>(colL + 4 * kn);
            rn = *reinterpret_cast<const float4*>(colR + 4 * kn);
        }
        const float lv[4] = {l4.x, l4.y, l4.z, l4.w};
        const float rv[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int j = REV ? 3 - jj : jj;
            const int r = 4 * k + j;
            const float cur = o[r];
            const float up = REV ? (r > 0 ? o[r - 1] : up_edge) : prev;
            const float dn = REV ? prev : (r < R - 1 ? o[r + 1] : dn_edge);
            float sum = rv[j] + lv[j];
            if (EDGE == SE_ROWS) {
                if (r <= rD) sum += dn;
                if (r >= rU) sum += up;
            } else {
                if (!(EDGE == SE_BOTTOM && r >= R - 2)) sum += dn;
                if (!(EDGE == SE_TOP && r == 0)) sum += up;
            }
            const float X = __builtin_fmaf(-q[r], cur, sum);
            const float sg = __builtin_fmaf(step, X, -kk[r]);
            const float on = cur + sg;
            o[r] = MOM ? __builtin_fmaf(mom, u[r], on) : on;
            u[r] = sg;
            prev = cur;
        }
        // publish the chunk's new values for the next sweep
        *reinterpret_cast<float4*>(nxt + 4 * k) = make_float4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
        l4 = ln;
        r4 = rn;
    }
}

// LDS row stride of a strip (floats): an odd number of 16-byte groups
template <int R> constexpr int sb_rs() { return ((R / 4) % 2 == 0) ? R + 4 : R + 8; }

// One tile of the temporally blocked solver: `sweeps` Jacobi sweeps of
// channel plane tile `tile` (tiles_x x tiles_y tiles per channel, channels
// one after the other) of a planar 3-channel level.  sO: the block's LDS
// exchange buffer.  Every thread of the block calls it with the same tile.
template <int R, int STRIPS, bool FMA>
__device__ __forceinline__ void consist_tile(
    const float* __restrict__ LapP, const float* __restrict__ A, const float* __restrict__ Wt, const float* Oin,
    float* Oout, const float* Uin, float* Uout, int W, int H, int sweeps, int first_sweep, float step, float mom,
    int halo, int tiles_x, int tiles_y, int store_u, int tile, float (*sO)[STRIPS][SB_TW + 1][sb_rs<R>()])
{
    constexpr int TW = SB_TW;
    constexpr int TH = STRIPS * R;
    constexpr int RS = sb_rs<R>();
    static_assert(R % 4 == 0 && (RS / 4) % 2 == 1, "strip stride must be an odd number of 16-byte groups");
    const int c = tile / (tiles_x * tiles_y);
    const int bxy = tile - c * tiles_x * tiles_y;
    const int bty = bxy / tiles_x, btx = bxy - bty * tiles_x;
    const long long plane = (long long)W * H;
    const float* __restrict__ Lc = LapP + c * plane;
    const float* __restrict__ Ac = A + c * plane;
    const float* __restrict__ Wc = Wt + c * plane;
    const float* __restrict__ Oi = Oin + c * plane;
    float* __restrict__ Oo = Oout + c * plane;
    const float* __restrict__ Ui = Uin + c * plane;
    float* __restrict__ Uo = Uout + c * plane;

    const int tx = threadIdx.x % TW;
    // TW is a multiple of the wave size, so a strip is wave-uniform: keep the
    // row bookkeeping in scalar registers.
    const int strip = __builtin_amdgcn_readfirstlane(threadIdx.x / TW);
    const int wave_tx0 = __builtin_amdgcn_readfirstlane((threadIdx.x % TW) & ~63);
    const int ty0 = strip * R;
    const TileSpan sx = tile_span(btx, tiles_x, W, TW, halo);
    const TileSpan sy = tile_span(bty, tiles_y, H, TH, halo);
    const int gx0 = sx.start, gy0 = sy.start;
    const int gx = gx0 + tx;
    const bool hr = (gx + 1) < (W - 1), hl = (gx - 1) >= 0;
    const int gys = gy0 + ty0;
    // no pixel of this wave (64 columns x R rows) lies in the image
    const bool wave_out = (gx0 + wave_tx0 >= W) || (gx0 + wave_tx0 + 63 < 0) || (gys >= H) || (gys + R - 1 < 0);
    // Which rows of this wave's strip lack a vertical neighbour.  With tiles
    // flush with the image (H >= TH), only the top tile's first strip (row 0)
    // and the bottom tile's last strip (rows R-2, R-1) hold such rows whose
    // values can reach a stored pixel: a tile's other rows with a missing
    // neighbour lie on its own edge towards a neighbouring tile, inside the
    // halo, whose values are not exact anyway.
    int edge;
    if (H >= TH)
        edge = gys == 0 ? SE_TOP : gys + R == H ? SE_BOTTOM : SE_NONE;
    else
        edge = (gys >= 1 && gys + R - 1 <= H - 3) ? SE_NONE : SE_ROWS;
    const int cntc = (hr ? 1 : 0) + (hl ? 1 : 0);

    // FMA: w holds q = cnt + w, lp holds kk; a is unused (and not loaded)
    float o[R], u[R], a[FMA ? 1 : R], w[R], lp[R];
    // Loads are unconditional, from clamped positions, so that all of them
    // issue back to back: a tile pixel outside the image holds some finite
    // in-image value that no in-image pixel ever reads (the neighbour sets
    // exclude out-of-image positions) and that is never stored.
    // 32-bit element offsets (a plane holds < 2^31 pixels).
    const int gxc = min(max(gx, 0), W - 1);
    const bool load_u = first_sweep > 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gyc = min(max(gys + r, 0), H - 1);
        const int i = gyc * W + gxc;
        o[r] = Oi[i];
        if (!FMA) a[r] = Ac[i];
        w[r] = Wc[i];
        lp[r] = Lc[i];
    }
    if (load_u) {
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = Ui[min(max(gys + r, 0), H - 1) * W + gxc];
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = 0.0f;
    }

    // zero columns, published by the first sweep's barrier
    for (int i = threadIdx.x; i < 2 * STRIPS * RS; i += TW * STRIPS) sO[i / (STRIPS * RS)][(i / RS) % STRIPS][TW][i % RS] = 0.0f;
    // neighbour columns (clamped at the tile edge, where the lane is halo)
    const int txl = !hl ? TW : tx > 0 ? tx - 1 : 0;
    const int txr = !hr ? TW : tx < TW - 1 ? tx + 1 : TW - 1;
    const int stU = strip > 0 ? strip - 1 : 0;
    const int stD = strip < STRIPS - 1 ? strip + 1 : STRIPS - 1;
    // row r has an up neighbour iff r >= rU and a down neighbour iff r <= rD
    const int rU0 = 1 - gys;
    const int rD0 = (H - 3) - gys;

    // One sweep: once every strip's values are in plane s&1, update this
    // strip from its neighbours there, writing each 4-row chunk's new values
    // into plane (s+1)&1 as soon as they are computed (the writes spread over
    // the sweep instead of bunching before its barrier).
    auto sweep = [&](int s, auto mom_c, auto rev_c, auto edge_c) {
        constexpr bool MOM = decltype(mom_c)::value, REV = decltype(rev_c)::value;
        constexpr int EDGE = decltype(edge_c)::value;
        float (*buf)[TW + 1][RS] = sO[s & 1];
        __syncthreads();
        const float up_edge = buf[stU][tx][R - 1];
        const float dn_edge = buf[stD][tx][0];
        const float* colL = buf[strip][txl];
        const float* colR = buf[strip][txr];
        float* nxt = sO[(s + 1) & 1][strip][tx];
        int rU = rU0, rD = rD0;
        // opaque per sweep: keeps the compiler from hoisting 2*R row
        // predicates out of the sweep loop into scalar registers
        asm volatile("" : "+s"(rU), "+s"(rD));
        if constexpr (FMA)
            sweep_strip_fma<R, MOM, EDGE, REV>(o, u, w, lp, colL, colR, nxt, up_edge, dn_edge, step, mom, rU, rD);
        else
            sweep_strip<R, MOM, EDGE, REV>(o, u, a, w, lp, colL, colR, nxt, up_edge, dn_edge, step, mom, cntc, rU,
                                           rD);
    };
    // The level's first sweep has no momentum term; after it, ascending and
    // descending sweeps alternate in pairs.  The initial values go to plane 0.
    auto run = [&](auto edge_c) {
        const std::true_type yes{};
        const std::false_type no{};
        {
            float* own = sO[0][strip][tx];
#pragma unroll
            for (int k = 0; k < R / 4; ++k)
                *reinterpret_cast<float4*>(own + 4 * k) =
                    make_float4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
        }
        int s = 0;
        if (first_sweep == 0) {
            sweep(0, no, no, edge_c);
            s = 1;
        }
        for (; s + 1 < sweeps; s += 2) {
            sweep(s, yes, no, edge_c);
            sweep(s + 1, yes, yes, edge_c);
        }
        if (s < sweeps) sweep(s, yes, no, edge_c);
    };
    // Every wave passes one barrier per sweep, whichever branch it takes.
    if (wave_out) {
        for (int s = 0; s < sweeps; ++s) __syncthreads();
    } else if (edge == SE_NONE) {
        run(std::integral_constant<int, SE_NONE>{});
    } else if (edge == SE_TOP) {
        run(std::integral_constant<int, SE_TOP>{});
    } else if (edge == SE_BOTTOM) {
        run(std::integral_constant<int, SE_BOTTOM>{});
    } else {
        run(std::integral_constant<int, SE_ROWS>{});
    }

    // each image pixel is stored by exactly one tile (its span's output range)
    if (gx >= sx.lo && gx < sx.hi) {
        const int wbase = gys * W + gx;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int gy = gys + r;
            if (gy >= sy.lo && gy < sy.hi) {
                Oo[wbase + r * W] = o[r];
                if (store_u) Uo[wbase + r * W] = u[r];
            }
        }
    }
}

template <int R, int STRIPS, bool FMA>
__global__ __launch_bounds__(SB_TW * STRIPS) void k_consist_blocked(
    const float* __restrict__ LapP, const float* __restrict__ A, const float* __restrict__ Wt, const float* Oin,
    float* Oout, const float* Uin, float* Uout, int W, int H, int sweeps, int first_sweep, float step, float mom,
    int halo, int tiles_x, int tiles_y, int per_xcd, int store_u)
{
    // column SB_TW of every plane and strip holds zeros: the excluded neighbour
    __shared__ __attribute__((aligned(16))) float sO[2][STRIPS][SB_TW + 1][sb_rs<R>()];
    // XCD-aware tile order: blocks are dispatched to the 8 XCDs round robin
    // (block b runs on XCD b % 8), so give XCD x the contiguous run of tiles
    // [x*per_xcd, (x+1)*per_xcd) in raster order; neighbouring tiles then
    // share that XCD's L2 for their overlapping halo reads.
    const int tile = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (tile >= tiles_x * tiles_y * 3) return;
    consist_tile<R, STRIPS, FMA>(LapP, A, Wt, Oin, Oout, Uin, Uout, W, H, sweeps, first_sweep, step, mom, halo,
                                 tiles_x, tiles_y, store_u, tile, sO);
}

// lapP of one planar level: the processed-frame half of the Laplacian in
// kernel_consist_out (flowconsistency.cu:215-238), constant over the sweeps.
__global__ __launch_bounds__(256) void k_lap(const float* __restrict__ P, float* __restrict__ L, int W, int H)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    const int c = blockIdx.z;
    if (x >= W || y >= H) return;
    const float* Pc = P + (long long)c * W * H;
    const int i = y * W + x;
    int cnt = 0;
    float l = 0.0f;
    if ((x + 1) < (W - 1)) { l += Pc[i + 1]; cnt += 1; }
    if ((x - 1) >= 0) { l += Pc[i - 1]; cnt += 1; }
    if ((y + 1) < (H - 1)) { l += Pc[i + W]; cnt += 1; }
    if ((y - 1) >= 0) { l += Pc[i - W]; cnt += 1; }
    l -= (float)cnt * Pc[i];
    L[(long long)c * W * H + i] = l;
}

int launch_lap(const float* P, float* L, int W, int H, hipStream_t s)
{
    hipLaunchKernelGGL(k_lap, dim3((W + BX - 1) / BX, (H + BY - 1) / BY, 3), dim3(BX, BY), 0, s, P, L, W, H);
    return check_launch("lap");
}

// FMA-form constants of one planar level, in place: q = cnt + w into Wt and
// kk = step*fma(-w, A, lapP) into L (cnt: the neighbour count of
// kernel_consist_out, flowconsistency.cu:215-236).
__global__ __launch_bounds__(256) void k_fma_consts(float* __restrict__ L, const float* __restrict__ A,
                                                     float* __restrict__ Wt, int W, int H, float step)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= W || y >= H) return;
    const long long i = (long long)blockIdx.z * W * H + (long long)y * W + x;
    const int cnt = ((x + 1) < (W - 1) ? 1 : 0) + ((x - 1) >= 0 ? 1 : 0) + ((y + 1) < (H - 1) ? 1 : 0) +
                    ((y - 1) >= 0 ? 1 : 0);
    const float w = Wt[i];
    L[i] = step * __builtin_fmaf(-w, A[i], L[i]);
    Wt[i] = (float)cnt + w;
}

int launch_fma_consts(float* L, const float* A, float* Wt, int W, int H, float step, hipStream_t s)
{
    hipLaunchKernelGGL(k_fma_consts, dim3((W + BX - 1) / BX, (H + BY - 1) / BY, 3), dim3(BX, BY), 0, s, L, A, Wt, W,
                       H, step);
    return check_launch("fma_consts");
}

// Launch plan of one pyramid level: the halo (= the most sweeps per launch)
// and the tile height, from a cost model of the kernel fitted on MI355X.  A
// launch costs A + r * sweeps, in units of one sweep over one tile per CU:
// A ~ 24 (the tiles' load and store and the launch itself: 1080p launches of
// 72.4 us for 12 sweeps and 50.4 us for 6 sweeps over 2.45 tiles per CU) and
// r = the tiles per CU, half as the mean tile count and half as its ceiling
// (blocks of the last partial round start late but still lengthen it).  A
// wider halo means fewer launches per level but more, mostly redundant, tile
// area.  Half-height tiles (two blocks per CU) when full-height ones would
// leave a quarter of the CUs or more idle (at 960x540, 198 and 252 full-height
// tiles, full height measured faster: equal alone, 2% faster beside the flow
// network of frame pipelining).  Launches of 209 to 256 tiles (one nearly
// full round) are taken only when no other halo fits: they cost a whole round
// anyway and leave almost no CU to work on another stream (at 960x540, 252
// tiles in 3 launches at halo 25 against 198 in 4 at halo 19: 0.8% faster
// alone, 4% slower frame-pipelined).
namespace {

BlockedPlan best_plan(int W, int H, int iters, int R, int strips, long long* tiles_out)
{
    constexpr double A = 24.0;
    BlockedPlan best{12, strips}, best_band{12, strips};
    double best_cost = 1e300, band_cost = 1e300;
    long long band_tiles = 0;
    for (int h = 4; h <= 30; ++h) {
        if (strips * R - 2 * h < 8) break;
        const long long tiles = 3LL * tile_count(W, SB_TW, h) * tile_count(H, strips * R, h);
        const double per_cu = 0.5 * ((double)tiles / 256.0 + (double)((tiles + 255) / 256)) * strips / SB_STRIPS;
        const int full = iters / h, rem = iters % h;
        const double cost = full * (A + per_cu * h) + (rem ? A + per_cu * rem : 0.0);
        const bool near_full = tiles > 208 && tiles <= 256;
        if (near_full) {
            if (cost < band_cost - 1e-9) {
                band_cost = cost;
                best_band = BlockedPlan{h, strips};
                band_tiles = tiles;
            }
        } else if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = BlockedPlan{h, strips};
            *tiles_out = tiles;
        }
    }
    if (best_cost < 1e299) return best;
    *tiles_out = band_tiles;
    return best_band;
}

}  // namespace

BlockedPlan plan_consist_blocked(int W, int H, int iters, int fma)
{
    const int R = fma ? SB_R_FMA : SB_R;
    long long tiles = 0;
    const BlockedPlan full = best_plan(W, H, iters, R, SB_STRIPS, &tiles);
    if (tiles >= 192) return full;
    return best_plan(W, H, iters, R, 2, &tiles);
}

namespace {

template <int R, bool FMA>
int launch_blocked(const float* lapP, const float* A, const float* Wt, const float* Oin, float* Oout, const float* Uin,
                   float* Uout, int W, int H, int sweeps, int first_sweep, float step, float mom,
                   const BlockedPlan& plan, int store_u, hipStream_t s)
{
    const int tiles_x = tile_count(W, SB_TW, plan.halo);
    const int tiles_y = tile_count(H, plan.strips * R, plan.halo);
    const int per_xcd = (tiles_x * tiles_y * 3 + 7) / 8;
    // algorithmic bytes: read O, U and the per-pixel constants (lapP, A, w; FMA
    // form: q, kk) and write O (and U unless it is the level's last launch)
    // once per pixel and channel
    const int slot =
        prof_start(PROF_CONSIST_BLOCKED, ((FMA ? 6.0 : 7.0) - (store_u ? 0.0 : 1.0)) * 3 * 4 * (double)W * H, s);
    if (plan.strips == SB_STRIPS)
        hipLaunchKernelGGL((k_consist_blocked<R, SB_STRIPS, FMA>), dim3(8 * per_xcd), dim3(SB_TW * SB_STRIPS), 0, s,
                           lapP, A, Wt, Oin, Oout, Uin, Uout, W, H, sweeps, first_sweep, step, mom, plan.halo, tiles_x,
                           tiles_y, per_xcd, store_u);
    else
        hipLaunchKernelGGL((k_consist_blocked<R, 2, FMA>), dim3(8 * per_xcd), dim3(SB_TW * 2), 0, s, lapP, A, Wt, Oin,
                           Oout, Uin, Uout, W, H, sweeps, first_sweep, step, mom, plan.halo, tiles_x, tiles_y, per_xcd,
                           store_u);
    prof_stop(slot, s);
    return check_launch("consist_blocked");
}

}  // namespace

int launch_consist_blocked(const float* lapP, const float* A, const float* Wt, const float* Oin, float* Oout,
                           const float* Uin, float* Uout, int W, int H, int sweeps, int first_sweep, float step,
                           float mom, int fma, const BlockedPlan& plan, hipStream_t s, int store_u)
{
    const int R = fma ? SB_R_FMA : SB_R;
    if (plan.strips != SB_STRIPS && plan.strips != 2)
        return set_error(VSC_E_INVALID, "consist_blocked: strips must be 2 or 4");
    if (plan.halo < 1 || SB_TW - 2 * plan.halo < 1 || plan.strips * R - 2 * plan.halo < 1)
        return set_error(VSC_E_INVALID, "consist_blocked: halo leaves no output tile");
    if (sweeps < 1 || sweeps > plan.halo) return set_error(VSC_E_INVALID, "consist_blocked: sweeps out of range");
    if (fma)
        return launch_blocked<SB_R_FMA, true>(lapP, A, Wt, Oin, Oout, Uin, Uout, W, H, sweeps, first_sweep, step, mom,
                                              plan, store_u, s);
    return launch_blocked<SB_R, false>(lapP, A, Wt, Oin, Oout, Uin, Uout, W, H, sweeps, first_sweep, step, mom, plan,
                                       store_u, s);
}

// ---------------------------------------------------------------------------
// fused front pass: 5 warps + adaptive combination + consistency weights for
// one pixel (all 3 channels), reading the RGBA8 frames, the flows and the
// previous output once.  Writes level-0 lapP, A, w (planar; P itself when the
// pyramid needs it) and, for an exact 2x pyramid (even W and H:
// kernel_bilinear is then pure decimation, fx = fy = 0), the level-1 lapP, A,
// w and the level-1 initial solution O1 = P1.  RGBA8 pixels are read as one
// 32-bit word each.
// ---------------------------------------------------------------------------
struct Rgba32 {
    const uint32_t* p;
    int w;
    __device__ inline uint32_t px(int x, int y) const { return p[y * w + x]; }
    __device__ inline float ch(uint32_t v, int c) const { return u8_to_f32((v >> (8 * c)) & 0xFFu); }
};

// bilinear warp sample of all 3 channels of an RGBA8 frame (kernel_warp order)
__device__ inline void warp_rgba(const Rgba32& im, const WarpTap& t, float out[3])
{
    const uint32_t v00 = im.px(t.ix, t.iy), v10 = im.px(t.ix + 1, t.iy);
    const uint32_t v01 = im.px(t.ix, t.iy + 1), v11 = im.px(t.ix + 1, t.iy + 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float t1 = im.ch(v00, c) * (1.0f - t.fx) + im.ch(v10, c) * t.fx;
        const float t2 = im.ch(v01, c) * (1.0f - t.fx) + im.ch(v11, c) * t.fx;
        out[c] = t1 * (1.0f - t.fy) + t2 * t.fy;
    }
}

// lapP of the processed frame at (x, y) of a level that samples the frame
// every `st` pixels (st = 1: level 0; st = 2: the decimated level 1 of size
// Wl x Hl), for all 3 channels.
__device__ inline int lap_rgba(const Rgba32& im, int x, int y, int st, int xl, int yl, int Wl, int Hl, float out[3])
{
    const uint32_t v = im.px(x, y);
    const bool hr = (xl + 1) < (Wl - 1), hl = (xl - 1) >= 0, hd = (yl + 1) < (Hl - 1), hu = (yl - 1) >= 0;
    const uint32_t vr = hr ? im.px(x + st, y) : 0u, vl = hl ? im.px(x - st, y) : 0u;
    const uint32_t vd = hd ? im.px(x, y + st) : 0u, vu = hu ? im.px(x, y - st) : 0u;
    const int cnt = (hr ? 1 : 0) + (hl ? 1 : 0) + (hd ? 1 : 0) + (hu ? 1 : 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float l = 0.0f;
        if (hr) l += im.ch(vr, c);
        if (hl) l += im.ch(vl, c);
        if (hd) l += im.ch(vd, c);
        if (hu) l += im.ch(vu, c);
        l -= (float)cnt * im.ch(v, c);
        out[c] = l;
    }
    return cnt;
}

// the fused front pass at pixel (x, y) (inside the frame)
__device__ __forceinline__ void front_px(const FrontArgs& a, int x, int y)
{
    const int W = a.A0.w, H = a.A0.h;
    const Rgba32 oPrev{reinterpret_cast<const uint32_t*>(a.oPrev.p), W};
    const Rgba32 oCur{reinterpret_cast<const uint32_t*>(a.oCur.p), W};
    const Rgba32 oNext{reinterpret_cast<const uint32_t*>(a.oNext.p), W};
    const Rgba32 pPrev{reinterpret_cast<const uint32_t*>(a.pPrev.p), W};
    const Rgba32 pCur{reinterpret_cast<const uint32_t*>(a.pCur.p), W};
    const Rgba32 pNext{reinterpret_cast<const uint32_t*>(a.pNext.p), W};
    const View32 fwd = v32(a.fwd), bwd = v32(a.bwd), last = v32(a.last);
    const View32 L0 = v32(a.L0), A0 = v32(a.A0), W0 = v32(a.W0), P0 = v32(a.P0);
    const View32 L1 = v32(a.L1), A1 = v32(a.A1), W1 = v32(a.W1), O1 = v32(a.O1), P1 = v32(a.P1);
    const WarpTap tb = warp_tap(x, y, bwd.rd(x, y, 0), bwd.rd(x, y, 1), W, H);
    const WarpTap tf = warp_tap(x, y, fwd.rd(x, y, 0), fwd.rd(x, y, 1), W, H);
    float pwi[3], pwp[3], nwi[3], nwp[3], lap0[3];
    warp_rgba(oPrev, tb, pwi);
    warp_rgba(pPrev, tb, pwp);
    warp_rgba(oNext, tf, nwi);
    warp_rgba(pNext, tf, nwp);
    const int cnt0 = lap_rgba(pCur, x, y, 1, x, y, W, H, lap0);
    const uint32_t vci = oCur.px(x, y), vcp = pCur.px(x, y);
    const bool dec = a.decimate && ((x | y) & 1) == 0 && (x >> 1) < a.A1.w && (y >> 1) < a.A1.h;
    const int x1 = x >> 1, y1 = y >> 1;
    float lap1[3];
    int cnt1 = 0;
    if (dec) cnt1 = lap_rgba(pCur, x, y, 2, x1, y1, a.A1.w, a.A1.h, lap1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float lw = warp_sample(last, tb, c);
        const float ci = oCur.ch(vci, c);
        const float cp = pCur.ch(vcp, c);
        float adIn, adPr;
        adap_comb(ci, cp, pwi[c], pwp[c], nwi[c], nwp[c], lw, a.alpha, adIn, adPr);
        const float wt = consist_wt(ci, adIn, a.beta, a.gamma);
        if (a.fma) {
            L0.wr(x, y, c, a.step * __builtin_fmaf(-wt, adPr, lap0[c]));
            W0.wr(x, y, c, (float)cnt0 + wt);
        } else {
            L0.wr(x, y, c, lap0[c]);
            A0.wr(x, y, c, adPr);
            W0.wr(x, y, c, wt);
        }
        if (a.write_p) P0.wr(x, y, c, cp);
        if (dec) {
            if (a.fma) {
                L1.wr(x1, y1, c, a.step * __builtin_fmaf(-wt, adPr, lap1[c]));
                W1.wr(x1, y1, c, (float)cnt1 + wt);
            } else {
                L1.wr(x1, y1, c, lap1[c]);
                A1.wr(x1, y1, c, adPr);
                W1.wr(x1, y1, c, wt);
            }
            O1.wr(x1, y1, c, cp);
            if (a.write_p) P1.wr(x1, y1, c, cp);
        }
    }
}

__global__ __launch_bounds__(256) void k_front(FrontArgs a)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= a.A0.w || y >= a.A0.h) return;
    front_px(a, x, y);
}

int launch_front_pass(const FrontArgs& a, hipStream_t s)
{
    // algorithmic bytes per pixel: 6 RGBA8 frames (24), previous output (12),
    // 2 flows (2 * 4 * channels), level-0 lapP, A, w (36) and, when decimating,
    // a quarter pixel of level-1 lapP, A, w, O (48 / 4)
    const double px = (double)a.A0.w * a.A0.h;
    // (FMA form: 2 constants per pixel and channel instead of 3)
    const double per_level = a.fma ? 24.0 : 36.0;
    const double bytes = px * (24.0 + 12.0 + 4.0 * (a.fwd.ch + a.bwd.ch) + per_level +
                               (a.decimate ? (per_level + 12.0) / 4.0 : 0.0) +
                               (a.write_p ? (a.decimate ? 15.0 : 12.0) : 0.0));
    for (const View* v : {&a.fwd, &a.bwd, &a.last, &a.L0, &a.A0, &a.W0, &a.P0, &a.L1, &a.A1, &a.W1, &a.O1, &a.P1})
        if (!fits32(*v)) return set_error(VSC_E_INVALID, "front_pass: image too large for 32-bit offsets");
    const int slot = prof_start(PROF_FRONT_PASS, bytes, s);
    hipLaunchKernelGGL(k_front, grid2d(a.A0.w, a.A0.h), dim3(BX, BY), 0, s, a);
    prof_stop(slot, s);
    return check_launch("front_pass");
}

// ---------------------------------------------------------------------------
// Batched frames: persistent kernels over a work list (layout in
// batch_worklist_bytes / batch_build).  Each kernel walks "virtual blocks"
// (front pass, resample, RGBA8) or solver tiles of all frames of the batch;
// a frame's share is found by binary search in the work list's prefix table.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t round256(size_t b) { return (b + 255) / 256 * 256; }

struct BatchLayout {
    size_t frames, front, up, rgba, tiles, total;
    int cap;
    __host__ __device__ BatchLayout(int c) : cap(c)
    {
        frames = round256(sizeof(BatchHeader));
        const size_t tab = round256((size_t)(c + 1) * sizeof(int));
        front = frames + round256((size_t)c * sizeof(BatchFrame));
        up = front + tab;
        rgba = up + tab;
        tiles = rgba + tab;
        total = tiles + 2 * BATCH_MAX_LAUNCH * tab;
    }
    __host__ __device__ size_t tile_table(int slot, int launch) const
    {
        return tiles + (size_t)(slot * BATCH_MAX_LAUNCH + launch) * round256((size_t)(cap + 1) * sizeof(int));
    }
};

// block-uniform values read from the work list, moved to scalar registers
__device__ inline int ufirst(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float ufirst(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
template <typename T> __device__ inline T* ufirst(T* p)
{
    const uint64_t u = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// largest f in [0, n) with prefix[f] <= v (prefix[0] = 0 <= v < prefix[n])
__device__ inline int batch_find(const int* prefix, int n, int v)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (prefix[m] <= v) lo = m;
        else hi = m;
    }
    return lo;
}

int blocks2d(int w, int h) { return ((w + BX - 1) / BX) * ((h + BY - 1) / BY); }
int rgba_blocks(const BatchFrame& f)
{
    const int W = f.front.A0.w, H = f.front.A0.h;
    return f.rgba_vec ? ((W / 4 + 63) / 64) * ((H + 3) / 4) : blocks2d(W, H);
}

__global__ __launch_bounds__(256) void k_front_batch(const unsigned char* __restrict__ wl)
{
    const BatchHeader* hd = reinterpret_cast<const BatchHeader*>(wl);
    const BatchLayout lay(hd->cap);
    const BatchFrame* fr = reinterpret_cast<const BatchFrame*>(wl + lay.frames);
    const int* pre = reinterpret_cast<const int*>(wl + lay.front);
    const int n = hd->n_frames, total = pre[n];
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const int f = batch_find(pre, n, vb);
        const FrontArgs& a = fr[f].front;
        const int local = vb - pre[f], gx = (a.A0.w + BX - 1) / BX;
        const int x = (local % gx) * BX + (int)(threadIdx.x % BX), y = (local / gx) * BY + (int)(threadIdx.x / BX);
        if (x < a.A0.w && y < a.A0.h) front_px(a, x, y);
    }
}

__global__ __launch_bounds__(256) void k_up2_batch(const unsigned char* __restrict__ wl)
{
    const BatchHeader* hd = reinterpret_cast<const BatchHeader*>(wl);
    const BatchLayout lay(hd->cap);
    const BatchFrame* fr = reinterpret_cast<const BatchFrame*>(wl + lay.frames);
    const int* pre = reinterpret_cast<const int*>(wl + lay.up);
    const int n = hd->n_frames, total = pre[n];
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const int f = batch_find(pre, n, vb);
        const BatchLevel& l0 = fr[f].lev[0];
        const BatchLevel& l1 = fr[f].lev[1];
        const View in = planar(l1.O[fr[f].up_idx], l1.w, l1.h, 3), out = planar(l0.O[0], l0.w, l0.h, 3);
        const int local = vb - pre[f], gx = (l1.w + BX - 1) / BX;
        const int x = (local % gx) * BX + (int)(threadIdx.x % BX), y = (local / gx) * BY + (int)(threadIdx.x / BX);
        if (x < l1.w && y < l1.h) up2_px<true>(in, out, x, y);
    }
}

__global__ __launch_bounds__(256) void k_rgba_batch(const unsigned char* __restrict__ wl)
{
    const BatchHeader* hd = reinterpret_cast<const BatchHeader*>(wl);
    const BatchLayout lay(hd->cap);
    const BatchFrame* fr = reinterpret_cast<const BatchFrame*>(wl + lay.frames);
    const int* pre = reinterpret_cast<const int*>(wl + lay.rgba);
    const int n = hd->n_frames, total = pre[n];
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const int f = batch_find(pre, n, vb);
        const BatchLevel& l0 = fr[f].lev[0];
        const View res = planar(l0.O[fr[f].out_idx], l0.w, l0.h, 3);
        const int local = vb - pre[f];
        if (fr[f].rgba_vec) {
            const int gx = (l0.w / 4 + 63) / 64;
            const int x0 = 4 * ((local % gx) * 64 + (int)(threadIdx.x % 64)), y = (local / gx) * 4 + (int)(threadIdx.x / 64);
            if (x0 < l0.w && y < l0.h) to_rgba4_px(res, fr[f].out_rgba, x0, y);
        } else {
            const int gx = (l0.w + BX - 1) / BX;
            const int x = (local % gx) * BX + (int)(threadIdx.x % BX), y = (local / gx) * BY + (int)(threadIdx.x / BX);
            if (x < l0.w && y < l0.h) to_rgba_px(res, fr[f].out_rgba, x, y);
        }
    }
}

// Launch `launch` of level slot `slot` (0: level 0, 1: level 1) for every
// frame of the batch: persistent blocks walk the tiles, XCD-aware (block b
// runs on XCD b % 8 and takes tiles from that XCD's contiguous share).
template <int R, bool FMA>
__global__ __launch_bounds__(SB_TW * SB_STRIPS) void k_consist_batch(const unsigned char* __restrict__ wl, int slot,
                                                                     int launch)
{
    __shared__ __attribute__((aligned(16))) float sO[2][SB_STRIPS][SB_TW + 1][sb_rs<R>()];
    const BatchHeader* hd = reinterpret_cast<const BatchHeader*>(wl);
    const BatchLayout lay(hd->cap);
    const BatchFrame* fr = reinterpret_cast<const BatchFrame*>(wl + lay.frames);
    const int* pre = reinterpret_cast<const int*>(wl + lay.tile_table(slot, launch));
    const int n = ufirst(hd->n_frames), total = ufirst(pre[n]);
    const int per_xcd = (total + 7) / 8;
    const int xcd = blockIdx.x % 8, nk = gridDim.x / 8;
    for (int i = blockIdx.x / 8; i < per_xcd; i += nk) {
        const int t = xcd * per_xcd + i;
        if (t >= total) break;
        const int f = ufirst(batch_find(pre, n, t));
        const BatchLevel& l = fr[f].lev[slot];
        const int iters = ufirst(l.iters);
        const int first = launch * BATCH_HALO;
        const int sweeps = min(BATCH_HALO, iters - first);
        const int in = ufirst(l.start) ^ (launch & 1), ui = launch & 1;
        __syncthreads();   // the previous tile's last readers of sO are done
        consist_tile<R, SB_STRIPS, FMA>(ufirst(l.L), ufirst(l.A), ufirst(l.Wt), ufirst(l.O[in]), ufirst(l.O[in ^ 1]),
                                        ufirst(l.U[ui]), ufirst(l.U[ui ^ 1]), ufirst(l.w), ufirst(l.h), sweeps, first,
                                        ufirst(fr[f].step), ufirst(fr[f].mom), BATCH_HALO, ufirst(l.tiles_x),
                                        ufirst(l.tiles_y), first + sweeps < iters ? 1 : 0, t - ufirst(pre[f]), sO);
    }
}

}  // namespace

size_t batch_worklist_bytes(int cap) { return BatchLayout(cap).total; }

void batch_tiles(int w, int h, int fma, int* tiles_x, int* tiles_y)
{
    *tiles_x = tile_count(w, SB_TW, BATCH_HALO);
    *tiles_y = tile_count(h, SB_STRIPS * (fma ? SB_R_FMA : SB_R), BATCH_HALO);
}

void batch_build(unsigned char* host, int cap, const BatchFrame* frames, int n, int fma)
{
    (void)fma;
    const BatchLayout lay(cap);
    BatchHeader* hd = reinterpret_cast<BatchHeader*>(host);
    hd->n_frames = n;
    hd->cap = cap;
    hd->n_launch[0] = hd->n_launch[1] = 0;
    BatchFrame* fr = reinterpret_cast<BatchFrame*>(host + lay.frames);
    int* front = reinterpret_cast<int*>(host + lay.front);
    int* up = reinterpret_cast<int*>(host + lay.up);
    int* rgba = reinterpret_cast<int*>(host + lay.rgba);
    front[0] = up[0] = rgba[0] = 0;
    for (int f = 0; f < n; ++f) {
        fr[f] = frames[f];
        front[f + 1] = front[f] + blocks2d(frames[f].front.A0.w, frames[f].front.A0.h);
        up[f + 1] = up[f] + (frames[f].levels == 2 ? blocks2d(frames[f].lev[1].w, frames[f].lev[1].h) : 0);
        rgba[f + 1] = rgba[f] + (frames[f].out_rgba ? rgba_blocks(frames[f]) : 0);
        for (int j = 0; j < frames[f].levels; ++j)
            hd->n_launch[j] = std::max(hd->n_launch[j], (frames[f].lev[j].iters + BATCH_HALO - 1) / BATCH_HALO);
    }
    for (int slot = 0; slot < 2; ++slot)
        for (int k = 0; k < hd->n_launch[slot]; ++k) {
            int* t = reinterpret_cast<int*>(host + lay.tile_table(slot, k));
            t[0] = 0;
            for (int f = 0; f < n; ++f) {
                const BatchLevel& l = frames[f].lev[slot];
                const bool on = slot < frames[f].levels && k * BATCH_HALO < l.iters;
                t[f + 1] = t[f] + (on ? 3 * l.tiles_x * l.tiles_y : 0);
            }
        }
}

int launch_batch(const unsigned char* dev, const int n_launch[2], int fma, int grid, hipStream_t s)
{
    grid = std::max(8, grid / 8 * 8);
    hipLaunchKernelGGL(k_front_batch, dim3(8 * grid), dim3(256), 0, s, dev);
    for (int k = 0; k < n_launch[1]; ++k) {
        if (fma) hipLaunchKernelGGL((k_consist_batch<SB_R_FMA, true>), dim3(grid), dim3(SB_TW * SB_STRIPS), 0, s, dev, 1, k);
        else hipLaunchKernelGGL((k_consist_batch<SB_R, false>), dim3(grid), dim3(SB_TW * SB_STRIPS), 0, s, dev, 1, k);
    }
    hipLaunchKernelGGL(k_up2_batch, dim3(4 * grid), dim3(256), 0, s, dev);
    for (int k = 0; k < n_launch[0]; ++k) {
        if (fma) hipLaunchKernelGGL((k_consist_batch<SB_R_FMA, true>), dim3(grid), dim3(SB_TW * SB_STRIPS), 0, s, dev, 0, k);
        else hipLaunchKernelGGL((k_consist_batch<SB_R, false>), dim3(grid), dim3(SB_TW * SB_STRIPS), 0, s, dev, 0, k);
    }
    hipLaunchKernelGGL(k_rgba_batch, dim3(4 * grid), dim3(256), 0, s, dev);
    return check_launch("batched frames");
}

}  // namespace vsc

// ===========================================================================
// C ABI: per-stage entry points
// ===========================================================================
using namespace vsc;

static int check_img(const vsc_image& im, const char* name, int min_ch)
{
    if (!im.data || im.width < 1 || im.height < 1 || im.channels < min_ch)
        return set_error(VSC_E_INVALID, std::string(name) + ": invalid image (null data, size < 1 or too few channels)");
    return VSC_OK;
}

static int same_size(const vsc_image& a, const vsc_image& b, const char* what)
{
    if (a.width != b.width || a.height != b.height)
        return set_error(VSC_E_INVALID, std::string(what) + ": image sizes differ");
    return VSC_OK;
}

extern "C" int vsc_stab_warp(vsc_image in, vsc_image flow, vsc_image out, vsc_hip_stream stream)
{
    int rc;
    if ((rc = check_img(in, "vsc_stab_warp in", 3)) || (rc = check_img(flow, "vsc_stab_warp flow", 2)) ||
        (rc = check_img(out, "vsc_stab_warp out", 3)))
        return rc;
    if ((rc = same_size(in, out, "vsc_stab_warp")) || (rc = same_size(flow, out, "vsc_stab_warp flow")))
        return rc;
    if (out.width < 2 || out.height < 2) return set_error(VSC_E_INVALID, "vsc_stab_warp: image must be at least 2x2");
    return launch_stab_warp(hwc(in), nullptr, hwc(flow), hwc(out), as_stream(stream));
}

extern "C" int vsc_stab_adap_comb(vsc_image crnt_in, vsc_image crnt_pr, vsc_image prev_warp_in,
                                  vsc_image prev_warp_pr, vsc_image next_warp_in, vsc_image next_warp_pr,
                                  vsc_image adap_cmb_in, vsc_image adap_cmb_pr, vsc_image last_stab_warp, float alpha,
                                  vsc_hip_stream stream)
{
    const vsc_image* ims[9] = {&crnt_in,      &crnt_pr,     &prev_warp_in, &prev_warp_pr,  &next_warp_in,
                               &next_warp_pr, &adap_cmb_in, &adap_cmb_pr,  &last_stab_warp};
    for (int i = 0; i < 9; ++i) {
        int rc = check_img(*ims[i], "vsc_stab_adap_comb", 3);
        if (!rc) rc = same_size(*ims[i], crnt_in, "vsc_stab_adap_comb");
        if (rc) return rc;
    }
    AdapArgs a{hwc(crnt_in),     hwc(crnt_pr),     hwc(prev_warp_in), hwc(prev_warp_pr), hwc(next_warp_in),
               hwc(next_warp_pr), hwc(adap_cmb_in), hwc(adap_cmb_pr),  hwc(last_stab_warp), alpha};
    hipLaunchKernelGGL(k_adap_comb, grid2d(crnt_in.width, crnt_in.height), dim3(BX, BY), 0, as_stream(stream), a);
    return check_launch("adap_comb");
}

extern "C" int vsc_stab_consist_wt(vsc_image adap_cmb_in, vsc_image crnt_in, vsc_image consist_wt, float beta,
                                   float gamma, vsc_hip_stream stream)
{
    int rc;
    if ((rc = check_img(adap_cmb_in, "vsc_stab_consist_wt", 3)) || (rc = check_img(crnt_in, "vsc_stab_consist_wt", 3)) ||
        (rc = check_img(consist_wt, "vsc_stab_consist_wt", 3)))
        return rc;
    if ((rc = same_size(adap_cmb_in, crnt_in, "vsc_stab_consist_wt")) ||
        (rc = same_size(consist_wt, crnt_in, "vsc_stab_consist_wt")))
        return rc;
    hipLaunchKernelGGL(k_consist_wt, grid2d(crnt_in.width, crnt_in.height), dim3(BX, BY), 0, as_stream(stream),
                       hwc(crnt_in), hwc(adap_cmb_in), hwc(consist_wt), beta, gamma);
    return check_launch("consist_wt");
}

extern "C" size_t vsc_stab_consist_out_workspace_size(int width, int height)
{
    if (width < 1 || height < 1) return 0;
    return (size_t)2 * width * height * 3 * sizeof(float);
}

extern "C" int vsc_stab_consist_out(vsc_image crnt_pr, vsc_image prev_stab_warp, vsc_image consist_wt, int num_iter,
                                    float step_size, float mom_fac, vsc_image consist_out, void* workspace,
                                    vsc_hip_stream stream)
{
    int rc;
    if ((rc = check_img(crnt_pr, "vsc_stab_consist_out", 3)) || (rc = check_img(prev_stab_warp, "vsc_stab_consist_out", 3)) ||
        (rc = check_img(consist_wt, "vsc_stab_consist_out", 3)) || (rc = check_img(consist_out, "vsc_stab_consist_out", 3)))
        return rc;
    if ((rc = same_size(crnt_pr, consist_out, "vsc_stab_consist_out")) ||
        (rc = same_size(prev_stab_warp, consist_out, "vsc_stab_consist_out")) ||
        (rc = same_size(consist_wt, consist_out, "vsc_stab_consist_out")))
        return rc;
    if (num_iter < 0) return set_error(VSC_E_INVALID, "vsc_stab_consist_out: num_iter < 0");
    if (num_iter == 0) return VSC_OK;
    hipStream_t s = as_stream(stream);
    const int W = consist_out.width, H = consist_out.height;
    const size_t n = (size_t)W * H * 3;
    float* ws = static_cast<float*>(workspace);
    bool owned = false;
    if (!ws) {
        rc = check_hip(hipMallocAsync(reinterpret_cast<void**>(&ws), vsc_stab_consist_out_workspace_size(W, H), s),
                       "vsc_stab_consist_out workspace");
        if (rc) return rc;
        owned = true;
    }
    View O0 = hwc(consist_out);
    View O1{ws, W, H, 3, 1, 3};
    View U{ws + n, W, H, 3, 1, 3};
    View P = hwc(crnt_pr), A = hwc(prev_stab_warp), Wt = hwc(consist_wt);
    View cur = O0, nxt = O1;
    for (int k = 0; k < num_iter && !rc; ++k) {
        rc = launch_consist_sweep(P, A, Wt, cur, nxt, U, step_size, mom_fac, k > 0 ? 1 : 0, s);
        std::swap(cur, nxt);
    }
    if (!rc && cur.d != consist_out.data) rc = launch_copy_view(cur, O0, s);
    if (owned) {
        int rc2 = check_hip(hipFreeAsync(ws, s), "vsc_stab_consist_out workspace free");
        if (!rc) rc = rc2;
    }
    return rc;
}

extern "C" int vsc_stab_blocked_plan(int width, int height, int num_iter, int fma, int* halo, int* tile_rows,
                                     int* tiles_x, int* tiles_y)
{
    if (!halo || !tile_rows || !tiles_x || !tiles_y) return set_error(VSC_E_INVALID, "vsc_stab_blocked_plan: NULL");
    if (width <= 0 || height <= 0 || num_iter < 1)
        return set_error(VSC_E_INVALID, "vsc_stab_blocked_plan: empty level or no sweeps");
    const BlockedPlan plan = plan_consist_blocked(width, height, num_iter, fma);
    const int th = plan.strips * (fma ? SB_R_FMA : SB_R);
    *halo = plan.halo;
    *tile_rows = th;
    *tiles_x = tile_count(width, SB_TW, plan.halo);
    *tiles_y = tile_count(height, th, plan.halo);
    return VSC_OK;
}

extern "C" int vsc_stab_tile_span(int length, int tile, int halo, int i, int* count, int* start, int* lo, int* hi)
{
    if (!count || !start || !lo || !hi) return set_error(VSC_E_INVALID, "vsc_stab_tile_span: NULL");
    if (length <= 0 || halo < 0 || tile - 2 * halo < 1)
        return set_error(VSC_E_INVALID, "vsc_stab_tile_span: no output per tile");
    *count = tile_count(length, tile, halo);
    if (i < 0 || i >= *count) return set_error(VSC_E_INVALID, "vsc_stab_tile_span: tile index out of range");
    const TileSpan t = tile_span(i, *count, length, tile, halo);
    *start = t.start;
    *lo = t.lo;
    *hi = t.hi;
    return VSC_OK;
}

extern "C" size_t vsc_stab_consist_out_planar_workspace_size(int width, int height)
{
    if (width < 1 || height < 1) return 0;
    return (size_t)4 * width * height * 3 * sizeof(float);
}

extern "C" int vsc_stab_consist_out_planar(const float* crnt_pr, const float* prev_stab_warp, const float* consist_wt,
                                           int width, int height, int num_iter, float step_size, float mom_fac,
                                           float* consist_out, void* workspace, vsc_hip_stream stream)
{
    if (!crnt_pr || !prev_stab_warp || !consist_wt || !consist_out)
        return set_error(VSC_E_INVALID, "vsc_stab_consist_out_planar: null pointer");
    if (width < 1 || height < 1) return set_error(VSC_E_INVALID, "vsc_stab_consist_out_planar: empty image");
    if (num_iter < 0) return set_error(VSC_E_INVALID, "vsc_stab_consist_out_planar: num_iter < 0");
    if (num_iter == 0) return VSC_OK;
    hipStream_t s = as_stream(stream);
    const size_t n = (size_t)width * height * 3;
    float* ws = static_cast<float*>(workspace);
    bool owned = false;
    int rc = VSC_OK;
    if (!ws) {
        rc = check_hip(hipMallocAsync(reinterpret_cast<void**>(&ws),
                                      vsc_stab_consist_out_planar_workspace_size(width, height), s),
                       "vsc_stab_consist_out_planar workspace");
        if (rc) return rc;
        owned = true;
    }
    float* O[2] = {consist_out, ws};
    float* U[2] = {ws + n, ws + 2 * n};
    float* lapP = ws + 3 * n;
    rc = launch_lap(crnt_pr, lapP, width, height, s);
    const BlockedPlan plan = plan_consist_blocked(width, height, num_iter, 0);
    int cur = 0;
    for (int k = 0; k < num_iter && !rc;) {
        const int m = std::min(plan.halo, num_iter - k);
        rc = launch_consist_blocked(lapP, prev_stab_warp, consist_wt, O[cur], O[cur ^ 1], U[cur], U[cur ^ 1], width,
                                    height, m, k, step_size, mom_fac, 0, plan, s, k + m < num_iter);
        cur ^= 1;
        k += m;
    }
    if (!rc && cur == 1)
        rc = check_hip(hipMemcpyAsync(consist_out, ws, n * sizeof(float), hipMemcpyDeviceToDevice, s),
                       "vsc_stab_consist_out_planar copy");
    if (owned) {
        int rc2 = check_hip(hipFreeAsync(ws, s), "vsc_stab_consist_out_planar workspace free");
        if (!rc) rc = rc2;
    }
    return rc;
}

extern "C" int vsc_stab_bilinear(vsc_image in, vsc_image out, vsc_hip_stream stream)
{
    int rc;
    if ((rc = check_img(in, "vsc_stab_bilinear in", 1)) || (rc = check_img(out, "vsc_stab_bilinear out", 1))) return rc;
    if (out.channels > in.channels)
        return set_error(VSC_E_INVALID, "vsc_stab_bilinear: output has more channels than input");
    return launch_bilinear(hwc(in), hwc(out), as_stream(stream));
}

extern "C" int vsc_average_frames(vsc_image processed_prev, vsc_image processed_next, vsc_image out,
                                  vsc_hip_stream stream)
{
    int rc;
    if ((rc = check_img(processed_prev, "vsc_average_frames", 3)) || (rc = check_img(processed_next, "vsc_average_frames", 3)) ||
        (rc = check_img(out, "vsc_average_frames", 3)))
        return rc;
    if ((rc = same_size(processed_prev, out, "vsc_average_frames")) ||
        (rc = same_size(processed_next, out, "vsc_average_frames")))
        return rc;
    hipLaunchKernelGGL(k_average, grid2d(out.width, out.height), dim3(BX, BY), 0, as_stream(stream),
                       hwc(processed_prev), hwc(processed_next), hwc(out));
    return check_launch("average_frames");
}

extern "C" int vsc_rgba_to_float(const uint8_t* rgba, vsc_image out, vsc_hip_stream stream)
{
    int rc;
    if (!rgba) return set_error(VSC_E_INVALID, "vsc_rgba_to_float: null input");
    if ((rc = check_img(out, "vsc_rgba_to_float", 3))) return rc;
    return launch_rgba_to_view(rgba, hwc(out), as_stream(stream));
}

extern "C" int vsc_float_to_rgba(vsc_image in, uint8_t* rgba, vsc_hip_stream stream)
{
    int rc;
    if (!rgba) return set_error(VSC_E_INVALID, "vsc_float_to_rgba: null output");
    if ((rc = check_img(in, "vsc_float_to_rgba", 3))) return rc;
    return launch_to_rgba(hwc(in), rgba, as_stream(stream));
}
