// Stabilization kernels for gfx950: the per-stage entry points of
// src/stabilization/flowconsistency.cuh and gpuimage.cu, the fused front pass
// and the consistency solver (single-sweep and temporally blocked forms).
#include "vsc_internal.h"

#include <algorithm>

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

int launch_bilinear(View in, View out, hipStream_t s)
{
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

__global__ __launch_bounds__(256) void k_to_rgba(View in, uint8_t* out)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    uchar4 v;
    v.x = to_u8(in.rd(x, y, 0));
    v.y = to_u8(in.rd(x, y, 1));
    v.z = to_u8(in.rd(x, y, 2));
    v.w = 1;
    reinterpret_cast<uchar4*>(out)[(long long)y * in.w + x] = v;
}

int launch_to_rgba(View in, uint8_t* out, hipStream_t s)
{
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
// A 512-thread block owns a TW x TH tile of one channel plane (TW = 128
// columns, TH = 4*R rows) that overlaps its neighbours by a halo of HALO
// pixels.  Thread (tx, strip) keeps the R pixels of column tx in rows
// strip*R .. strip*R+R-1 in registers: O, U, lapP, A, w.  Each sweep writes O
// to one of two LDS planes (ping-pong, one barrier per sweep), reads the left
// and right neighbours from LDS and the up/down neighbours from its own
// registers (LDS only at the strip ends).  O and U are read with halo from
// the *in* planes and written to the *out* planes (neighbouring blocks read the
// pixels a block writes).  After `sweeps` <= HALO sweeps the
// inner (TW-2*HALO) x (TH-2*HALO) pixels are exact and written back; halo
// pixels whose out-of-tile neighbours are missing never reach the output.
// Arithmetic per pixel and sweep is that of k_consist_sweep (lapP is computed
// once per launch with the same operations), so the result is identical.
// ---------------------------------------------------------------------------
static constexpr int SB_TW = 128;
static constexpr int SB_THREADS = 512;

template <int R, int HALO>
__global__ __launch_bounds__(SB_THREADS) void k_consist_blocked(const float* __restrict__ P, const float* __restrict__ A,
                                                                const float* __restrict__ Wt, const float* Oin,
                                                                float* Oout, const float* Uin, float* Uout, int W,
                                                                int H, int sweeps, int first_sweep, float step,
                                                                float mom)
{
    constexpr int TW = SB_TW;
    constexpr int STRIPS = SB_THREADS / TW;
    constexpr int TH = STRIPS * R;
    constexpr int OW = TW - 2 * HALO, OH = TH - 2 * HALO;
    __shared__ float sO[2][TH][TW];

    const long long plane = (long long)W * H;
    const int c = blockIdx.z;
    const float* __restrict__ Pc = P + c * plane;
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
    const int ty0 = strip * R;
    const int gx0 = blockIdx.x * OW - HALO;
    const int gy0 = blockIdx.y * OH - HALO;
    const int gx = gx0 + tx;
    const bool colin = gx >= 0 && gx < W;
    const bool hr = (gx + 1) < (W - 1), hl = (gx - 1) >= 0;
    const bool interior = gx0 >= 1 && gx0 + TW - 1 <= W - 3 && gy0 >= 1 && gy0 + TH - 1 <= H - 3;

    float o[R], u[R], a[R], w[R], lp[R];
    // 32-bit element offsets (a plane holds < 2^31 pixels); row r of the strip
    // is at base + r*W, so no per-row address has to stay live.
    const int base = (gy0 + ty0) * W + gx;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gy = gy0 + ty0 + r;
        const bool in = colin && gy >= 0 && gy < H;
        const int i = base + r * W;
        o[r] = in ? Oi[i] : 0.0f;
        u[r] = (in && first_sweep > 0) ? Ui[i] : 0.0f;
        a[r] = in ? Ac[i] : 0.0f;
        w[r] = in ? Wc[i] : 0.0f;
        float l = 0.0f;
        if (in) {
            int cnt = 0;
            if (hr) { l += Pc[i + 1]; cnt += 1; }
            if (hl) { l += Pc[i - 1]; cnt += 1; }
            if ((gy + 1) < (H - 1)) { l += Pc[i + W]; cnt += 1; }
            if ((gy - 1) >= 0) { l += Pc[i - W]; cnt += 1; }
            l -= (float)cnt * Pc[i];
        }
        lp[r] = l;
    }

    const int txl = tx > 0 ? tx - 1 : 0;
    const int txr = tx < TW - 1 ? tx + 1 : TW - 1;
    const int rowU = ty0 > 0 ? ty0 - 1 : 0;
    const int rowD = ty0 + R < TH ? ty0 + R : TH - 1;

    for (int s = 0; s < sweeps; ++s) {
        float (*buf)[TW] = sO[s & 1];
#pragma unroll
        for (int r = 0; r < R; ++r) buf[ty0 + r][tx] = o[r];
        __syncthreads();
        const bool momk = (first_sweep + s) > 0;
        float up_old = buf[rowU][tx];
        const float below = buf[rowD][tx];
        if (interior) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float cur = o[r];
                const float dn = (r < R - 1) ? o[r + 1] : below;
                float lo = buf[ty0 + r][txr];
                lo += buf[ty0 + r][txl];
                lo += dn;
                lo += up_old;
                lo -= 4.0f * cur;
                const float t1 = w[r] * (cur - a[r]);
                const float t2 = lo - lp[r];
                const float g = t2 - t1;
                const float sg = step * g;
                o[r] = momk ? (cur + sg) + mom * u[r] : cur + sg;
                u[r] = sg;
                up_old = cur;
            }
        } else {
            // image-border tiles.  Column conditions are per lane; the row
            // conditions reduce to two wave-uniform bounds: row r has an up
            // neighbour iff r >= rU and a down neighbour iff r <= rD.
            const int gys = gy0 + ty0;
            int rU = 1 - gys;
            int rD = (H - 3) - gys;
            // opaque per sweep: keeps the compiler from hoisting 2*R row
            // predicates out of the sweep loop into scalar registers
            asm volatile("" : "+s"(rU), "+s"(rD));
            const int ccol = (hr ? 1 : 0) + (hl ? 1 : 0);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float cur = o[r];
                const float dn = (r < R - 1) ? o[r + 1] : below;
                const float vr = buf[ty0 + r][txr];
                const float vl = buf[ty0 + r][txl];
                float lo = 0.0f;
                lo = hr ? lo + vr : lo;
                lo = hl ? lo + vl : lo;
                int cnt = ccol;
                if (r <= rD) { lo += dn; cnt += 1; }
                if (r >= rU) { lo += up_old; cnt += 1; }
                lo -= (float)cnt * cur;
                const float t1 = w[r] * (cur - a[r]);
                const float t2 = lo - lp[r];
                const float g = t2 - t1;
                const float sg = step * g;
                o[r] = momk ? (cur + sg) + mom * u[r] : cur + sg;
                u[r] = sg;
                up_old = cur;
            }
        }
    }

    if (tx >= HALO && tx < TW - HALO && colin) {
        const int wbase = (gy0 + ty0) * W + gx;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int ty = ty0 + r;
            const int gy = gy0 + ty;
            if (ty >= HALO && ty < TH - HALO && gy >= 0 && gy < H) {
                Oo[wbase + r * W] = o[r];
                Uo[wbase + r * W] = u[r];
            }
        }
    }
}

static constexpr int SB_R = 24;
static constexpr int SB_HALO = 8;

int consist_blocked_halo() { return SB_HALO; }

int launch_consist_blocked(const float* P, const float* A, const float* Wt, const float* Oin, float* Oout,
                           const float* Uin, float* Uout, int W, int H, int sweeps, int first_sweep, float step,
                           float mom, hipStream_t s)
{
    if (sweeps < 1 || sweeps > SB_HALO) return set_error(VSC_E_INVALID, "consist_blocked: sweeps out of range");
    constexpr int TH = (SB_THREADS / SB_TW) * SB_R;
    constexpr int OW = SB_TW - 2 * SB_HALO, OH = TH - 2 * SB_HALO;
    dim3 grid((W + OW - 1) / OW, (H + OH - 1) / OH, 3);
    // algorithmic bytes: read O, U, P, A, w and write O, U once per pixel and channel
    const int slot = prof_start(PROF_CONSIST_BLOCKED, 7.0 * 3 * 4 * (double)W * H, s);
    hipLaunchKernelGGL((k_consist_blocked<SB_R, SB_HALO>), grid, dim3(SB_THREADS), 0, s, P, A, Wt, Oin, Oout, Uin, Uout,
                       W, H, sweeps, first_sweep, step, mom);
    prof_stop(slot, s);
    return check_launch("consist_blocked");
}

// ---------------------------------------------------------------------------
// fused front pass: 5 warps + adaptive combination + consistency weights for
// one pixel (all 3 channels), reading the u8 frames, the flows and the previous
// output once; writes level-0 P, A, w (planar) and, for an exact 2x pyramid
// (even W and H: kernel_bilinear is then pure decimation, fx = fy = 0),
// the level-1 P, A, w and the level-1 initial solution O1 = P1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_front(FrontArgs a)
{
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    const int W = a.P0.w, H = a.P0.h;
    if (x >= W || y >= H) return;
    const WarpTap tb = warp_tap(x, y, a.bwd.rd(x, y, 0), a.bwd.rd(x, y, 1), W, H);
    const WarpTap tf = warp_tap(x, y, a.fwd.rd(x, y, 0), a.fwd.rd(x, y, 1), W, H);
    const bool dec = a.decimate && ((x | y) & 1) == 0;
    const int x1 = x >> 1, y1 = y >> 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float pwi = warp_sample(a.oPrev, tb, c);
        const float pwp = warp_sample(a.pPrev, tb, c);
        const float nwi = warp_sample(a.oNext, tf, c);
        const float nwp = warp_sample(a.pNext, tf, c);
        const float lw = warp_sample(a.last, tb, c);
        const float ci = a.oCur.rd(x, y, c);
        const float cp = a.pCur.rd(x, y, c);
        float adIn, adPr;
        adap_comb(ci, cp, pwi, pwp, nwi, nwp, lw, a.alpha, adIn, adPr);
        const float wt = consist_wt(ci, adIn, a.beta, a.gamma);
        a.P0.wr(x, y, c, cp);
        a.A0.wr(x, y, c, adPr);
        a.W0.wr(x, y, c, wt);
        if (dec && x1 < a.P1.w && y1 < a.P1.h) {
            a.P1.wr(x1, y1, c, cp);
            a.A1.wr(x1, y1, c, adPr);
            a.W1.wr(x1, y1, c, wt);
            a.O1.wr(x1, y1, c, cp);
        }
    }
}

int launch_front_pass(const FrontArgs& a, hipStream_t s)
{
    // algorithmic bytes per pixel: 6 RGBA8 frames (24), previous output (12),
    // 2 flows (2 * 4 * channels), level-0 P, A, w (36) and, when decimating,
    // a quarter pixel of level-1 P, A, w, O (48 / 4)
    const double px = (double)a.P0.w * a.P0.h;
    const double bytes = px * (24.0 + 12.0 + 4.0 * (a.fwd.ch + a.bwd.ch) + 36.0 + (a.decimate ? 12.0 : 0.0));
    const int slot = prof_start(PROF_FRONT_PASS, bytes, s);
    hipLaunchKernelGGL(k_front, grid2d(a.P0.w, a.P0.h), dim3(BX, BY), 0, s, a);
    prof_stop(slot, s);
    return check_launch("front_pass");
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
