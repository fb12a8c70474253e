// Per-stream stabilizer: the device-resident state and frame schedule of
// VideoStabilizer (src/stabilization/videostabilizer.cpp:46-265) behind the
// C ABI.  All state lives in one device allocation made at create; a frame
// allocates nothing (the staged parity mode allocates its HWC images lazily on
// its first use).
#include "vsc_internal.h"

#include <algorithm>
#include <new>
#include <string>
#include <vector>

using namespace vsc;

struct vsc_stream {
    int W = 0, H = 0, fw = 0, fh = 0;
    int nlev = 0;                      // pyramid levels allocated (<= VSC_MAX_LEVELS)
    int lw[VSC_MAX_LEVELS] = {}, lh[VSC_MAX_LEVELS] = {};
    float* P[VSC_MAX_LEVELS] = {};     // planar 3-channel level images
    float* L[VSC_MAX_LEVELS] = {};     // lapP: Laplacian of P, constant per level
    float* A[VSC_MAX_LEVELS] = {};
    float* Wt[VSC_MAX_LEVELS] = {};
    float* O[VSC_MAX_LEVELS][2] = {};  // solution ping-pong
    float* U[VSC_MAX_LEVELS][2] = {};  // momentum ping-pong
    float* flow_up[2] = {};            // fwd, bwd at frame size (FLOWDOWNSCALE), HWC 3 ch
    int last_idx = 0;                  // O[0][last_idx] holds the previous stabilized frame
    bool has_last = false;
    void* block = nullptr;
    size_t bytes = 0;
    // staged (per-stage reference sequence) images, allocated on first use
    void* staged_block = nullptr;
    float* st[32] = {};
};

namespace {

int level_dims(int W, int H, int* lw, int* lh)
{
    int n = 0, w = W, h = H;
    while (n < VSC_MAX_LEVELS && w >= 1 && h >= 1) {
        lw[n] = w;
        lh[n] = h;
        ++n;
        w /= 2;
        h /= 2;
    }
    return n;
}

// Bytes one carved piece of n floats occupies: every piece starts on a 256-byte
// boundary.  footprint() and vsc_stream_create's carve both use it, so the
// block's size is exactly the sum of the pieces carved from it.
inline size_t piece_bytes(size_t nfloats) { return ((nfloats * sizeof(float) + 255) / 256) * 256; }

size_t footprint(int W, int H, int fw, int fh, int* nlev_out = nullptr)
{
    int lw[VSC_MAX_LEVELS], lh[VSC_MAX_LEVELS];
    const int n = level_dims(W, H, lw, lh);
    size_t bytes = 0;
    for (int j = 0; j < n; ++j) bytes += 8 * piece_bytes((size_t)3 * lw[j] * lh[j]);
    if (fw != W || fh != H) bytes += 2 * piece_bytes((size_t)3 * W * H);
    if (nlev_out) *nlev_out = n;
    return bytes;
}

inline View pl(float* d, int w, int h) { return planar(d, w, h, 3); }

}  // namespace

extern "C" size_t vsc_stream_footprint(int width, int height, int flow_width, int flow_height)
{
    if (width < 4 || height < 4 || flow_width < 2 || flow_height < 2) return 0;
    return footprint(width, height, flow_width, flow_height);
}

extern "C" int vsc_stream_create(int width, int height, int flow_width, int flow_height, vsc_stream** out)
{
    if (!out) return set_error(VSC_E_INVALID, "vsc_stream_create: null output");
    *out = nullptr;
    if (width < 4 || height < 4)
        return set_error(VSC_E_INVALID, "vsc_stream_create: frame must be at least 4x4");
    if (flow_width < 2 || flow_height < 2 || flow_width > width || flow_height > height)
        return set_error(VSC_E_INVALID, "vsc_stream_create: flow size must be in [2, frame size]");
    vsc_stream* s = new (std::nothrow) vsc_stream();
    if (!s) return set_error(VSC_E_NOMEM, "vsc_stream_create: host allocation failed");
    s->W = width;
    s->H = height;
    s->fw = flow_width;
    s->fh = flow_height;
    s->bytes = footprint(width, height, flow_width, flow_height, &s->nlev);
    level_dims(width, height, s->lw, s->lh);
    int rc = check_hip(hipMalloc(&s->block, s->bytes), "vsc_stream_create");
    if (rc) {
        delete s;
        return rc;
    }
    char* p = static_cast<char*>(s->block);
    auto take = [&](size_t nfloats) {
        float* r = reinterpret_cast<float*>(p);
        p += piece_bytes(nfloats);
        return r;
    };
    for (int j = 0; j < s->nlev; ++j) {
        const size_t n = (size_t)3 * s->lw[j] * s->lh[j];
        s->P[j] = take(n);
        s->L[j] = take(n);
        s->A[j] = take(n);
        s->Wt[j] = take(n);
        s->O[j][0] = take(n);
        s->O[j][1] = take(n);
        s->U[j][0] = take(n);
        s->U[j][1] = take(n);
    }
    if (flow_width != width || flow_height != height) {
        s->flow_up[0] = take((size_t)3 * width * height);
        s->flow_up[1] = take((size_t)3 * width * height);
    }
    if ((size_t)(p - static_cast<char*>(s->block)) != s->bytes) {
        (void)hipFree(s->block);
        delete s;
        return set_error(VSC_E_INVALID, "vsc_stream_create: internal error: carve does not match the footprint");
    }
    *out = s;
    return VSC_OK;
}

extern "C" int vsc_stream_destroy(vsc_stream* s)
{
    if (!s) return VSC_OK;
    int rc = VSC_OK;
    if (s->staged_block) rc = check_hip(hipFree(s->staged_block), "vsc_stream_destroy");
    if (s->block) {
        int rc2 = check_hip(hipFree(s->block), "vsc_stream_destroy");
        if (!rc) rc = rc2;
    }
    delete s;
    return rc;
}

extern "C" int vsc_stream_set_last_rgba(vsc_stream* s, const uint8_t* processed_rgba, vsc_hip_stream stream)
{
    if (!s || !processed_rgba) return set_error(VSC_E_INVALID, "vsc_stream_set_last_rgba: null argument");
    s->last_idx = 0;
    int rc = launch_rgba_to_view(processed_rgba, pl(s->O[0][0], s->W, s->H), as_stream(stream));
    if (!rc) s->has_last = true;
    return rc;
}

extern "C" int vsc_stream_set_last(vsc_stream* s, vsc_image last, vsc_hip_stream stream)
{
    if (!s || !last.data) return set_error(VSC_E_INVALID, "vsc_stream_set_last: null argument");
    if (last.width != s->W || last.height != s->H || last.channels < 3)
        return set_error(VSC_E_INVALID, "vsc_stream_set_last: size mismatch");
    s->last_idx = 0;
    int rc = launch_copy_view(hwc(last), pl(s->O[0][0], s->W, s->H), as_stream(stream));
    if (!rc) s->has_last = true;
    return rc;
}

extern "C" int vsc_stream_copy_last(vsc_stream* s, vsc_image out, vsc_hip_stream stream)
{
    if (!s || !out.data) return set_error(VSC_E_INVALID, "vsc_stream_copy_last: null argument");
    if (!s->has_last) return set_error(VSC_E_INVALID, "vsc_stream_copy_last: no frame yet");
    if (out.width != s->W || out.height != s->H || out.channels < 3)
        return set_error(VSC_E_INVALID, "vsc_stream_copy_last: size mismatch");
    return launch_copy_view(pl(s->O[0][s->last_idx], s->W, s->H), hwc(out), as_stream(stream));
}

namespace {

int check_frame_in(const vsc_stream* s, const vsc_frame_in* in)
{
    if (!in) return set_error(VSC_E_INVALID, "vsc_stab_frame: null frame input");
    const void* ptrs[8] = {in->orig_prev, in->orig_cur, in->orig_next, in->proc_prev,
                           in->proc_cur,  in->proc_next, in->flow_fwd, in->flow_bwd};
    for (const void* p : ptrs)
        if (!p) return set_error(VSC_E_INVALID, "vsc_stab_frame: null frame or flow pointer");
    if (in->flow_channels != 2 && in->flow_channels != 3)
        return set_error(VSC_E_INVALID, "vsc_stab_frame: flow_channels must be 2 or 3");
    const void* frames[6] = {in->orig_prev, in->orig_cur, in->orig_next, in->proc_prev, in->proc_cur, in->proc_next};
    for (const void* f : frames)
        if (reinterpret_cast<uintptr_t>(f) % 4 != 0)
            return set_error(VSC_E_INVALID, "vsc_stab_frame: RGBA8 frames must be 4-byte aligned");
    if (!s->has_last)
        return set_error(VSC_E_INVALID, "vsc_stab_frame: no previous stabilized frame (call vsc_stream_set_last*)");
    return VSC_OK;
}

int check_params(const vsc_stream* s, const vsc_params& p)
{
    if (p.pyramid_levels < 1 || p.pyramid_levels > s->nlev)
        return set_error(VSC_E_INVALID, "vsc_stab_frame: pyramid_levels out of range for this frame size");
    if (p.num_iter < 0) return set_error(VSC_E_INVALID, "vsc_stab_frame: num_iter < 0");
    return VSC_OK;
}

}  // namespace

// The checks vsc_stab_frame makes before it enqueues anything, for callers
// that change their own state first (the pipeline, the server): parameters
// (NULL = defaults) valid for this stream, and a previous frame set.
namespace vsc {

int stream_params_check(const vsc_stream* s, const vsc_params* params)
{
    if (!s) return set_error(VSC_E_INVALID, "vsc_stab_frame: null stream");
    vsc_params p;
    if (params) p = *params;
    else vsc_default_params(&p);
    return check_params(s, p);
}

int stream_frame_check(const vsc_stream* s, const vsc_params* params)
{
    int rc = stream_params_check(s, params);
    if (rc) return rc;
    if (!s->has_last)
        return set_error(VSC_E_INVALID, "vsc_stab_frame: no previous stabilized frame (call vsc_stream_set_last*)");
    return VSC_OK;
}

}  // namespace vsc

namespace {

// Flows at frame size: either the caller's buffers or the get_bilinear
// upsample of the FLOWDOWNSCALE flows (flowmodel.cpp:158-161; vectors are not
// rescaled, as in the reference).
int prepare_flows(vsc_stream* s, const vsc_frame_in* in, View* fwd, View* bwd, hipStream_t st)
{
    const int fc = in->flow_channels;
    View f{const_cast<float*>(in->flow_fwd), s->fw, s->fh, fc, 1, fc};
    View b{const_cast<float*>(in->flow_bwd), s->fw, s->fh, fc, 1, fc};
    if (s->fw == s->W && s->fh == s->H) {
        *fwd = f;
        *bwd = b;
        return VSC_OK;
    }
    *fwd = View{s->flow_up[0], s->W, s->H, fc, 1, fc};
    *bwd = View{s->flow_up[1], s->W, s->H, fc, 1, fc};
    int rc = launch_bilinear(f, *fwd, st);
    if (!rc) rc = launch_bilinear(b, *bwd, st);
    return rc;
}

// Pyramid solve, levels L-1 .. 0 (videostabilizer.cpp:219-227) with the
// temporally blocked solver.  Level j initial solution: O[j][0] (level L-1)
// or bilinear(up) of the level above.  Returns the index of O[0] holding the
// result.
int solve_pyramid(vsc_stream* s, int L, const vsc_params& p, int fma, hipStream_t st, int* final_idx)
{
    int cur = 0;
    for (int j = L - 1; j >= 0; --j) {
        const int w = s->lw[j], h = s->lh[j];
        const int iters = p.num_iter / (j + 1);
        int rc;
        if (j != L - 1) {
            cur = 0;
            rc = launch_bilinear(pl(s->O[j + 1][*final_idx], s->lw[j + 1], s->lh[j + 1]), pl(s->O[j][0], w, h), st);
            if (rc) return rc;
        }
        const BlockedPlan plan = plan_consist_blocked(w, h, iters, fma);
        int ucur = 0;
        for (int k = 0; k < iters;) {
            const int n = std::min(plan.halo, iters - k);
            rc = launch_consist_blocked(s->L[j], s->A[j], s->Wt[j], s->O[j][cur], s->O[j][cur ^ 1], s->U[j][ucur],
                                        s->U[j][ucur ^ 1], w, h, n, k, p.step_size, p.mom_fac, fma, plan, st,
                                        k + n < iters);
            if (rc) return rc;
            cur ^= 1;
            ucur ^= 1;
            k += n;
        }
        *final_idx = cur;
    }
    return VSC_OK;
}

// fma: solver in FMA form (VSC_FRAME_FMA).  Its per-pixel constants are
// written by the front pass when it produces every level (one level, or two
// with exact decimation); otherwise the levels are built as for the reference
// order and converted in place afterwards.
int frame_fused(vsc_stream* s, const vsc_frame_in* in, const vsc_params& p, uint8_t* out_rgba, float* out_float,
                int fma, hipStream_t st)
{
    View fwd, bwd;
    int rc = prepare_flows(s, in, &fwd, &bwd, st);
    if (rc) return rc;
    const int L = p.pyramid_levels;
    const int W = s->W, H = s->H;
    FrontArgs a;
    a.oPrev = Rgba{in->orig_prev, W, H};
    a.oCur = Rgba{in->orig_cur, W, H};
    a.oNext = Rgba{in->orig_next, W, H};
    a.pPrev = Rgba{in->proc_prev, W, H};
    a.pCur = Rgba{in->proc_cur, W, H};
    a.pNext = Rgba{in->proc_next, W, H};
    a.last = pl(s->O[0][s->last_idx], W, H);
    a.fwd = fwd;
    a.bwd = bwd;
    a.L0 = pl(s->L[0], W, H);
    a.A0 = pl(s->A[0], W, H);
    a.W0 = pl(s->Wt[0], W, H);
    a.P0 = pl(s->P[0], W, H);
    // kernel_bilinear from W x H to W/2 x H/2 is exact decimation when W, H are even
    a.decimate = (L >= 2 && (W % 2) == 0 && (H % 2) == 0) ? 1 : 0;
    // P itself is needed to initialise a single level or to build levels
    // that the front pass does not decimate
    a.write_p = (L == 1 || L > 2 || !a.decimate) ? 1 : 0;
    if (L >= 2) {
        const int w1 = s->lw[1], h1 = s->lh[1];
        a.L1 = pl(s->L[1], w1, h1);
        a.A1 = pl(s->A[1], w1, h1);
        a.W1 = pl(s->Wt[1], w1, h1);
        a.O1 = pl(s->O[1][0], w1, h1);
        a.P1 = pl(s->P[1], w1, h1);
    } else {
        a.L1 = a.A1 = a.W1 = a.O1 = a.P1 = a.A0;
    }
    a.alpha = p.alpha;
    a.beta = p.beta;
    a.gamma = p.gamma;
    const bool direct = fma && (L == 1 || (L == 2 && a.decimate));
    a.fma = direct ? 1 : 0;
    a.step = p.step_size;
    if ((rc = launch_front_pass(a, st))) return rc;
    // remaining pyramid levels (videostabilizer.cpp:203-217): level j from j-1,
    // the initial solution O_j = bilinear(O_{j-1}) = bilinear(P_{j-1}) = P_j
    for (int j = 1; j < L; ++j) {
        if (j == 1 && a.decimate) continue;
        const int w0 = s->lw[j - 1], h0 = s->lh[j - 1], w1 = s->lw[j], h1 = s->lh[j];
        if ((rc = launch_bilinear(pl(s->P[j - 1], w0, h0), pl(s->P[j], w1, h1), st))) return rc;
        if ((rc = launch_bilinear(pl(s->A[j - 1], w0, h0), pl(s->A[j], w1, h1), st))) return rc;
        if ((rc = launch_bilinear(pl(s->Wt[j - 1], w0, h0), pl(s->Wt[j], w1, h1), st))) return rc;
        if ((rc = launch_bilinear(pl(s->P[j - 1], w0, h0), pl(s->O[j][0], w1, h1), st))) return rc;
        if ((rc = launch_lap(s->P[j], s->L[j], w1, h1, st))) return rc;
    }
    if (fma && !direct)
        for (int j = 0; j < L; ++j)
            if ((rc = launch_fma_consts(s->L[j], s->A[j], s->Wt[j], s->lw[j], s->lh[j], p.step_size, st))) return rc;
    if (L == 1) {
        // single level: the solution starts at the processed frame (videostabilizer.cpp:209)
        if ((rc = launch_copy_view(pl(s->P[0], W, H), pl(s->O[0][0], W, H), st))) return rc;
    }
    int fin = 0;
    if ((rc = solve_pyramid(s, L, p, fma, st, &fin))) return rc;
    const View res = pl(s->O[0][fin], W, H);
    if (out_rgba && (rc = launch_to_rgba(res, out_rgba, st))) return rc;
    if (out_float && (rc = launch_copy_view(res, View{out_float, W, H, 3, 1, 3}, st))) return rc;
    s->last_idx = fin;
    return VSC_OK;
}

// The reference's stage-by-stage sequence on HWC images (one launch per
// stage, single-sweep Jacobi solver).  Used to check the fused path.
enum {
    ST_OPREV, ST_OCUR, ST_ONEXT, ST_PPREV, ST_PCUR, ST_PNEXT, ST_LAST, ST_PWI, ST_PWP, ST_NWI, ST_NWP, ST_LW,
    ST_ADIN, ST_ADPR, ST_WT, ST_SOLVE_WS, ST_LEVEL0  // then 4 images per level (P, A, W, O)
};

int ensure_staged(vsc_stream* s)
{
    if (s->staged_block) return VSC_OK;
    const size_t n0 = (size_t)3 * s->W * s->H;
    size_t total = 0;
    auto rnd = [](size_t b) { return ((b + 255) / 256) * 256; };
    total += (size_t)ST_SOLVE_WS * rnd(n0 * 4);
    total += rnd(2 * n0 * 4);
    for (int j = 0; j < s->nlev; ++j) total += 4 * rnd((size_t)3 * s->lw[j] * s->lh[j] * 4);
    int rc = check_hip(hipMalloc(&s->staged_block, total), "vsc_stab_frame (staged buffers)");
    if (rc) return rc;
    char* p = static_cast<char*>(s->staged_block);
    for (int i = 0; i < ST_SOLVE_WS; ++i) { s->st[i] = reinterpret_cast<float*>(p); p += rnd(n0 * 4); }
    s->st[ST_SOLVE_WS] = reinterpret_cast<float*>(p);
    p += rnd(2 * n0 * 4);
    for (int j = 0; j < s->nlev; ++j)
        for (int k = 0; k < 4; ++k) {
            s->st[ST_LEVEL0 + 4 * j + k] = reinterpret_cast<float*>(p);
            p += rnd((size_t)3 * s->lw[j] * s->lh[j] * 4);
        }
    return VSC_OK;
}

int frame_staged(vsc_stream* s, const vsc_frame_in* in, const vsc_params& p, uint8_t* out_rgba, float* out_float,
                 hipStream_t st)
{
    int rc = ensure_staged(s);
    if (rc) return rc;
    const int W = s->W, H = s->H;
    auto img = [&](int i, int w, int h) { return vsc_image{s->st[i], w, h, 3}; };
    auto full = [&](int i) { return img(i, W, H); };
    View fwdv, bwdv;
    if ((rc = prepare_flows(s, in, &fwdv, &bwdv, st))) return rc;
    vsc_image fwd{fwdv.d, W, H, fwdv.ch}, bwd{bwdv.d, W, H, bwdv.ch};
    const uint8_t* frames[6] = {in->orig_prev, in->orig_cur, in->orig_next, in->proc_prev, in->proc_cur, in->proc_next};
    vsc_hip_stream vs = reinterpret_cast<vsc_hip_stream>(st);
    for (int i = 0; i < 6; ++i)
        if ((rc = vsc_rgba_to_float(frames[i], full(ST_OPREV + i), vs))) return rc;
    if ((rc = launch_copy_view(pl(s->O[0][s->last_idx], W, H), hwc(full(ST_LAST)), st))) return rc;
    // videostabilizer.cpp:182-198
    if ((rc = vsc_stab_warp(full(ST_OPREV), bwd, full(ST_PWI), vs))) return rc;
    if ((rc = vsc_stab_warp(full(ST_PPREV), bwd, full(ST_PWP), vs))) return rc;
    if ((rc = vsc_stab_warp(full(ST_ONEXT), fwd, full(ST_NWI), vs))) return rc;
    if ((rc = vsc_stab_warp(full(ST_PNEXT), fwd, full(ST_NWP), vs))) return rc;
    if ((rc = vsc_stab_warp(full(ST_LAST), bwd, full(ST_LW), vs))) return rc;
    if ((rc = vsc_stab_adap_comb(full(ST_OCUR), full(ST_PCUR), full(ST_PWI), full(ST_PWP), full(ST_NWI), full(ST_NWP),
                                 full(ST_ADIN), full(ST_ADPR), full(ST_LW), p.alpha, vs)))
        return rc;
    if ((rc = vsc_stab_consist_wt(full(ST_ADIN), full(ST_OCUR), full(ST_WT), p.beta, p.gamma, vs))) return rc;
    // pyramid (videostabilizer.cpp:203-217)
    const int L = p.pyramid_levels;
    auto lvl = [&](int j, int k) { return img(ST_LEVEL0 + 4 * j + k, s->lw[j], s->lh[j]); };
    for (int j = 0; j < L; ++j) {
        if (j == 0) {
            if ((rc = launch_copy_view(hwc(full(ST_PCUR)), hwc(lvl(0, 0)), st))) return rc;
            if ((rc = launch_copy_view(hwc(full(ST_ADPR)), hwc(lvl(0, 1)), st))) return rc;
            if ((rc = launch_copy_view(hwc(full(ST_WT)), hwc(lvl(0, 2)), st))) return rc;
            if ((rc = launch_copy_view(hwc(full(ST_PCUR)), hwc(lvl(0, 3)), st))) return rc;
        } else {
            for (int k = 0; k < 4; ++k)
                if ((rc = vsc_stab_bilinear(lvl(j - 1, k), lvl(j, k), vs))) return rc;
        }
    }
    // videostabilizer.cpp:219-227
    for (int j = L - 1; j >= 0; --j) {
        if (j != L - 1 && (rc = vsc_stab_bilinear(lvl(j + 1, 3), lvl(j, 3), vs))) return rc;
        if ((rc = vsc_stab_consist_out(lvl(j, 0), lvl(j, 1), lvl(j, 2), p.num_iter / (j + 1), p.step_size, p.mom_fac,
                                       lvl(j, 3), s->st[ST_SOLVE_WS], vs)))
            return rc;
    }
    const vsc_image res = lvl(0, 3);
    if (out_rgba && (rc = vsc_float_to_rgba(res, out_rgba, vs))) return rc;
    if (out_float && (rc = launch_copy_view(hwc(res), View{out_float, W, H, 3, 1, 3}, st))) return rc;
    // last <- result (videostabilizer.cpp:247)
    const int nl = s->last_idx ^ 1;
    if ((rc = launch_copy_view(hwc(res), pl(s->O[0][nl], W, H), st))) return rc;
    s->last_idx = nl;
    return VSC_OK;
}

}  // namespace

namespace vsc {

// The fused frame of frame_fused as a batch item (the common forms: one level,
// or two levels whose level 1 is the exact decimation of level 0).  The front
// pass writes every level's constants; a single level's initial solution is P
// written straight into the O buffer that does not hold `last` (the front pass
// reads `last` while it runs).
int stream_batch_frame(vsc_stream* s, const vsc_frame_in* in, const vsc_params* params, uint8_t* out_rgba, int fma,
                       hipStream_t prep, BatchFrame* bf)
{
    if (!s || !bf) return set_error(VSC_E_INVALID, "vsc_stab_frame: null stream");
    int rc = check_frame_in(s, in);
    if (rc) return rc;
    vsc_params p;
    if (params) p = *params;
    else vsc_default_params(&p);
    if ((rc = check_params(s, p))) return rc;
    const int L = p.pyramid_levels;
    const int W = s->W, H = s->H;
    if (L > 2 || (L == 2 && !batch_exact_2x(W, H)) || p.num_iter > BATCH_MAX_LAUNCH * BATCH_HALO)
        return set_error(VSC_E_UNSUPPORTED, "batched frame: needs 1 level, or 2 with even frame sizes, and at most " +
                                                std::to_string(BATCH_MAX_LAUNCH * BATCH_HALO) + " sweeps");
    View fwd, bwd;
    if ((rc = prepare_flows(s, in, &fwd, &bwd, prep))) return rc;
    *bf = BatchFrame{};
    FrontArgs& a = bf->front;
    a.oPrev = Rgba{in->orig_prev, W, H};
    a.oCur = Rgba{in->orig_cur, W, H};
    a.oNext = Rgba{in->orig_next, W, H};
    a.pPrev = Rgba{in->proc_prev, W, H};
    a.pCur = Rgba{in->proc_cur, W, H};
    a.pNext = Rgba{in->proc_next, W, H};
    a.last = pl(s->O[0][s->last_idx], W, H);
    a.fwd = fwd;
    a.bwd = bwd;
    a.L0 = pl(s->L[0], W, H);
    a.A0 = pl(s->A[0], W, H);
    a.W0 = pl(s->Wt[0], W, H);
    a.decimate = L == 2 ? 1 : 0;
    a.write_p = L == 1 ? 1 : 0;
    const int start0 = L == 1 ? (s->last_idx ^ 1) : 0;
    a.P0 = pl(s->O[0][start0], W, H);   // L == 1: P is the initial solution
    if (L == 2) {
        const int w1 = s->lw[1], h1 = s->lh[1];
        a.L1 = pl(s->L[1], w1, h1);
        a.A1 = pl(s->A[1], w1, h1);
        a.W1 = pl(s->Wt[1], w1, h1);
        a.O1 = pl(s->O[1][0], w1, h1);
        a.P1 = a.O1;
    } else {
        a.L1 = a.A1 = a.W1 = a.O1 = a.P1 = a.A0;
    }
    a.alpha = p.alpha;
    a.beta = p.beta;
    a.gamma = p.gamma;
    a.fma = fma ? 1 : 0;
    a.step = p.step_size;
    for (const View* v : {&a.fwd, &a.bwd, &a.last, &a.L0, &a.A0, &a.W0, &a.P0, &a.L1, &a.A1, &a.W1, &a.O1, &a.P1})
        if (!fits32(*v)) return set_error(VSC_E_INVALID, "batched frame: image too large for 32-bit offsets");
    bf->levels = L;
    int fin[2] = {0, 0};
    for (int j = 0; j < L; ++j) {
        BatchLevel& l = bf->lev[j];
        l.L = s->L[j];
        l.A = s->A[j];
        l.Wt = s->Wt[j];
        for (int k = 0; k < 2; ++k) {
            l.O[k] = s->O[j][k];
            l.U[k] = s->U[j][k];
        }
        l.w = s->lw[j];
        l.h = s->lh[j];
        l.iters = p.num_iter / (j + 1);
        l.start = j == 0 ? start0 : 0;
        batch_tiles(l.w, l.h, fma, &l.tiles_x, &l.tiles_y);
        const int launches = (l.iters + BATCH_HALO - 1) / BATCH_HALO;
        fin[j] = l.start ^ (launches & 1);
    }
    bf->up_idx = fin[1];
    bf->out_idx = fin[0];
    bf->rgba_vec = (W % 4) == 0 && reinterpret_cast<uintptr_t>(out_rgba) % 16 == 0 ? 1 : 0;
    bf->step = p.step_size;
    bf->mom = p.mom_fac;
    bf->out_rgba = out_rgba;
    s->last_idx = fin[0];
    return VSC_OK;
}

}  // namespace vsc

extern "C" int vsc_stab_frame(vsc_stream* s, const vsc_frame_in* in, const vsc_params* params, uint8_t* out_rgba,
                              float* out_float, int mode, vsc_hip_stream stream)
{
    if (!s) return set_error(VSC_E_INVALID, "vsc_stab_frame: null stream");
    int rc = check_frame_in(s, in);
    if (rc) return rc;
    vsc_params p;
    if (params) p = *params;
    else vsc_default_params(&p);
    if ((rc = check_params(s, p))) return rc;
    hipStream_t st = as_stream(stream);
    if (mode == VSC_FRAME_FUSED) return frame_fused(s, in, p, out_rgba, out_float, 0, st);
    if (mode == VSC_FRAME_FMA) return frame_fused(s, in, p, out_rgba, out_float, 1, st);
    if (mode == VSC_FRAME_STAGED) return frame_staged(s, in, p, out_rgba, out_float, st);
    return set_error(VSC_E_INVALID, "vsc_stab_frame: unknown mode");
}
