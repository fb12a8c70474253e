// Per-stream frame pipeline: VideoStabilizer::doOneStep with its flow
// acquisition (retrieveOpticalFlow, videostabilizer.cpp:267-279, through
// FlowModel::run, flowmodel.cpp:121-168) on one device stream:
//   1. flow network features of the new frame (cached per frame: each frame
//      is extracted once and serves three windows),
//   2. forward (cur -> next) and backward flows in one batched decode, at the
//      FLOWDOWNSCALE resolution,
//   3. upsampling to the frame: the reference's get_bilinear (vectors
//      unscaled, as flowmodel.cpp:158-161, or scaled by the factor), or the
//      learned upsampler gated per frame by the forward-backward check,
//   4. the fused stabilization of vsc_stab_frame.
// All buffers are allocated at creation; a frame allocates nothing.  The
// flows of frame t (and its fallback gate) live in buffer set t mod 2B (B =
// cfg.flow_batch), so that with overlap the flows of the next frame or batch
// (internal stream sf: decode, consistency gate) are computed while the
// current one is upsampled and stabilized (internal stream ss), and the new
// frames' features of a batch (internal stream se) while the previous batch
// decodes.  A batch of n <= B frames (vsc_pipeline_frames) decodes the
// forward and backward flows of all its frames in one network batch of 2n
// pairs.
#include "vsc_internal.h"

#include <algorithm>
#include <cmath>
#include <new>

constexpr int PIPE_MAX_BATCH = 4, PIPE_MAX_SETS = 2 * PIPE_MAX_BATCH, PIPE_MAX_SLOTS = 2 * PIPE_MAX_BATCH + 2;

struct vsc_pipeline {
    int W, H, Wd, Hd;
    vsc_pipeline_cfg cfg;
    int sets = 2, slots = 3;                       // flow buffer sets (2B), feature slots (B + 2; overlap 2B + 2)
    vsc_flownet* net = nullptr;
    vsc_upsampler* ups = nullptr;
    vsc_stream* stab = nullptr;
    float* lo[PIPE_MAX_SETS][2] = {};              // [frame mod sets][fwd, bwd] flows at the flow resolution (HWC 2)
    float* hi[PIPE_MAX_SETS][2] = {};              // [frame mod sets][fwd, bwd] at frame size when upsampling here
    uint32_t* count = nullptr;                     // [set] inconsistent-pixel counts (64 B apart)
    float* gate = nullptr;                         // [set] fallback gates (64 B apart)
    long long slot_frame[PIPE_MAX_SLOTS];           // frame index held by each feature slot (-1: none)
    long long prev_fwd_frame = -2;                 // last frame whose flows were computed
    // overlap: flows on sf, stabilization on ss; ev_s[k] = the stabilization
    // that last read buffer set k, ev_f = the last flow work
    hipStream_t sf = nullptr, ss = nullptr;
    hipEvent_t ev_in = nullptr, ev_f = nullptr, ev_s[PIPE_MAX_SETS] = {};
    bool f_rec = false, s_rec[PIPE_MAX_SETS] = {};
    int last_set = 0;                              // buffer set of the last frame
    // overlap: features on se.  ev_e = the last extraction; ev_d[c & 1] = the
    // decode of overlapped call c; slot_read[j] = the call whose decode last
    // read feature slot j (-1: none).  With 2B + 2 slots the slots a batch's
    // extraction overwrites were last read two calls before, so it runs
    // beside the previous call's decode.
    hipStream_t se = nullptr;
    hipEvent_t ev_e = nullptr, ev_d[2] = {};
    bool e_rec = false, d_rec[2] = {};
    long long calls = 0, slot_read[PIPE_MAX_SLOTS];
};

namespace vsc {
namespace {

__global__ void k_flow_upscale(const float* __restrict__ lo, int lw, int lh, float* hi, int W, int H, float scale)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const BiTap t = bilinear_tap(x, y, lw, lh, W, H);
    View fl{const_cast<float*>(lo), lw, lh, 2, 1, 2};
    const long long p = (long long)y * W + x;
    hi[p * 2] = bilinear_sample(fl, t, 0) * scale;
    hi[p * 2 + 1] = bilinear_sample(fl, t, 1) * scale;
}

__global__ void k_set_gate(float* g, float v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) g[0] = v;
}

int dalloc(void** p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) return set_error(VSC_E_NOMEM, "vsc_pipeline_create: device allocation failed");
    return VSC_OK;
}

bool learned(const vsc_pipeline_cfg& c) { return c.downscale > 1 && c.upsample == 1; }
bool upsample_here(const vsc_pipeline_cfg& c) { return c.downscale > 1 && (c.upsample == 1 || c.scale_vectors); }

int pmod(long long v, int m) { return (int)(((v % m) + m) % m); }

bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// overlap mode: make `s` wait for all outstanding work of the pipeline
int join_all(vsc_pipeline* p, hipStream_t s)
{
    if (!p->cfg.overlap) return VSC_OK;
    if (p->f_rec && check_hip(hipStreamWaitEvent(s, p->ev_f, 0), "vsc_pipeline join")) return VSC_E_HIP;
    if (p->e_rec && check_hip(hipStreamWaitEvent(s, p->ev_e, 0), "vsc_pipeline join")) return VSC_E_HIP;
    for (int d = 0; d < 2; ++d)
        if (p->d_rec[d] && check_hip(hipStreamWaitEvent(s, p->ev_d[d], 0), "vsc_pipeline join")) return VSC_E_HIP;
    for (int k = 0; k < p->sets; ++k)
        if (p->s_rec[k] && check_hip(hipStreamWaitEvent(s, p->ev_s[k], 0), "vsc_pipeline join")) return VSC_E_HIP;
    return VSC_OK;
}

}  // namespace
}  // namespace vsc

using namespace vsc;

extern "C" void vsc_pipeline_default_cfg(vsc_pipeline_cfg* c)
{
    if (!c) return;
    c->flow_dtype = VSC_F32;
    c->downscale = 1;
    c->upsample = 0;
    c->scale_vectors = 0;
    c->fb_max_inconsistent = 0.25f;
    c->ref_bwd_index = 0;
    c->seed = 0;
    c->solver = 0;
    c->overlap = 0;
    c->flow_batch = 1;
    c->corr_dtype = VSC_F32;
}

static int cfg_ok(int W, int H, const vsc_pipeline_cfg* c)
{
    if (!c || W <= 0 || H <= 0) return set_error(VSC_E_INVALID, "vsc_pipeline: bad size or NULL config");
    if (c->flow_dtype != VSC_F32 && c->flow_dtype != VSC_BF16)
        return set_error(VSC_E_DTYPE, "vsc_pipeline: flow_dtype must be VSC_F32 or VSC_BF16");
    if (c->downscale != 1 && c->downscale != 2 && c->downscale != 4)
        return set_error(VSC_E_INVALID, "vsc_pipeline: downscale must be 1, 2 or 4");
    if (c->upsample != 0 && c->upsample != 1) return set_error(VSC_E_INVALID, "vsc_pipeline: upsample must be 0 or 1");
    if (c->solver != 0 && c->solver != 1) return set_error(VSC_E_INVALID, "vsc_pipeline: solver must be 0 or 1");
    if (c->overlap != 0 && c->overlap != 1) return set_error(VSC_E_INVALID, "vsc_pipeline: overlap must be 0 or 1");
    if (c->flow_batch < 1 || c->flow_batch > PIPE_MAX_BATCH)
        return set_error(VSC_E_INVALID, "vsc_pipeline: flow_batch must be in [1, 4]");
    if (c->corr_dtype != VSC_F32 && c->corr_dtype != VSC_BF16)
        return set_error(VSC_E_DTYPE, "vsc_pipeline: corr_dtype must be VSC_F32 or VSC_BF16");
    return VSC_OK;
}

extern "C" size_t vsc_pipeline_footprint(int width, int height, const vsc_pipeline_cfg* cfg)
{
    if (cfg_ok(width, height, cfg)) return 0;
    const int Wd = (int)std::lround((double)width / cfg->downscale), Hd = (int)std::lround((double)height / cfg->downscale);
    const int B = cfg->flow_batch;
    size_t t = flownet_footprint_pairs(width, height, cfg->downscale, cfg->flow_dtype, cfg->overlap ? 2 * B + 2 : B + 2,
                                       2 * B);
    const bool here = upsample_here(*cfg);
    t += vsc_stream_footprint(width, height, here ? width : Wd, here ? height : Hd);
    t += 2 * 2 * B * (size_t)Wd * Hd * 2 * 4;
    if (here) t += 2 * 2 * B * (size_t)width * height * 2 * 4;
    if (learned(*cfg)) t += 4 * 3202 * 4 + 512;
    return t + 64;
}

extern "C" int vsc_pipeline_create(int width, int height, const vsc_pipeline_cfg* cfg, vsc_pipeline** out)
{
    if (!out) return set_error(VSC_E_INVALID, "vsc_pipeline_create: out is NULL");
    *out = nullptr;
    int rc = cfg_ok(width, height, cfg);
    if (rc) return rc;
    vsc_pipeline* p = new (std::nothrow) vsc_pipeline();
    if (!p) return set_error(VSC_E_NOMEM, "vsc_pipeline_create: host memory");
    p->W = width; p->H = height; p->cfg = *cfg;
    p->sets = 2 * cfg->flow_batch;
    p->slots = cfg->overlap ? 2 * cfg->flow_batch + 2 : cfg->flow_batch + 2;
    for (int i = 0; i < PIPE_MAX_SLOTS; ++i) p->slot_frame[i] = p->slot_read[i] = -1;
    if ((rc = flownet_create_pairs(width, height, cfg->downscale, cfg->flow_dtype, cfg->seed, p->slots, 2 * cfg->flow_batch,
                                   &p->net)) ||
        (rc = vsc_flownet_set_corr_dtype(p->net, cfg->flow_dtype == VSC_BF16 ? VSC_BF16 : cfg->corr_dtype)) ||
        (rc = vsc_flownet_flow_size(p->net, &p->Wd, &p->Hd))) {
        vsc_pipeline_destroy(p);
        return rc;
    }
    const bool here = upsample_here(*cfg);
    if ((rc = vsc_stream_create(width, height, here ? width : p->Wd, here ? height : p->Hd, &p->stab))) {
        vsc_pipeline_destroy(p);
        return rc;
    }
    const size_t lo_b = (size_t)p->Wd * p->Hd * 2 * 4, hi_b = (size_t)width * height * 2 * 4;
    for (int i = 0; i < 2 * p->sets && !rc; ++i) rc = dalloc(reinterpret_cast<void**>(&p->lo[i / 2][i % 2]), lo_b);
    if (here)
        for (int i = 0; i < 2 * p->sets && !rc; ++i) rc = dalloc(reinterpret_cast<void**>(&p->hi[i / 2][i % 2]), hi_b);
    if (!rc && cfg->overlap) {
        // default-priority streams; events without timing
        if (check_hip(hipStreamCreateWithFlags(&p->sf, hipStreamNonBlocking), "vsc_pipeline_create") ||
            check_hip(hipStreamCreateWithFlags(&p->ss, hipStreamNonBlocking), "vsc_pipeline_create") ||
            check_hip(hipStreamCreateWithFlags(&p->se, hipStreamNonBlocking), "vsc_pipeline_create") ||
            check_hip(hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming), "vsc_pipeline_create") ||
            check_hip(hipEventCreateWithFlags(&p->ev_f, hipEventDisableTiming), "vsc_pipeline_create") ||
            check_hip(hipEventCreateWithFlags(&p->ev_e, hipEventDisableTiming), "vsc_pipeline_create") ||
            check_hip(hipEventCreateWithFlags(&p->ev_d[0], hipEventDisableTiming), "vsc_pipeline_create") ||
            check_hip(hipEventCreateWithFlags(&p->ev_d[1], hipEventDisableTiming), "vsc_pipeline_create"))
            rc = VSC_E_HIP;
        for (int k = 0; k < p->sets && !rc; ++k)
            if (check_hip(hipEventCreateWithFlags(&p->ev_s[k], hipEventDisableTiming), "vsc_pipeline_create"))
                rc = VSC_E_HIP;
    }
    const size_t gb = 64 * PIPE_MAX_SETS;
    if (!rc) rc = dalloc(reinterpret_cast<void**>(&p->count), gb);
    if (!rc) rc = dalloc(reinterpret_cast<void**>(&p->gate), gb);
    // forms without the learned upsampler report gate 0 (bilinear) and count 0
    if (!rc && (check_hip(hipMemset(p->count, 0, gb), "vsc_pipeline_create") ||
                check_hip(hipMemset(p->gate, 0, gb), "vsc_pipeline_create")))
        rc = VSC_E_HIP;
    if (!rc && learned(*cfg)) {
        rc = vsc_upsampler_create(cfg->seed, &p->ups);
    }
    if (rc) {
        vsc_pipeline_destroy(p);
        return rc;
    }
    *out = p;
    return VSC_OK;
}

extern "C" int vsc_pipeline_destroy(vsc_pipeline* p)
{
    if (!p) return VSC_OK;
    for (hipStream_t h : {p->sf, p->ss, p->se})
        if (h) (void)hipStreamSynchronize(h);
    if (p->net) vsc_flownet_destroy(p->net);
    if (p->ups) vsc_upsampler_destroy(p->ups);
    if (p->stab) vsc_stream_destroy(p->stab);
    for (int k = 0; k < PIPE_MAX_SETS; ++k)
        for (int d = 0; d < 2; ++d) {
            if (p->lo[k][d]) (void)hipFree(p->lo[k][d]);
            if (p->hi[k][d]) (void)hipFree(p->hi[k][d]);
        }
    for (void* b : {static_cast<void*>(p->count), static_cast<void*>(p->gate)})
        if (b) (void)hipFree(b);
    for (hipEvent_t e : {p->ev_in, p->ev_f, p->ev_e, p->ev_d[0], p->ev_d[1]})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : p->ev_s)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t h : {p->sf, p->ss, p->se})
        if (h) (void)hipStreamDestroy(h);
    delete p;
    return VSC_OK;
}

extern "C" int vsc_pipeline_set_upsampler_params(vsc_pipeline* p, const float* host)
{
    if (!p || !host) return set_error(VSC_E_INVALID, "vsc_pipeline_set_upsampler_params: NULL argument");
    if (!p->ups) return set_error(VSC_E_INVALID, "vsc_pipeline_set_upsampler_params: pipeline has no learned upsampler");
    return vsc_upsampler_set_params(p->ups, host);
}

extern "C" int vsc_pipeline_flownet(vsc_pipeline* p, vsc_flownet** net)
{
    if (!p || !net) return set_error(VSC_E_INVALID, "vsc_pipeline_flownet: NULL argument");
    *net = p->net;
    return VSC_OK;
}

extern "C" int vsc_pipeline_stream(vsc_pipeline* p, vsc_stream** s)
{
    if (!p || !s) return set_error(VSC_E_INVALID, "vsc_pipeline_stream: NULL argument");
    *s = p->stab;
    return VSC_OK;
}

// the forward and backward pairs of window `frame` (feature slots of frames
// frame-1 .. frame+1, extracted before)
static void frame_pairs(const vsc_pipeline* p, long long frame, int* first, int* second)
{
    const int sp = pmod(frame - 1, p->slots), sc = pmod(frame, p->slots), sn = pmod(frame + 1, p->slots);
    // forward cur -> next; backward cur -> prev (or next -> cur as videostabilizer.cpp:271-272)
    first[0] = sc;
    second[0] = sn;
    first[1] = p->cfg.ref_bwd_index ? sn : sc;
    second[1] = p->cfg.ref_bwd_index ? sc : sp;
}

// the network part for the n consecutive windows of frames first .. first+n-1:
// features of frames first-1 .. first+n (each extracted once), then one
// batched decode of their 2n flow pairs into outs (fwd, bwd per frame) on s.
// es: the extraction stream (overlap: p->se, which the caller has ordered
// after its inputs; else s).
static int flows_batch_on(vsc_pipeline* p, const vsc_window* w, int n, long long first, float* const* outs,
                          hipStream_t s, hipStream_t es)
{
    int rc;
    const bool split = es != s;
    const long long call = p->calls;
    bool extracted = false, waited[2] = {false, false};
    // frames first-1 .. first+n: the ones not cached are extracted in batches
    // of consecutive frames whose slots do not wrap
    const uint8_t* rgba[PIPE_MAX_BATCH + 2];
    rgba[0] = w[0].orig_prev;
    for (int i = 0; i < n; ++i) rgba[i + 1] = w[i].orig_cur;
    rgba[n + 1] = w[n - 1].orig_next;
    for (int j = 0; j < n + 2;) {
        const long long f = first - 1 + j;
        if (p->slot_frame[pmod(f, p->slots)] == f) {
            ++j;
            continue;
        }
        int m = 1;
        while (j + m < n + 2 && pmod(f + m, p->slots) != 0 && p->slot_frame[pmod(f + m, p->slots)] != f + m) ++m;
        m = std::min(m, p->cfg.flow_batch);
        if (split) {
            // the slots overwritten must be done with: wait for the decode
            // that last read each (the latest record of its event is that
            // decode or a later one on the same stream)
            for (int i = 0; i < m; ++i) {
                const long long r = p->slot_read[pmod(f + i, p->slots)];
                if (r < 0 || waited[r & 1] || !p->d_rec[r & 1]) continue;
                if (check_hip(hipStreamWaitEvent(es, p->ev_d[r & 1], 0), "vsc_pipeline_frame")) return VSC_E_HIP;
                waited[r & 1] = true;
            }
        }
        if ((rc = flownet_extract_frames(p->net, rgba + j, m, pmod(f, p->slots), es))) return rc;
        for (int i = 0; i < m; ++i) p->slot_frame[pmod(f + i, p->slots)] = f + i;
        extracted = true;
        j += m;
    }
    if (split && extracted) {
        if (check_hip(hipEventRecord(p->ev_e, es), "vsc_pipeline_frame")) return VSC_E_HIP;
        p->e_rec = true;
        if (check_hip(hipStreamWaitEvent(s, p->ev_e, 0), "vsc_pipeline_frame")) return VSC_E_HIP;
    }
    int a[2 * PIPE_MAX_BATCH], b[2 * PIPE_MAX_BATCH];
    for (int i = 0; i < n; ++i) frame_pairs(p, first + i, a + 2 * i, b + 2 * i);
    if ((rc = vsc_flownet_flow(p->net, 2 * n, a, b, outs, 2, reinterpret_cast<vsc_hip_stream>(s)))) return rc;
    if (split) {
        if (check_hip(hipEventRecord(p->ev_d[call & 1], s), "vsc_pipeline_frame")) return VSC_E_HIP;
        p->d_rec[call & 1] = true;
        for (int i = 0; i < 2 * n; ++i) p->slot_read[a[i]] = p->slot_read[b[i]] = call;
        p->calls = call + 1;
    }
    return VSC_OK;
}

static int flows_on(vsc_pipeline* p, const vsc_window* w, long long frame, float* fwd_out, float* bwd_out,
                    hipStream_t s)
{
    float* outs[2] = {fwd_out, bwd_out};
    return flows_batch_on(p, w, 1, frame, outs, s, s);
}

extern "C" int vsc_pipeline_flows(vsc_pipeline* p, const vsc_window* w, long long frame, float* fwd_out,
                                  float* bwd_out, vsc_hip_stream stream)
{
    if (!p || !w || !w->orig_prev || !w->orig_cur || !w->orig_next)
        return set_error(VSC_E_INVALID, "vsc_pipeline_flows: NULL argument");
    hipStream_t s = as_stream(stream);
    int rc = join_all(p, s);   // the flow network's buffers may still be in use by frame work
    if (rc) return rc;
    const int k = pmod(frame, p->sets);
    return flows_on(p, w, frame, fwd_out ? fwd_out : p->lo[k][0], bwd_out ? bwd_out : p->lo[k][1], s);
}

namespace {

// Host state of a pipeline that a frame's enqueueing changes; restored when
// enqueueing (or capturing) fails part-way, so a later call recomputes what
// the failed one did not finish.  A feature slot the failed call began to
// fill is marked empty rather than given back its old frame: an eager copy
// into it may already have been enqueued (a captured one is discarded, and
// the slot is then merely extracted again).
struct HostState {
    long long slot_frame[PIPE_MAX_SLOTS];
    long long prev_fwd;
    int set;
    explicit HostState(const vsc_pipeline* p) : prev_fwd(p->prev_fwd_frame), set(p->last_set)
    {
        for (int i = 0; i < PIPE_MAX_SLOTS; ++i) slot_frame[i] = p->slot_frame[i];
    }
    void restore(vsc_pipeline* p) const
    {
        for (int i = 0; i < PIPE_MAX_SLOTS; ++i) p->slot_frame[i] = p->slot_frame[i] == slot_frame[i] ? slot_frame[i] : -1;
        p->prev_fwd_frame = prev_fwd;
        p->last_set = set;
    }
};

// The fallback gate of window `frame` on fs (its flows are in set frame mod
// sets; the previous frame's forward flow, when computed, in its own set).
int gate_on(vsc_pipeline* p, long long frame, hipStream_t fs)
{
    const int k = pmod(frame, p->sets);
    int rc;
    const vsc_hip_stream fsv = reinterpret_cast<vsc_hip_stream>(fs);
    // previous forward flow for the consistency check of this frame's backward flow
    const bool have_prev = p->prev_fwd_frame == frame - 1;
    float* gate = p->gate + 16 * k;
    uint32_t* count = p->count + 16 * k;
    if (learned(p->cfg)) {
        if (p->cfg.fb_max_inconsistent < 0.0f || (!have_prev && !p->cfg.ref_bwd_index)) {
            hipLaunchKernelGGL(k_set_gate, dim3(1), dim3(64), 0, fs, gate, 1.0f);
            if ((rc = check_launch("pipeline gate"))) return rc;
        } else if (p->cfg.ref_bwd_index) {
            // fwd: cur -> next, bwd: next -> cur are each other's inverse
            if ((rc = vsc_flow_fb_gate(p->lo[k][0], p->lo[k][1], p->Wd, p->Hd, 2, p->cfg.fb_max_inconsistent, count,
                                       gate, fsv)))
                return rc;
        } else {
            // bwd: cur -> prev against the previous frame's fwd: prev -> cur
            if ((rc = vsc_flow_fb_gate(p->lo[k][1], p->lo[pmod(frame - 1, p->sets)][0], p->Wd, p->Hd, 2,
                                       p->cfg.fb_max_inconsistent, count, gate, fsv)))
                return rc;
        }
    }
    p->prev_fwd_frame = frame;
    p->last_set = k;
    return VSC_OK;
}

// Upsampling of window `frame`'s flows to the frame, on ss (it reads only this
// frame's buffer set).  *ff, *fb receive the frame-size (or, without
// upsampling here, flow-size) flows.
int upsample_on(vsc_pipeline* p, const vsc_window* w, long long frame, hipStream_t ss, const float** ff,
                const float** fb)
{
    const int f = p->cfg.downscale;
    const int k = pmod(frame, p->sets);
    int rc;
    *ff = p->lo[k][0];
    *fb = p->lo[k][1];
    if (learned(p->cfg)) {
        // the learned upsampler is trained on full-resolution vectors: always scaled
        const float sc = (float)f;
        const float* const lo[2] = {p->lo[k][0], p->lo[k][1]};
        float* const hi[2] = {p->hi[k][0], p->hi[k][1]};
        if ((rc = upsampler_run_flows(p->ups, 2, lo, p->Wd, p->Hd, 2, w->orig_cur, p->W, p->H, sc, hi, 2,
                                      p->gate + 16 * k, ss)))
            return rc;
        *ff = p->hi[k][0];
        *fb = p->hi[k][1];
    } else if (upsample_here(p->cfg)) {
        dim3 grid((p->W + 255) / 256, p->H);
        hipLaunchKernelGGL(k_flow_upscale, grid, dim3(256), 0, ss, p->lo[k][0], p->Wd, p->Hd, p->hi[k][0], p->W, p->H,
                           (float)f);
        hipLaunchKernelGGL(k_flow_upscale, grid, dim3(256), 0, ss, p->lo[k][1], p->Wd, p->Hd, p->hi[k][1], p->W, p->H,
                           (float)f);
        if ((rc = check_launch("pipeline flow upscale"))) return rc;
        *ff = p->hi[k][0];
        *fb = p->hi[k][1];
    }
    return VSC_OK;
}

// flows computed on fs are complete before ss continues (overlap)
int flows_done(vsc_pipeline* p, hipStream_t fs, hipStream_t ss)
{
    if (check_hip(hipEventRecord(p->ev_f, fs), "vsc_pipeline_frame") ||
        check_hip(hipStreamWaitEvent(ss, p->ev_f, 0), "vsc_pipeline_frame"))
        return VSC_E_HIP;
    p->f_rec = true;
    return VSC_OK;
}

// The flows of window `frame` at frame size: features, the batched
// forward/backward decode and the consistency gate on fs, the upsampling on
// ss (ov: the flows' completion is ordered before ss by ev_f).
int frame_flows(vsc_pipeline* p, const vsc_window* w, long long frame, hipStream_t fs, hipStream_t ss, bool ov,
                const float** ff, const float** fb)
{
    const int k = pmod(frame, p->sets);
    int rc;
    if ((rc = flows_on(p, w, frame, p->lo[k][0], p->lo[k][1], fs)) || (rc = gate_on(p, frame, fs))) return rc;
    if (ov && (rc = flows_done(p, fs, ss))) return rc;
    return upsample_on(p, w, frame, ss, ff, fb);
}

// the flows frame_flows leaves for window `frame` (no launch)
void frame_flow_ptrs(const vsc_pipeline* p, long long frame, const float** ff, const float** fb)
{
    const int k = pmod(frame, p->sets);
    const bool hi = learned(p->cfg) || upsample_here(p->cfg);
    *ff = hi ? p->hi[k][0] : p->lo[k][0];
    *fb = hi ? p->hi[k][1] : p->lo[k][1];
}

}  // namespace

extern "C" int vsc_pipeline_frame(vsc_pipeline* p, const vsc_window* w, long long frame, const vsc_params* params,
                                  uint8_t* out_rgba, float* out_float, vsc_hip_stream stream)
{
    return vsc_pipeline_frames(p, w, 1, frame, params, &out_rgba, &out_float, stream);
}

extern "C" int vsc_pipeline_frames(vsc_pipeline* p, const vsc_window* w, int n, long long first,
                                   const vsc_params* params, uint8_t* const* out_rgba, float* const* out_float,
                                   vsc_hip_stream stream)
{
    if (!p || !w) return set_error(VSC_E_INVALID, "vsc_pipeline_frame: NULL argument");
    if (n < 1 || n > p->cfg.flow_batch)
        return set_error(VSC_E_INVALID, "vsc_pipeline_frames: n must be in [1, flow_batch]");
    for (int i = 0; i < n; ++i)
        if (!w[i].orig_prev || !w[i].orig_cur || !w[i].orig_next)
            return set_error(VSC_E_INVALID, "vsc_pipeline_frame: NULL argument");
    // every check before any state change: a refused frame leaves the
    // pipeline as it was
    int rc = stream_frame_check(p->stab, params);
    if (rc) return rc;
    hipStream_t cs = as_stream(stream);
    const bool ov = p->cfg.overlap && !capturing(cs);
    hipStream_t fs = ov ? p->sf : cs, ss = ov ? p->ss : cs;
    const HostState saved(p);
    auto fail = [&](int code) {
        saved.restore(p);
        return code;
    };
    if (ov) {
        // flows wait for the caller's inputs and for the stabilizations that
        // last read the batch's buffer sets (frames first - 2B ..)
        if (check_hip(hipEventRecord(p->ev_in, cs), "vsc_pipeline_frame") ||
            check_hip(hipStreamWaitEvent(fs, p->ev_in, 0), "vsc_pipeline_frame") ||
            check_hip(hipStreamWaitEvent(p->se, p->ev_in, 0), "vsc_pipeline_frame"))
            return fail(VSC_E_HIP);
        for (int i = 0; i < n; ++i) {
            const int k = pmod(first + i, p->sets);
            if (p->s_rec[k] && check_hip(hipStreamWaitEvent(fs, p->ev_s[k], 0), "vsc_pipeline_frame"))
                return fail(VSC_E_HIP);
        }
    }
    float* outs[2 * PIPE_MAX_BATCH];
    for (int i = 0; i < n; ++i) {
        const int k = pmod(first + i, p->sets);
        outs[2 * i] = p->lo[k][0];
        outs[2 * i + 1] = p->lo[k][1];
    }
    if ((rc = flows_batch_on(p, w, n, first, outs, fs, ov ? p->se : fs))) return fail(rc);
    for (int i = 0; i < n; ++i)
        if ((rc = gate_on(p, first + i, fs))) return fail(rc);
    // the BF16 network's flows are upsampled on the flow stream, overlapping
    // the previous frames' stabilization (1080p: batches of two 774 vs 758
    // frames/s; one frame per call 809 vs 805 with the features extracted on
    // their own stream); the longer flow work of an FP32 network would delay
    // the next flows instead (474 vs 499 frames/s), so its flows are
    // upsampled on the stabilization stream
    const bool ups_on_flows = p->cfg.flow_dtype == VSC_BF16;
    const float* ffs[PIPE_MAX_BATCH] = {};
    const float* fbs[PIPE_MAX_BATCH] = {};
    if (ups_on_flows)
        for (int i = 0; i < n; ++i)
            if ((rc = upsample_on(p, &w[i], first + i, fs, &ffs[i], &fbs[i]))) return fail(rc);
    if (ov && (rc = flows_done(p, fs, ss))) return fail(rc);
    for (int i = 0; i < n; ++i) {
        const long long frame = first + i;
        const float *ff = ffs[i], *fb = fbs[i];
        if (!ups_on_flows && (rc = upsample_on(p, &w[i], frame, ss, &ff, &fb))) return fail(rc);
        vsc_frame_in in{w[i].orig_prev, w[i].orig_cur, w[i].orig_next, w[i].proc_prev, w[i].proc_cur,
                        w[i].proc_next, ff, fb, 2};
        rc = vsc_stab_frame(p->stab, &in, params, out_rgba ? out_rgba[i] : nullptr, out_float ? out_float[i] : nullptr,
                            p->cfg.solver ? VSC_FRAME_FMA : VSC_FRAME_FUSED, reinterpret_cast<vsc_hip_stream>(ss));
        if (rc) return fail(rc);
        if (!ov) continue;
        const int k = pmod(frame, p->sets);
        if (check_hip(hipEventRecord(p->ev_s[k], ss), "vsc_pipeline_frame")) return fail(VSC_E_HIP);
        p->s_rec[k] = true;
    }
    if (!ov) return VSC_OK;
    // the caller's stream is ordered after the frame before this call's
    const int kp = pmod(first - 1, p->sets);
    if (p->s_rec[kp] && check_hip(hipStreamWaitEvent(cs, p->ev_s[kp], 0), "vsc_pipeline_frame"))
        return fail(VSC_E_HIP);
    return VSC_OK;
}

namespace vsc {

int pipeline_batch_check(vsc_pipeline* p, const vsc_params* params)
{
    int rc = stream_frame_check(p->stab, params);
    if (rc) return rc;
    if (p->cfg.overlap) return set_error(VSC_E_UNSUPPORTED, "batched frame: the pipeline runs overlapped");
    // the stream must receive frame-size flows (the pipeline upsamples them)
    if (p->cfg.downscale > 1 && !upsample_here(p->cfg))
        return set_error(VSC_E_UNSUPPORTED, "batched frame: the stream upsamples its own flows");
    vsc_params q;
    if (params) q = *params;
    else vsc_default_params(&q);
    if (q.pyramid_levels > 2 || (q.pyramid_levels == 2 && !batch_exact_2x(p->W, p->H)) ||
        q.num_iter > BATCH_MAX_LAUNCH * BATCH_HALO)
        return set_error(VSC_E_UNSUPPORTED, "batched frame: needs 1 level, or 2 with even frame sizes");
    return VSC_OK;
}

int pipeline_flows_part(vsc_pipeline* p, const vsc_window* w, long long frame, hipStream_t s)
{
    if (!p || !w || !w->orig_prev || !w->orig_cur || !w->orig_next)
        return set_error(VSC_E_INVALID, "vsc_pipeline_frame: NULL argument");
    const HostState saved(p);
    const float *ff = nullptr, *fb = nullptr;
    const int rc = frame_flows(p, w, frame, s, s, false, &ff, &fb);
    if (rc) saved.restore(p);
    return rc;
}

int pipeline_batch_item(vsc_pipeline* p, const vsc_window* w, long long frame, const vsc_params* params,
                        uint8_t* out_rgba, BatchFrame* bf)
{
    const float *ff = nullptr, *fb = nullptr;
    frame_flow_ptrs(p, frame, &ff, &fb);
    vsc_frame_in in{w->orig_prev, w->orig_cur, w->orig_next, w->proc_prev, w->proc_cur, w->proc_next, ff, fb, 2};
    return stream_batch_frame(p->stab, &in, params, out_rgba, p->cfg.solver ? 1 : 0, nullptr, bf);
}

}  // namespace vsc

extern "C" int vsc_pipeline_sync(vsc_pipeline* p, vsc_hip_stream stream)
{
    if (!p) return set_error(VSC_E_INVALID, "vsc_pipeline_sync: NULL argument");
    return join_all(p, as_stream(stream));
}

extern "C" int vsc_pipeline_gate(const vsc_pipeline* p, float* gate_dev, uint32_t* count_dev, vsc_hip_stream stream)
{
    if (!p) return set_error(VSC_E_INVALID, "vsc_pipeline_gate: NULL argument");
    hipStream_t s = as_stream(stream);
    if (p->cfg.overlap && p->f_rec && check_hip(hipStreamWaitEvent(s, p->ev_f, 0), "vsc_pipeline_gate"))
        return VSC_E_HIP;
    const int k = p->last_set;
    if (gate_dev &&
        check_hip(hipMemcpyAsync(gate_dev, p->gate + 16 * k, 4, hipMemcpyDeviceToDevice, s), "vsc_pipeline_gate"))
        return VSC_E_HIP;
    if (count_dev &&
        check_hip(hipMemcpyAsync(count_dev, p->count + 16 * k, 4, hipMemcpyDeviceToDevice, s), "vsc_pipeline_gate"))
        return VSC_E_HIP;
    return VSC_OK;
}

namespace vsc {

// The host-side state changes of vsc_pipeline_frame(frame) in steady state
// (previous frame was frame - 1, features of frame - 1 and frame cached),
// without launching anything: used when the frame's work is replayed from a
// captured graph.  Returns VSC_E_INVALID when the pipeline is not in that state.
int pipeline_advance(vsc_pipeline* p, long long frame)
{
    const int sp = pmod(frame - 1, p->slots), sc = pmod(frame, p->slots), sn = pmod(frame + 1, p->slots);
    if (p->prev_fwd_frame != frame - 1 || p->slot_frame[sp] != frame - 1 || p->slot_frame[sc] != frame)
        return set_error(VSC_E_INVALID, "pipeline_advance: not in steady state");
    p->slot_frame[sn] = frame + 1;
    p->prev_fwd_frame = frame;
    p->last_set = pmod(frame, p->sets);
    return VSC_OK;
}

}  // namespace vsc
