// Multi-stream frame server: many concurrent video streams of different
// resolutions on one GPU.  Each admitted stream owns a vsc_pipeline, a 3-slot
// device ring of original and processed frames, a 6-slot output ring and its
// hyper-parameters.  Frames are submitted as they arrive; every tick takes the
// next ready window of every stream (continuous batching: a stream that
// arrives or is admitted between ticks joins at the next tick, while earlier
// work is still running).  Admission is by device-memory footprint against a
// budget; a refused stream can be opened again once others have closed.
//
// A tick, batched (the default): each stream's flow work (network features,
// forward/backward decode, consistency gate, upsampling) is enqueued on the
// stream's HIP stream from a pool, replayed from a captured graph in steady
// state; then the stabilization of every window of the tick (front pass,
// pyramid solve, RGBA8 out) runs as one captured graph on the server's tick
// stream: a copy of the tick's work list into device memory followed by
// persistent launches that walk it, one launch per solver step for all
// streams together (consecutive batches alternate between two tick streams,
// so one may start while the previous runs).  Unbatched (or for a window the batch cannot take: more
// than 2 pyramid levels, odd sizes at 2 levels), a stream's whole frame runs
// on its pool stream, replayed from per-stream graphs.
//
// Frame protocol per stream (k = 1, B = 1; videostabilizer.cpp:136-164): frame
// t's original and processed images are copied into ring slot t % 3 on the
// stream's HIP stream.  Submitting frame 2 sets `last` to processed frame 2
// (preloadProcessedFrames); from then on, submitting frame t makes window
// (t-2, t-1, t) ready, and a tick stabilizes frame t-1 into output slot
// (t-1) % 6 once that slot's previous frame has been reported by poll.  Frames
// 0 and the final frame are the processed frames unchanged, as in the
// reference (the caller owns them).
#include "vsc_internal.h"

#include <deque>
#include <map>
#include <new>
#include <string>
#include <tuple>
#include <vector>

namespace vsc {
namespace {

struct Done {
    long long tag;
    long long frame;
    int slot;   // output slot; its completion event is Stream::ev_done[slot]
};

// 6 output slots: every per-frame pointer then repeats with period 6 in the
// frame index (input ring t % 3, ping-pongs t % 2), so 6 graphs replay forever
constexpr int R_OUT = 6;
constexpr int N_PHASES = 6;

struct Stream {
    bool open = false;
    int W = 0, H = 0;
    size_t bytes = 0;
    vsc_pipeline* pipe = nullptr;
    uint8_t* orig[3] = {nullptr, nullptr, nullptr};
    uint8_t* proc[3] = {nullptr, nullptr, nullptr};
    uint8_t* out[R_OUT] = {};
    hipGraphExec_t ge[N_PHASES] = {};   // whole frames (unbatched)
    hipGraphExec_t gf[N_PHASES] = {};   // flow work only (batched)
    // completion event per output slot (a slot holds at most one frame in
    // flight), one for the flow work of the latest batched window, one for
    // reads: created at open, reused for every frame
    hipEvent_t ev_done[R_OUT] = {};
    hipEvent_t ev_flow = nullptr;
    hipEvent_t ev_read = nullptr;
    hipStream_t hs = nullptr;       // borrowed from the pool
    long long next_frame = 0;       // index of the next submitted frame
    long long ready = -1;           // highest window centre ready to run
    long long done_to = 0;          // highest window centre launched
    long long polled_to = 0;        // highest window centre reported by poll
    std::deque<long long> tags;     // tag of each ready window, in order
    vsc_params params{};
    std::vector<Done> inflight;
};

}  // namespace
}  // namespace vsc

struct vsc_server {
    vsc_server_cfg cfg;
    std::vector<float> ups_params;
    std::vector<vsc::Stream> s;
    std::vector<hipStream_t> pool;
    size_t used = 0;
    int rr = 0;
    // batched ticks: batches go round robin to tick streams, each with its
    // device work list and two pinned host work lists used alternately
    struct TickLane {
        hipStream_t ts = nullptr;
        unsigned char* wl_dev = nullptr;
        unsigned char* wl_host[2] = {};
        hipEvent_t ev_wl[2] = {};             // the copy that last read wl_host[i]
        bool wl_used[2] = {false, false};
        int wl_next = 0;
    };
    std::vector<TickLane> lanes;
    int lane_next = 0;
    int grid = 256;                           // blocks of a persistent launch (one per CU)
    std::vector<vsc::BatchFrame> items;
    std::vector<int> item_stream;
    // (lane, host list, launches at level 0, level 1)
    std::map<std::tuple<int, int, int, int>, hipGraphExec_t> batch_graphs;
};

namespace vsc {
namespace {

size_t ring_bytes(int W, int H) { return (size_t)W * H * 4 * (6 + R_OUT); }

void drop_graphs(Stream& st, bool flows_too)
{
    bool any = false;
    for (int i = 0; i < N_PHASES; ++i) any = any || st.ge[i] || (flows_too && st.gf[i]);
    if (any && st.hs) (void)hipStreamSynchronize(st.hs);
    for (int i = 0; i < N_PHASES; ++i) {
        if (st.ge[i]) (void)hipGraphExecDestroy(st.ge[i]);
        st.ge[i] = nullptr;
        if (flows_too) {
            if (st.gf[i]) (void)hipGraphExecDestroy(st.gf[i]);
            st.gf[i] = nullptr;
        }
    }
}

void release(Stream& st)
{
    drop_graphs(st, true);
    st.inflight.clear();
    for (int i = 0; i < R_OUT; ++i) {
        if (st.ev_done[i]) (void)hipEventDestroy(st.ev_done[i]);
        st.ev_done[i] = nullptr;
    }
    for (hipEvent_t* e : {&st.ev_flow, &st.ev_read}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    if (st.pipe) vsc_pipeline_destroy(st.pipe);
    st.pipe = nullptr;
    if (st.orig[0]) (void)hipFree(st.orig[0]);
    for (int i = 0; i < 3; ++i) st.orig[i] = st.proc[i] = nullptr;
    for (int i = 0; i < R_OUT; ++i) st.out[i] = nullptr;
    st.tags.clear();
    st.open = false;
}

// make `h` wait until window `c` of the stream has been stabilized (if it was launched)
int wait_window(const Stream& st, long long c, hipStream_t h)
{
    if (c < 1 || c > st.done_to) return VSC_OK;
    return check_hip(hipStreamWaitEvent(h, st.ev_done[c % R_OUT], 0), "vsc_server: wait for a window");
}

vsc_window window_of(const Stream& st, long long c)
{
    return vsc_window{st.orig[(c - 1) % 3], st.orig[c % 3], st.orig[(c + 1) % 3],
                      st.proc[(c - 1) % 3], st.proc[c % 3], st.proc[(c + 1) % 3]};
}

// capture `enqueue` (which enqueues on h) into a new graph exec
template <class F>
int capture(hipStream_t h, hipGraphExec_t* out, F enqueue)
{
    hipGraph_t g = nullptr;
    if (check_hip(hipStreamBeginCapture(h, hipStreamCaptureModeThreadLocal), "vsc_server_tick capture"))
        return VSC_E_HIP;
    const int rc = enqueue();
    const int rce = check_hip(hipStreamEndCapture(h, &g), "vsc_server_tick capture end");
    if (rc || rce) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : rce;
    }
    const int rci = check_hip(hipGraphInstantiate(out, g, nullptr, nullptr, 0), "vsc_server_tick instantiate");
    (void)hipGraphDestroy(g);
    if (rci) *out = nullptr;
    return rci;
}

// A whole frame of window c on the stream's pool stream (the unbatched path).
int frame_alone(vsc_server* sv, Stream& st, long long c)
{
    const vsc_window w = window_of(st, c);
    int rc;
    // the previous frame may have been stabilized in a batch on the tick stream
    if ((rc = wait_window(st, c - 1, st.hs))) return rc;
    const int ph = (int)(c % N_PHASES);
    auto enqueue = [&]() {
        return vsc_pipeline_frame(st.pipe, &w, c, &st.params, st.out[c % R_OUT], nullptr,
                                  reinterpret_cast<vsc_hip_stream>(st.hs));
    };
    if (sv->cfg.use_graphs && c >= 3 && st.ge[ph]) {
        // steady state: step the host state, replay the phase's graph
        if ((rc = pipeline_advance(st.pipe, c))) return rc;
        if (check_hip(hipGraphLaunch(st.ge[ph], st.hs), "vsc_server_tick graph launch")) return VSC_E_HIP;
    } else if (sv->cfg.use_graphs && c >= 3) {
        // a failed capture leaves the pipeline's host state as it was, but
        // for the feature slot it began to fill, which reads as empty
        // (vsc_pipeline_frame restores it), so the frame is retried whole
        if ((rc = capture(st.hs, &st.ge[ph], enqueue))) return rc;
        if (check_hip(hipGraphLaunch(st.ge[ph], st.hs), "vsc_server_tick graph launch")) return VSC_E_HIP;
    } else if ((rc = enqueue())) {
        return rc;
    }
    return check_hip(hipEventRecord(st.ev_done[c % R_OUT], st.hs), "vsc_server_tick");
}

// The flow work of window c on the stream's pool stream, and its batch item
// for the tick stream ts.
int frame_flows_batched(vsc_server* sv, Stream& st, long long c, hipStream_t ts, BatchFrame* item)
{
    const vsc_window w = window_of(st, c);
    int rc;
    // the flows of window c overwrite the buffers window c-2 stabilized from;
    // the batch continues from window c-1 (perhaps on another tick stream)
    if ((rc = wait_window(st, c - 2, st.hs)) || (rc = wait_window(st, c - 1, ts))) return rc;
    const int ph = (int)(c % N_PHASES);
    auto enqueue = [&]() { return pipeline_flows_part(st.pipe, &w, c, st.hs); };
    if (sv->cfg.use_graphs && c >= 3 && st.gf[ph]) {
        if ((rc = pipeline_advance(st.pipe, c))) return rc;
        if (check_hip(hipGraphLaunch(st.gf[ph], st.hs), "vsc_server_tick graph launch")) return VSC_E_HIP;
    } else if (sv->cfg.use_graphs && c >= 3) {
        if ((rc = capture(st.hs, &st.gf[ph], enqueue))) return rc;
        if (check_hip(hipGraphLaunch(st.gf[ph], st.hs), "vsc_server_tick graph launch")) return VSC_E_HIP;
    } else if ((rc = enqueue())) {
        return rc;
    }
    if (check_hip(hipEventRecord(st.ev_flow, st.hs), "vsc_server_tick") ||
        check_hip(hipStreamWaitEvent(ts, st.ev_flow, 0), "vsc_server_tick"))
        return VSC_E_HIP;
    return pipeline_batch_item(st.pipe, &w, c, &st.params, st.out[c % R_OUT], item);
}

// The stabilization of this round's batch items on tick lane `li`.
int run_batch(vsc_server* sv, int li)
{
    vsc_server::TickLane& lane = sv->lanes[li];
    const int n = (int)sv->items.size();
    const int fma = sv->cfg.form.solver ? 1 : 0;
    const int b = lane.wl_next;
    lane.wl_next ^= 1;
    // the host list is refilled only after its previous copy has run
    if (lane.wl_used[b] && check_hip(hipEventSynchronize(lane.ev_wl[b]), "vsc_server_tick")) return VSC_E_HIP;
    // the layout has room for every stream, so a captured graph depends only
    // on the lane, the host buffer and the launch counts, not on how many
    // streams this round batches
    const int cap = sv->cfg.max_streams;
    batch_build(lane.wl_host[b], cap, sv->items.data(), n, fma);
    const BatchHeader* hd = reinterpret_cast<const BatchHeader*>(lane.wl_host[b]);
    const int nl[2] = {hd->n_launch[0], hd->n_launch[1]};
    const size_t bytes = batch_worklist_bytes(cap);
    auto enqueue = [&]() {
        if (check_hip(hipMemcpyAsync(lane.wl_dev, lane.wl_host[b], bytes, hipMemcpyHostToDevice, lane.ts),
                      "vsc_server_tick work list"))
            return VSC_E_HIP;
        return launch_batch(lane.wl_dev, nl, fma, sv->grid, lane.ts);
    };
    int rc;
    if (sv->cfg.use_graphs) {
        const auto key = std::make_tuple(li, b, nl[0], nl[1]);
        auto it = sv->batch_graphs.find(key);
        if (it == sv->batch_graphs.end()) {
            hipGraphExec_t ge = nullptr;
            if ((rc = capture(lane.ts, &ge, enqueue))) return rc;
            it = sv->batch_graphs.emplace(key, ge).first;
        }
        if (check_hip(hipGraphLaunch(it->second, lane.ts), "vsc_server_tick batch launch")) return VSC_E_HIP;
    } else if ((rc = enqueue())) {
        return rc;
    }
    if (check_hip(hipEventRecord(lane.ev_wl[b], lane.ts), "vsc_server_tick")) return VSC_E_HIP;
    lane.wl_used[b] = true;
    return VSC_OK;
}

void launched(Stream& st, long long c)
{
    st.inflight.push_back(Done{st.tags.front(), c, (int)(c % R_OUT)});
    st.tags.pop_front();
    st.done_to = c;
}

bool has_window(const Stream& st)
{
    return st.open && st.done_to < st.ready && st.done_to + 1 - R_OUT <= st.polled_to;
}

}  // namespace
}  // namespace vsc

using namespace vsc;

extern "C" void vsc_server_default_cfg(vsc_server_cfg* c)
{
    if (!c) return;
    c->max_streams = 64;
    c->hbm_budget = (size_t)64 << 30;
    c->hip_streams = 4;
    vsc_pipeline_default_cfg(&c->form);
    c->form.flow_dtype = VSC_BF16;
    c->form.downscale = 2;
    c->form.upsample = 1;
    c->use_graphs = 1;
    c->batch = 1;
}

extern "C" int vsc_server_create(const vsc_server_cfg* cfg, const float* upsampler_params, vsc_server** out)
{
    if (!cfg || !out) return set_error(VSC_E_INVALID, "vsc_server_create: NULL argument");
    *out = nullptr;
    if (cfg->max_streams < 1 || cfg->max_streams > 4096 || cfg->hip_streams < 1 || cfg->hip_streams > 16)
        return set_error(VSC_E_INVALID, "vsc_server_create: max_streams in [1, 4096], hip_streams in [1, 16]");
    if (cfg->form.overlap != 0)
        return set_error(VSC_E_INVALID, "vsc_server_create: form.overlap must be 0 (the server orders each stream's "
                                        "frames on its own HIP stream)");
    if (cfg->form.flow_batch != 1)
        return set_error(VSC_E_INVALID, "vsc_server_create: form.flow_batch must be 1 (a stream's window advances "
                                        "one frame per tick; the tick batches across streams instead)");
    if (cfg->batch != 0 && cfg->batch != 1) return set_error(VSC_E_INVALID, "vsc_server_create: batch must be 0 or 1");
    if (cfg->form.downscale > 1 && cfg->form.upsample == 1 && !upsampler_params)
        return set_error(VSC_E_INVALID, "vsc_server_create: the learned upsampler needs its parameters");
    vsc_server* sv = new (std::nothrow) vsc_server();
    if (!sv) return set_error(VSC_E_NOMEM, "vsc_server_create: host memory");
    sv->cfg = *cfg;
    if (upsampler_params) sv->ups_params.assign(upsampler_params, upsampler_params + vsc_upsampler_param_count());
    sv->s.resize(cfg->max_streams);
    auto fail = [&](const char* what) {
        vsc_server_destroy(sv);
        return set_error(VSC_E_HIP, std::string("vsc_server_create: ") + what);
    };
    for (int i = 0; i < cfg->hip_streams; ++i) {
        hipStream_t h = nullptr;
        if (hipStreamCreateWithFlags(&h, hipStreamNonBlocking) != hipSuccess) return fail("stream creation failed");
        sv->pool.push_back(h);
    }
    if (cfg->batch) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return fail("device query failed");
        sv->grid = cus >= 8 ? cus / 8 * 8 : 8;
        const size_t wl = batch_worklist_bytes(cfg->max_streams);
        // two tick streams: a batch may start while the previous one runs
        // when their streams differ (a stream's frames stay in order through
        // the wait on its previous window)
        sv->lanes.resize(2);
        for (vsc_server::TickLane& lane : sv->lanes) {
            if (hipStreamCreateWithFlags(&lane.ts, hipStreamNonBlocking) != hipSuccess)
                return fail("stream creation failed");
            if (hipMalloc(&lane.wl_dev, wl) != hipSuccess) return fail("work list allocation failed");
            for (int i = 0; i < 2; ++i)
                if (hipHostMalloc(&lane.wl_host[i], wl, hipHostMallocDefault) != hipSuccess ||
                    hipEventCreateWithFlags(&lane.ev_wl[i], hipEventDisableTiming) != hipSuccess)
                    return fail("work list allocation failed");
        }
        sv->items.reserve(cfg->max_streams);
        sv->item_stream.reserve(cfg->max_streams);
    }
    *out = sv;
    return VSC_OK;
}

extern "C" int vsc_server_destroy(vsc_server* sv)
{
    if (!sv) return VSC_OK;
    for (hipStream_t h : sv->pool) (void)hipStreamSynchronize(h);
    for (vsc_server::TickLane& lane : sv->lanes)
        if (lane.ts) (void)hipStreamSynchronize(lane.ts);
    for (Stream& st : sv->s)
        if (st.open) release(st);
    for (auto& kv : sv->batch_graphs) (void)hipGraphExecDestroy(kv.second);
    for (vsc_server::TickLane& lane : sv->lanes) {
        for (int i = 0; i < 2; ++i) {
            if (lane.wl_host[i]) (void)hipHostFree(lane.wl_host[i]);
            if (lane.ev_wl[i]) (void)hipEventDestroy(lane.ev_wl[i]);
        }
        if (lane.wl_dev) (void)hipFree(lane.wl_dev);
        if (lane.ts) (void)hipStreamDestroy(lane.ts);
    }
    for (hipStream_t h : sv->pool) (void)hipStreamDestroy(h);
    delete sv;
    return VSC_OK;
}

extern "C" size_t vsc_server_stream_footprint(const vsc_server* sv, int width, int height)
{
    if (!sv || width <= 0 || height <= 0) return 0;
    return vsc_pipeline_footprint(width, height, &sv->cfg.form) + ring_bytes(width, height);
}

extern "C" size_t vsc_server_used_bytes(const vsc_server* sv) { return sv ? sv->used : 0; }

extern "C" int vsc_server_open(vsc_server* sv, int width, int height, int* id)
{
    if (!sv || !id) return set_error(VSC_E_INVALID, "vsc_server_open: NULL argument");
    *id = -1;
    if (width <= 0 || height <= 0) return set_error(VSC_E_INVALID, "vsc_server_open: bad frame size");
    int slot = -1;
    for (int i = 0; i < (int)sv->s.size(); ++i)
        if (!sv->s[i].open) { slot = i; break; }
    if (slot < 0) return set_error(VSC_E_NOMEM, "vsc_server_open: all stream slots are in use");
    const size_t need = vsc_server_stream_footprint(sv, width, height);
    if (need == 0) return set_error(VSC_E_INVALID, "vsc_server_open: invalid configuration");
    if (sv->used + need > sv->cfg.hbm_budget)
        return set_error(VSC_E_NOMEM, "vsc_server_open: device memory budget exhausted (retry after a stream closes)");
    Stream& st = sv->s[slot];
    st = Stream();
    st.W = width; st.H = height;
    int rc = vsc_pipeline_create(width, height, &sv->cfg.form, &st.pipe);
    if (!rc && !sv->ups_params.empty() && sv->cfg.form.upsample == 1 && sv->cfg.form.downscale > 1)
        rc = vsc_pipeline_set_upsampler_params(st.pipe, sv->ups_params.data());
    if (!rc) {
        uint8_t* ring = nullptr;
        const size_t fb = (size_t)width * height * 4;
        if (hipMalloc(&ring, fb * (6 + R_OUT)) != hipSuccess) {
            rc = set_error(VSC_E_NOMEM, "vsc_server_open: ring allocation failed");
        } else {
            for (int i = 0; i < 3; ++i) {
                st.orig[i] = ring + fb * i;
                st.proc[i] = ring + fb * (3 + i);
            }
            for (int i = 0; i < R_OUT; ++i) st.out[i] = ring + fb * (6 + i);
        }
    }
    for (int i = 0; i < R_OUT && !rc; ++i)
        rc = check_hip(hipEventCreateWithFlags(&st.ev_done[i], hipEventDisableTiming), "vsc_server_open");
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&st.ev_flow, hipEventDisableTiming), "vsc_server_open");
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&st.ev_read, hipEventDisableTiming), "vsc_server_open");
    if (rc) {
        release(st);
        return rc;
    }
    vsc_default_params(&st.params);
    st.hs = sv->pool[sv->rr++ % sv->pool.size()];
    st.bytes = need;
    st.open = true;
    sv->used += need;
    *id = slot;
    return VSC_OK;
}

static Stream* get(vsc_server* sv, int id)
{
    if (!sv || id < 0 || id >= (int)sv->s.size() || !sv->s[id].open) return nullptr;
    return &sv->s[id];
}

extern "C" int vsc_server_close(vsc_server* sv, int id)
{
    Stream* st = get(sv, id);
    if (!st) return set_error(VSC_E_INVALID, "vsc_server_close: no such stream");
    (void)hipStreamSynchronize(st->hs);
    for (vsc_server::TickLane& lane : sv->lanes)
        if (lane.ts) (void)hipStreamSynchronize(lane.ts);
    sv->used -= st->bytes;
    release(*st);
    return VSC_OK;
}

extern "C" int vsc_server_set_params(vsc_server* sv, int id, const vsc_params* p)
{
    Stream* st = get(sv, id);
    if (!st || !p) return set_error(VSC_E_INVALID, "vsc_server_set_params: no such stream or NULL params");
    vsc_stream* vs = nullptr;
    int rc = vsc_pipeline_stream(st->pipe, &vs);
    if (rc || (rc = stream_params_check(vs, p))) return rc;   // refused: the stream keeps its parameters
    st->params = *p;   // read by the next tick (hyperParams are copied per frame, videostabilizer.cpp:192)
    drop_graphs(*st, false);   // whole-frame graphs hold the old values as kernel arguments
    return VSC_OK;
}

extern "C" int vsc_server_submit(vsc_server* sv, int id, const uint8_t* orig, const uint8_t* proc, long long tag)
{
    Stream* st = get(sv, id);
    if (!st || !orig || !proc) return set_error(VSC_E_INVALID, "vsc_server_submit: no such stream or NULL frame");
    const long long t = st->next_frame;
    if (t >= 3 && st->done_to < t - 2)
        return set_error(VSC_E_AGAIN, "vsc_server_submit: window t-2 not launched yet (poll, tick, retry)");
    const size_t fb = (size_t)st->W * st->H * 4;
    // the slot's previous frame (t - 3) was last read by window t - 2 (its
    // flow work on this HIP stream, its stabilization here or on the tick
    // stream): the copies wait for that window
    if (wait_window(*st, t - 2, st->hs) ||
        check_hip(hipMemcpyAsync(st->orig[t % 3], orig, fb, hipMemcpyDefault, st->hs), "vsc_server_submit") ||
        check_hip(hipMemcpyAsync(st->proc[t % 3], proc, fb, hipMemcpyDefault, st->hs), "vsc_server_submit"))
        return VSC_E_HIP;
    if (t == 2) {
        vsc_stream* vs = nullptr;
        int rc = vsc_pipeline_stream(st->pipe, &vs);
        if (!rc) rc = vsc_stream_set_last_rgba(vs, st->proc[2], reinterpret_cast<vsc_hip_stream>(st->hs));
        if (rc) return rc;
    }
    if (t >= 2) {
        st->ready = t - 1;
        st->tags.push_back(tag);
    }
    st->next_frame = t + 1;
    return VSC_OK;
}

extern "C" int vsc_server_tick(vsc_server* sv, int* launched_out)
{
    if (!sv) return set_error(VSC_E_INVALID, "vsc_server_tick: NULL server");
    // a stream whose frame fails keeps that window ready (it is retried at the
    // next tick) and does not hold up the other streams; the first failure's
    // status is returned, its message in vsc_last_error
    int n = 0, first_rc = VSC_OK;
    std::string first_msg;
    auto note = [&](int rc) {
        if (rc && first_rc == VSC_OK) {
            first_rc = rc;
            first_msg = vsc_last_error();
        }
    };
    std::vector<char> failed(sv->s.size(), 0);
    // rounds: each takes the next ready window of every stream
    for (bool more = true; more;) {
        more = false;
        sv->items.clear();
        sv->item_stream.clear();
        const int li = sv->lanes.empty() ? 0 : sv->lane_next;
        hipStream_t ts = sv->lanes.empty() ? nullptr : sv->lanes[li].ts;
        for (int id = 0; id < (int)sv->s.size(); ++id) {
            Stream& st = sv->s[id];
            if (failed[id] || !has_window(st)) continue;
            const long long c = st.done_to + 1;
            int rc;
            if (sv->cfg.batch && pipeline_batch_check(st.pipe, &st.params) == VSC_OK) {
                sv->items.emplace_back();
                rc = frame_flows_batched(sv, st, c, ts, &sv->items.back());
                if (rc) {
                    sv->items.pop_back();
                } else {
                    sv->item_stream.push_back(id);
                }
            } else {
                rc = frame_alone(sv, st, c);
                if (!rc) {
                    launched(st, c);
                    ++n;
                }
            }
            if (rc) {
                failed[id] = 1;
                note(rc);
            }
        }
        if (!sv->items.empty()) {
            const int rc = run_batch(sv, li);
            if (rc) return rc;   // the device is in an unknown state: report at once
            sv->lane_next = (li + 1) % (int)sv->lanes.size();
            for (int id : sv->item_stream) {
                Stream& st = sv->s[id];
                const long long c = st.done_to + 1;
                if (check_hip(hipEventRecord(st.ev_done[c % R_OUT], ts), "vsc_server_tick")) return VSC_E_HIP;
                launched(st, c);
                ++n;
            }
        }
        for (int id = 0; id < (int)sv->s.size() && !more; ++id)
            more = !failed[id] && has_window(sv->s[id]);
    }
    if (launched_out) *launched_out = n;
    if (first_rc) return set_error(first_rc, first_msg);
    return VSC_OK;
}

extern "C" int vsc_server_poll(vsc_server* sv, vsc_server_done* done, int max, int* count)
{
    if (!sv || !count || (max > 0 && !done)) return set_error(VSC_E_INVALID, "vsc_server_poll: NULL argument");
    int n = 0;
    for (int id = 0; id < (int)sv->s.size(); ++id) {
        Stream& st = sv->s[id];
        if (!st.open) continue;
        size_t k = 0;
        while (k < st.inflight.size() && n < max) {
            Done& d = st.inflight[k];
            const hipError_t q = hipEventQuery(st.ev_done[d.slot]);
            if (q == hipErrorNotReady) break;   // later frames of this stream are not done either
            if (q != hipSuccess) return check_hip(q, "vsc_server_poll");
            done[n].stream = id;
            done[n].tag = d.tag;
            done[n].frame = d.frame;
            done[n].out_rgba = st.out[d.slot];
            st.polled_to = d.frame;
            ++n;
            ++k;
        }
        st.inflight.erase(st.inflight.begin(), st.inflight.begin() + k);
    }
    *count = n;
    return VSC_OK;
}

extern "C" int vsc_server_drain(vsc_server* sv)
{
    if (!sv) return set_error(VSC_E_INVALID, "vsc_server_drain: NULL server");
    for (hipStream_t h : sv->pool)
        if (check_hip(hipStreamSynchronize(h), "vsc_server_drain")) return VSC_E_HIP;
    for (vsc_server::TickLane& lane : sv->lanes)
        if (check_hip(hipStreamSynchronize(lane.ts), "vsc_server_drain")) return VSC_E_HIP;
    return VSC_OK;
}

extern "C" int vsc_server_read(vsc_server* sv, int id, long long frame, uint8_t* dst, vsc_hip_stream stream)
{
    Stream* st = get(sv, id);
    if (!st || !dst) return set_error(VSC_E_INVALID, "vsc_server_read: no such stream or NULL destination");
    if (frame < 1 || frame > st->polled_to || frame + R_OUT <= st->done_to)
        return set_error(VSC_E_INVALID, "vsc_server_read: frame not reported or no longer in the output ring");
    const size_t fb = (size_t)st->W * st->H * 4;
    // the read event is re-recorded by each read: a wait enqueued on `stream`
    // holds the state it had when the wait was enqueued
    if (wait_window(*st, frame, st->hs) ||
        check_hip(hipMemcpyAsync(dst, st->out[frame % R_OUT], fb, hipMemcpyDefault, st->hs), "vsc_server_read") ||
        check_hip(hipEventRecord(st->ev_read, st->hs), "vsc_server_read") ||
        check_hip(hipStreamWaitEvent(as_stream(stream), st->ev_read, 0), "vsc_server_read"))
        return VSC_E_HIP;
    return VSC_OK;
}
