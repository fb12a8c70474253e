// The flow-file mode of the file stabilizer (`-f <directory>`): FileStabilizer
// (stabilizefiles.cpp:99-149) driving VideoStabilizer's frame loop
// (preloadProcessedFrames / doOneStep / outputFinalFrames,
// videostabilizer.cpp:136-238) with flows read from Middlebury .flo files:
// for frame i the forward flow (i -> i+1) is frame_<i+1>.flo and the backward
// flow (i -> i-1) is frame_<i>_bwd.flo, indices zero-padded to 6 digits
// (stabilizefiles.cpp:99-101,134-149).  A flow whose size differs from the
// frame is refused, as initializeFlowImage throws (imagehelpers.cpp:42-50).
// Frames arrive as host RGBA8 arrays (image decoding stays with the caller)
// and leave through a callback, in frame order: frame 0 and the last frame
// are the processed frames (through the float image and back, as the
// reference's gpuToImage writes them), frames 1..n-2 are stabilized.
#include "vsc_internal.h"

#include <cinttypes>
#include <cstdio>
#include <string>
#include <vector>

using vsc::check_hip;
using vsc::set_error;

namespace {

std::string flo_name(const char* dir, long long index, bool bwd)
{
    char name[64];
    std::snprintf(name, sizeof(name), "frame_%06lld%s.flo", index, bwd ? "_bwd" : "");
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    return d + name;
}

struct Buffers {
    uint8_t* dev = nullptr;     // 3 original + 3 processed RGBA8 frames, the output frame
    float* flow = nullptr;      // forward, backward (HWC, 2 channels), then one HWC 3-channel image
    void* host = nullptr;       // pinned: forward, backward flows, then the output frame
    ~Buffers()
    {
        if (dev) (void)hipFree(dev);
        if (flow) (void)hipFree(flow);
        if (host) (void)hipHostFree(host);
    }
};

}  // namespace

extern "C" int vsc_flow_dir_path(const char* dir, long long frame, int backward, char* path, size_t capacity)
{
    if (!dir || !path) return set_error(VSC_E_INVALID, "vsc_flow_dir_path: NULL argument");
    if (frame < 0) return set_error(VSC_E_INVALID, "vsc_flow_dir_path: negative frame index");
    const std::string p = flo_name(dir, backward ? frame : frame + 1, backward != 0);
    if (p.size() + 1 > capacity) return set_error(VSC_E_INVALID, "vsc_flow_dir_path: path buffer too small");
    std::snprintf(path, capacity, "%s", p.c_str());
    return VSC_OK;
}

extern "C" int vsc_flow_dir_read(const char* dir, long long frame, float* fwd, float* bwd, size_t capacity_floats,
                                 int* width, int* height)
{
    if (!dir || !width || !height) return set_error(VSC_E_INVALID, "vsc_flow_dir_read: NULL argument");
    if (frame < 0) return set_error(VSC_E_INVALID, "vsc_flow_dir_read: negative frame index");
    int wf = 0, hf = 0, wb = 0, hb = 0;
    const std::string pf = flo_name(dir, frame + 1, false), pb = flo_name(dir, frame, true);
    int rc = vsc_flo_read(pf.c_str(), nullptr, 0, &wf, &hf);
    if (!rc) rc = vsc_flo_read(pb.c_str(), nullptr, 0, &wb, &hb);
    if (rc) return rc;
    if (wf != wb || hf != hb)
        return set_error(VSC_E_INVALID, "vsc_flow_dir_read: forward and backward flows differ in size (" + pf + ", " +
                                            pb + ")");
    *width = wf;
    *height = hf;
    if (!fwd && !bwd) return VSC_OK;
    if ((fwd && (rc = vsc_flo_read(pf.c_str(), fwd, capacity_floats, &wf, &hf))) ||
        (bwd && (rc = vsc_flo_read(pb.c_str(), bwd, capacity_floats, &wb, &hb))))
        return rc;
    return VSC_OK;
}

extern "C" int vsc_stabilize_flow_dir(const uint8_t* const* orig_rgba, const uint8_t* const* proc_rgba, int nframes,
                                      int width, int height, const char* flow_dir, const vsc_params* params, int mode,
                                      vsc_frame_sink sink, void* user, vsc_hip_stream stream)
{
    if (!orig_rgba || !proc_rgba || !flow_dir || !sink)
        return set_error(VSC_E_INVALID, "vsc_stabilize_flow_dir: NULL argument");
    if (nframes < 3) return set_error(VSC_E_INVALID, "vsc_stabilize_flow_dir: needs at least 3 frames (k = 1)");
    if (mode != VSC_FRAME_FUSED && mode != VSC_FRAME_FMA)
        return set_error(VSC_E_INVALID, "vsc_stabilize_flow_dir: mode must be VSC_FRAME_FUSED or VSC_FRAME_FMA");
    for (int i = 0; i < nframes; ++i)
        if (!orig_rgba[i] || !proc_rgba[i]) return set_error(VSC_E_INVALID, "vsc_stabilize_flow_dir: NULL frame");
    hipStream_t st = vsc::as_stream(stream);
    vsc_stream* s = nullptr;
    int rc = vsc_stream_create(width, height, width, height, &s);
    if (rc) return rc;
    struct Guard {
        vsc_stream* s;
        ~Guard() { vsc_stream_destroy(s); }
    } guard{s};
    const size_t fb = (size_t)width * height * 4, flow_floats = (size_t)width * height * 2;
    Buffers b;
    if (check_hip(hipMalloc(&b.dev, fb * 7), "vsc_stabilize_flow_dir") ||
        check_hip(hipMalloc(&b.flow, (2 * flow_floats + (size_t)width * height * 3) * sizeof(float)),
                  "vsc_stabilize_flow_dir") ||
        check_hip(hipHostMalloc(&b.host, 2 * flow_floats * sizeof(float) + fb, hipHostMallocDefault),
                  "vsc_stabilize_flow_dir"))
        return VSC_E_HIP;
    uint8_t* orig[3] = {b.dev, b.dev + fb, b.dev + 2 * fb};
    uint8_t* proc[3] = {b.dev + 3 * fb, b.dev + 4 * fb, b.dev + 5 * fb};
    uint8_t* out = b.dev + 6 * fb;
    float* hfwd = static_cast<float*>(b.host);
    float* hbwd = hfwd + flow_floats;
    uint8_t* hout = reinterpret_cast<uint8_t*>(hbwd + flow_floats);
    auto upload = [&](int i) {
        return check_hip(hipMemcpyAsync(orig[i % 3], orig_rgba[i], fb, hipMemcpyHostToDevice, st), "upload") ||
               check_hip(hipMemcpyAsync(proc[i % 3], proc_rgba[i], fb, hipMemcpyHostToDevice, st), "upload");
    };
    // a processed frame leaves as the reference writes it: through the
    // device's float image and back (gpuToImage of the GPUImage, A11 then A12:
    // alpha becomes 1)
    const vsc_image img{b.flow + 2 * flow_floats, width, height, 3};
    auto emit_processed = [&](int i, const uint8_t* dev_rgba) {
        int r = vsc_rgba_to_float(dev_rgba, img, stream);
        if (!r) r = vsc_float_to_rgba(img, out, stream);
        if (!r && (check_hip(hipMemcpyAsync(hout, out, fb, hipMemcpyDeviceToHost, st), "vsc_stabilize_flow_dir") ||
                   check_hip(hipStreamSynchronize(st), "vsc_stabilize_flow_dir")))
            r = VSC_E_HIP;
        if (!r && sink(user, i, hout, width, height))
            r = set_error(VSC_E_INVALID, "vsc_stabilize_flow_dir: stopped by the sink");
        return r;
    };
    // preloadProcessedFrames (k = 1, B = 1): frame 0 leaves as processed, the
    // first `last` is processed frame 2k+B-1 = 2
    for (int i = 0; i < 3; ++i)
        if (upload(i)) return VSC_E_HIP;
    if ((rc = emit_processed(0, proc[0]))) return rc;
    if ((rc = vsc_stream_set_last_rgba(s, proc[2], stream))) return rc;
    for (int i = 1; i + 1 < nframes; ++i) {
        if (i >= 2 && upload(i + 1)) return VSC_E_HIP;   // slot (i+1) % 3 last held frame i-2, read by frame i-1
        int fw = 0, fh = 0;
        // the copies of the previous frame's flows must finish before the
        // pinned buffers are refilled (the output read below synchronised)
        if ((rc = vsc_flow_dir_read(flow_dir, i, hfwd, hbwd, flow_floats, &fw, &fh))) return rc;
        if (fw != width || fh != height)
            return set_error(VSC_E_INVALID, "vsc_stabilize_flow_dir: flow image size does not match image size (frame " +
                                                std::to_string(i) + ")");
        if (check_hip(hipMemcpyAsync(b.flow, hfwd, 2 * flow_floats * sizeof(float), hipMemcpyHostToDevice, st),
                      "vsc_stabilize_flow_dir"))
            return VSC_E_HIP;
        const vsc_frame_in in{orig[(i - 1) % 3], orig[i % 3], orig[(i + 1) % 3], proc[(i - 1) % 3], proc[i % 3],
                              proc[(i + 1) % 3], b.flow, b.flow + flow_floats, 2};
        if ((rc = vsc_stab_frame(s, &in, params, out, nullptr, mode, stream))) return rc;
        if (check_hip(hipMemcpyAsync(hout, out, fb, hipMemcpyDeviceToHost, st), "vsc_stabilize_flow_dir") ||
            check_hip(hipStreamSynchronize(st), "vsc_stabilize_flow_dir"))
            return VSC_E_HIP;
        if (sink(user, i, hout, width, height))
            return set_error(VSC_E_INVALID, "vsc_stabilize_flow_dir: stopped by the sink");
    }
    // outputFinalFrames: the last frame leaves as processed (it is in ring
    // slot (nframes-1) % 3 since the last window)
    return emit_processed(nframes - 1, proc[(nframes - 1) % 3]);
}
