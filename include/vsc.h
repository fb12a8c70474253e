/*
 * vsc.h — C ABI of the MI355X (gfx950) per-frame consistency hot path.
 *
 * Drop-in boundary for two parts of MaxReimann/video-stream-consistency:
 *   1. the ONNX Runtime custom ops "Correlation" and "Warp" of
 *      src/ort_custom_ops (same tensor layouts, attributes and output shapes);
 *   2. the per-frame stabilization entry points of src/stabilization
 *      (get_warp_result, get_adap_comb, get_consist_wt, get_consist_out,
 *      get_bilinear, perform_consistency, the GPUImage u8<->float conversions
 *      and VideoStabilizer::doOneStep with its hyperParams).
 *
 * Conventions (differences from the reference are deliberate, see each entry):
 *   - every entry point returns int status: VSC_OK (0) or a negative VSC_E_*;
 *     nothing throws or calls exit() across this boundary.  vsc_last_error()
 *     returns a thread-local message describing the last failure.
 *   - all pointers are device pointers unless stated otherwise; the caller owns
 *     every buffer it passes.  All work is enqueued asynchronously on the
 *     caller's hipStream_t (NULL = the null stream); nothing synchronises the
 *     device except where stated.
 *   - images are HWC interleaved float, element (x,y,c) at
 *     data[(y*width + x)*channels + c] (the GPUImageWrapper layout,
 *     src/stabilization/gpuimagewrapper.h:8-25).
 */
#ifndef VSC_H
#define VSC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t without pulling in the HIP headers. */
typedef struct ihipStream_t* vsc_hip_stream;

#define VSC_OK 0
#define VSC_E_INVALID (-1)   /* bad shape, attribute, null pointer or parameter   */
#define VSC_E_DTYPE (-2)     /* unsupported element type                         */
#define VSC_E_HIP (-3)       /* a HIP runtime call or kernel launch failed        */
#define VSC_E_NOMEM (-4)     /* device allocation failed                          */
#define VSC_E_UNSUPPORTED (-5) /* valid request this build does not implement     */
#define VSC_E_AGAIN (-6)     /* not now: a ring slot is still in use; poll, tick and retry */

typedef enum {
    VSC_F32 = 0,
    VSC_F64 = 1,
    VSC_BF16 = 2,
    VSC_F16 = 3,
    VSC_FP8E4M3 = 4
} vsc_dtype;

/* Device image view: replaces GPUImage / GPUImageWrapper
 * (src/stabilization/gpuimage.h:8-31, gpuimagewrapper.h:8-25). */
typedef struct {
    float* data;
    int width;
    int height;
    int channels;
} vsc_image;

/* Consistency hyper-parameters: replaces struct hyperParams
 * (src/stabilization/videostabilizer.h:38-46). */
typedef struct {
    float alpha;        /* photometric weight of the adaptive combination     */
    float beta;         /* photometric weight of the consistency weights      */
    float gamma;        /* consistency strength (the interactive flicker knob) */
    int pyramid_levels; /* solver pyramid depth (1..VSC_MAX_LEVELS)            */
    int num_iter;       /* sweeps at level 0; level j runs num_iter/(j+1)      */
    float step_size;    /* gradient step                                      */
    float mom_fac;      /* momentum factor                                    */
} vsc_params;

#define VSC_MAX_LEVELS 4

const char* vsc_last_error(void);
const char* vsc_version(void);

/* Defaults of VideoStabilizer::initHyperParams (videostabilizer.cpp:104-113):
 * alpha = beta = 6800, gamma = 2, 2 levels, 150 iterations, step = mom = 0.15. */
int vsc_default_params(vsc_params* out);

/* ---------------------------------------------------------------------------
 * ORT custom op "Correlation".
 * Replaces CorrelationKernel::Compute (ort_custom_ops/opticalflow/correlation.h:35-50)
 * -> ComputeCPU (src/opticalflow/correlation.cc:45-143) / ComputeCUDA
 * (src/opticalflow/correlation_cuda.cc:29-138).
 * f1, f2: NCHW [N,C,H,W] of in_dtype.  out: FP32 (FP64 when in_dtype is F64)
 *   non-legacy: [N, 2md+1, 2md+1, H, W], out = sum_c f1 * f2(shifted), 0 outside;
 *   legacy:     [N, (2md+1)^2, H, W],  out = (sum_c f1 * f2(shifted)) / C,
 *               zero padding, md <= 4 (correlation_cuda.cu:182-331,359-360).
 * in_dtype F32 accumulates sequentially in c in FP32 (the reference CPU op's
 * order); BF16 / F16 / FP8E4M3 inputs are NCHW arrays of that element type
 * (raw bit patterns, 2/2/1 bytes; FP8 is OCP E4M3, bias 7) multiplied on MFMA
 * with FP32 accumulation, max_displacement <= 4.
 * scale (FP8 only, nullable = 1): device pointer to one float s, out = s *
 *   sum_c q1_c * q2_c.  Per-channel scaling is folded into the operands by
 *   vsc_quantize_fp8_pair (equalisation), so one factor dequantises.
 * Errors: VSC_E_INVALID for N,C,H,W < 1, md < 0, legacy with md > 4;
 *         VSC_E_DTYPE for an unknown in_dtype (the reference throws
 *         "Unsupported input type", correlation.cc:140-141).                  */
int vsc_correlation_forward(const void* f1, const void* f2, void* out, int N, int C, int H, int W,
                            int max_displacement, int legacy, int in_dtype, const float* scale,
                            vsc_hip_stream stream);
/* The same op with a caller-owned workspace (no allocation in the call; the
 * reference allocates and frees its channels-last temporaries on every call,
 * correlation_cuda.cu:365-370,440-441).  BF16 / F16 / FP8E4M3 inputs whose
 * width is a multiple of 4 and whose pointers are 8-byte (FP8: 4-byte)
 * aligned are read in their NCHW layout by the MFMA kernel and need no
 * workspace (NULL is accepted); other reduced-precision inputs are copied
 * channels-last into the workspace (two tensors of N*H*W*ceil32(C)
 * elements), whose size vsc_correlation_workspace_size gives from the shape
 * alone.  F32 and F64 need none (size 0, workspace may be NULL).
 * vsc_correlation_forward above allocates that workspace stream-ordered
 * (hipMallocAsync / hipFreeAsync) when the type needs it.
 * Errors: as vsc_correlation_forward, and VSC_E_INVALID when the workspace is
 * NULL or smaller than vsc_correlation_workspace_size().                    */
size_t vsc_correlation_workspace_size(int N, int C, int H, int W, int in_dtype);
/* No counterpart in the reference: the Correlation op for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and errors as
 * vsc_correlation_forward; reduced-precision inputs that need channels-last
 * copies (width not a multiple of 4, or unaligned pointers) use a static
 * device scratch of VSC_CORRELATION_STATIC_SCRATCH bytes per GPU (held until
 * the process ends) instead of an allocation: operands needing more
 * (vsc_correlation_workspace_size) are refused with VSC_E_INVALID before any
 * GPU work.  Calls that overlap on one GPU (two host threads or two streams)
 * race on that scratch; serialise them.  Nothing is allocated or freed.    */
#define VSC_CORRELATION_STATIC_SCRATCH (32u << 20)
int vsc_correlation_forward_noalloc(const void* f1, const void* f2, void* out, int N, int C, int H, int W,
                                    int max_displacement, int legacy, int in_dtype, const float* scale,
                                    vsc_hip_stream stream);
int vsc_correlation_forward_ws(const void* f1, const void* f2, void* out, int N, int C, int H, int W,
                               int max_displacement, int legacy, int in_dtype, const float* scale, void* workspace,
                               size_t workspace_bytes, vsc_hip_stream stream);

/* Quantiser for the FP8 path, on the device: per-channel equalisation
 * r_c = sqrt(max|f2_c| / max|f1_c|) (SmoothQuant), f1' = f1*r_c, f2' = f2/r_c,
 * which leaves every product f1_c*f2_c unchanged, then one per-tensor scale
 * per operand (s1 = max|f1'|/448, s2 = max|f2'|/448) and round-to-nearest-even
 * E4M3.  Writes q1, q2 (1 byte per element) and scale[0] = s1*s2.
 * Allocates a small stream-ordered workspace (vsc_quantize_fp8_pair_ws takes
 * the caller's instead).                                                    */
int vsc_quantize_fp8_pair(const float* f1, const float* f2, uint8_t* q1, uint8_t* q2, float* scale,
                          int N, int C, int H, int W, vsc_hip_stream stream);
/* The quantiser with a caller-owned workspace of
 * vsc_quantize_fp8_pair_workspace_size(C) bytes (no allocation in the call). */
size_t vsc_quantize_fp8_pair_workspace_size(int C);
int vsc_quantize_fp8_pair_ws(const float* f1, const float* f2, uint8_t* q1, uint8_t* q2, float* scale, int N, int C,
                             int H, int W, void* workspace, size_t workspace_bytes, vsc_hip_stream stream);

/* FP32 -> BF16 / F16 / FP8E4M3 element conversion on the device, round to
 * nearest even (FP8 saturates at +-448).                                   */
int vsc_convert_f32(const float* in, void* out, size_t count, int out_dtype, vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * ORT custom op "Warp".
 * Replaces WarpKernel::Compute (ort_custom_ops/opticalflow/warp.h:11-36) ->
 * ComputeCPU (src/opticalflow/warp.cc:25-134) / ComputeCUDA (warp_cuda.cc:28-77).
 * in NCHW [N,C,H,W], flow [N,2,H,W] (x then y displacement), out NCHW.
 * Bilinear backward warp with zero padding; a sample whose in-bounds bilinear
 * mass is <= 0.999 is 0.  dtype: VSC_F32 (FP32 weights) or VSC_F64.        */
int vsc_warp_forward(const void* in, const void* flow, void* out, int N, int C, int H, int W, int dtype,
                     vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * Stabilization stages (src/stabilization/flowconsistency.cuh:3-50).
 * All images HWC float; sizes must match where the reference reads them
 * together.  The reference synchronises the device after each stage
 * (flowconsistency.cu:285-373); these entry points do not.                  */

/* get_warp_result (flowconsistency.cu:77-114,301-312): out = in sampled at
 * clamp(x+flow, 0, W-3/H-3), bilinear, 3 channels; flow has 2 or 3 channels. */
int vsc_stab_warp(vsc_image in, vsc_image flow, vsc_image out, vsc_hip_stream stream);

/* get_adap_comb (flowconsistency.cu:116-166,314-333). */
int vsc_stab_adap_comb(vsc_image crnt_in, vsc_image crnt_pr, vsc_image prev_warp_in, vsc_image prev_warp_pr,
                       vsc_image next_warp_in, vsc_image next_warp_pr, vsc_image adap_cmb_in,
                       vsc_image adap_cmb_pr, vsc_image last_stab_warp, float alpha, vsc_hip_stream stream);

/* get_consist_wt (flowconsistency.cu:168-191,335-348). */
int vsc_stab_consist_wt(vsc_image adap_cmb_in, vsc_image crnt_in, vsc_image consist_wt, float beta, float gamma,
                        vsc_hip_stream stream);

/* get_consist_out (flowconsistency.cu:193-258,350-374): num_iter gradient
 * sweeps with momentum; consist_out holds the initial solution on entry and
 * the result on return.  Jacobi semantics (each sweep reads only the previous
 * sweep) — the reference updates in place, which is schedule-dependent.
 * workspace: device scratch of vsc_stab_consist_out_workspace_size() bytes, or
 * NULL to let the call allocate it stream-ordered (hipMallocAsync/FreeAsync). */
size_t vsc_stab_consist_out_workspace_size(int width, int height);
int vsc_stab_consist_out(vsc_image crnt_pr, vsc_image prev_stab_warp, vsc_image consist_wt, int num_iter,
                         float step_size, float mom_fac, vsc_image consist_out, void* workspace,
                         vsc_hip_stream stream);

/* get_consist_out on planar images: each argument is a W*H*3 float array
 * holding three W x H planes (channel c at offset c*W*H), the layout the
 * per-stream driver keeps its solver state in.  Runs the temporally blocked
 * solver; same results as vsc_stab_consist_out.  workspace: device scratch of
 * vsc_stab_consist_out_planar_workspace_size() bytes, or NULL to let the call
 * allocate it stream-ordered. */
size_t vsc_stab_consist_out_planar_workspace_size(int width, int height);
int vsc_stab_consist_out_planar(const float* crnt_pr, const float* prev_stab_warp, const float* consist_wt,
                                int width, int height, int num_iter, float step_size, float mom_fac,
                                float* consist_out, void* workspace, vsc_hip_stream stream);

/* Launch plan of the temporally blocked solver (diagnostic; no counterpart in
 * the reference, which launches one kernel per sweep): for one pyramid level
 * of width x height pixels and num_iter sweeps, the most sweeps per launch
 * (halo), the tile height in rows and the tile grid (per channel); fma
 * selects the FMA-form solver (VSC_FRAME_FMA), else the reference order. */
int vsc_stab_blocked_plan(int width, int height, int num_iter, int fma, int* halo, int* tile_rows, int* tiles_x,
                          int* tiles_y);

/* Tile placement along one axis of `length` pixels (diagnostic): tiles of
 * `tile` pixels whose `halo` pixels facing another tile are not stored.
 * *count receives the number of tiles; for 0 <= i < *count, tile i computes
 * [*start, *start + tile) and stores [*lo, *hi). */
int vsc_stab_tile_span(int length, int tile, int halo, int i, int* count, int* start, int* lo, int* hi);

/* get_bilinear (flowconsistency.cu:50-75,288-299): resample in -> out. */
int vsc_stab_bilinear(vsc_image in, vsc_image out, vsc_hip_stream stream);

/* perform_consistency (flowconsistency.cu:35-48,269-286): out = 0.5*prev + 0.5*next
 * of the processed frames (the reference declares more arguments but only
 * these three are read; it has no caller). */
int vsc_average_frames(vsc_image processed_prev, vsc_image processed_next, vsc_image out, vsc_hip_stream stream);

/* char_to_cuda_float_image / kernel_to_float_image (gpuimage.cu:39-51,70-104):
 * RGBA8 (device, width*height*4 bytes) -> 3-channel float, v/255; alpha dropped. */
int vsc_rgba_to_float(const uint8_t* rgba, vsc_image out, vsc_hip_stream stream);

/* cuda_float_to_char_image / kernel_to_char_image (gpuimage.cu:54-67,107-135):
 * 3-channel float -> RGBA8, u8 = (uint8)floor(|v|*255) with uint32 saturation
 * then truncation mod 256 (the reference's wrap for v >= 256/255); alpha = 1. */
int vsc_float_to_rgba(vsc_image in, uint8_t* rgba, vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * Per-stream stabilizer: VideoStabilizer (videostabilizer.h:48-110) without
 * the Qt frame lists.  Owns the device-resident, double-buffered state (the
 * previous stabilized frame and every intermediate image); nothing is
 * allocated per frame.  One vsc_stream may be driven by one host thread at a
 * time; different streams are independent.                                 */
typedef struct vsc_stream vsc_stream;

/* flow_width/flow_height: resolution the flows arrive at (== width/height,
 * or the FLOWDOWNSCALE resolution round(W/f) x round(H/f), in which case
 * vsc_stab_frame upsamples them with get_bilinear as flowmodel.cpp:158-161). */
int vsc_stream_create(int width, int height, int flow_width, int flow_height, vsc_stream** out);
int vsc_stream_destroy(vsc_stream* s);
/* Device bytes held by a stream of this size (for admission control). */
size_t vsc_stream_footprint(int width, int height, int flow_width, int flow_height);

/* preloadProcessedFrames (videostabilizer.cpp:136-153): the first `last`
 * stabilized frame is processed frame 2k+B-1 (RGBA8 device image). */
int vsc_stream_set_last_rgba(vsc_stream* s, const uint8_t* processed_rgba, vsc_hip_stream stream);
/* Set `last` from a 3-channel float HWC device image. */
int vsc_stream_set_last(vsc_stream* s, vsc_image last, vsc_hip_stream stream);
/* Copy the current `last` stabilized frame into a 3-channel HWC float image
 * (stream-ordered). */
int vsc_stream_copy_last(vsc_stream* s, vsc_image out, vsc_hip_stream stream);

typedef struct {
    /* RGBA8 device frames (width*height*4 bytes, 4-byte aligned): the k=1 window. */
    const uint8_t* orig_prev;
    const uint8_t* orig_cur;
    const uint8_t* orig_next;
    const uint8_t* proc_prev;
    const uint8_t* proc_cur;
    const uint8_t* proc_next;
    /* HWC float device flows at flow_width x flow_height, flow_channels (2|3):
     * flow_fwd maps cur -> next, flow_bwd maps cur -> prev. */
    const float* flow_fwd;
    const float* flow_bwd;
    int flow_channels;
} vsc_frame_in;

#define VSC_FRAME_FUSED 0     /* fused front pass + temporally blocked solver */
#define VSC_FRAME_STAGED 1    /* one launch per reference stage (parity aid) */
#define VSC_FRAME_FMA 2       /* as FUSED, solver update in FMA form: per-pixel
                                 constants q = cnt + w, k = step*(lapP - w*A) and
                                 7 operations per pixel and sweep instead of 13;
                                 differs from the reference op order by rounding
                                 (<= 1 LSB in the RGBA8 output) */

/* Stabilization body of VideoStabilizer::doOneStep (videostabilizer.cpp:177-238):
 * 5 warps, adaptive combination (alpha), consistency weights (beta, gamma),
 * pyramid solve, RGBA8 output; then last <- stabilized frame.
 * out_rgba: device RGBA8 (may be NULL); out_float: optional HWC float copy.
 * params: NULL = defaults; read at call time (the interactive knob applies
 * from the next frame, as videostabilizer.cpp:192). */
int vsc_stab_frame(vsc_stream* s, const vsc_frame_in* in, const vsc_params* params, uint8_t* out_rgba,
                   float* out_float, int mode, vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * Flow network building blocks on channels-last (NHWC) device tensors.  In
 * the reference these layers are ONNX Conv nodes that ORT runs through cuDNN
 * (pwcnet.py:186-241 extractor, 511-535 DecoderLight, 625-644 Refiner); here
 * they are part of the library so the whole flow network runs on its kernels.
 * A tensor slice is (pointer, pixel stride cs, channel offset c0): channel c
 * of pixel (n, y, x) lives at p[((n*H + y)*W + x)*cs + c0 + c].             */

/* 3x3 convolution, padding `pad`, stride 1|2, dilation d (stride 2 needs d = 1):
 * out = act(sum_{ky,kx,ci} wt[co][ky][kx][ci] * in[y*s - pad + ky*d][x*s - pad + kx*d][ci]
 *           + bias[co]) (+ add), act 0 none, 1 LeakyReLU(0.1), 2 ReLU.
 * dtype VSC_F32 / VSC_BF16 / VSC_F16 for `in` and `wt` (FP32 accumulation on
 * MFMA); out_dtype = dtype or VSC_F32; `add` (nullable) has out_dtype.  Cin,
 * in_cs and in_c0 are multiples of 16 bytes' worth of elements (4 FP32 / 8
 * 16-bit); out_c0 and out_cs are multiples of 4 when Cout >= 4.  bias nullable. */
int vsc_conv3x3_nhwc(const void* in, int in_cs, int in_c0, int N, int H, int W, int Cin, const void* wt,
                     const float* bias, void* out, int out_cs, int out_c0, int Cout, int stride, int dilation,
                     int pad, int act, const void* add, int add_cs, int add_c0, int dtype, int out_dtype,
                     vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * Flow network: PWCNet-light ("pwcnet-my-light4dec", model_spec.py:32; the
 * network of pwcnet.py:787-1014 with Extractor, DecoderLight at levels 6..3
 * and Refiner) wrapped by the ONNX pre/post-processing of convert_onnx.py:
 * RGBA8 frames in, HWC flow (u, v[, 0]) out at the frame resolution.  Replaces
 * FlowModel::run's ORT session (flowmodel.cpp:121-168) for one frame pair.
 * The trained weights are not distributed with the reference; the network is
 * initialised from `seed` with the reference's initialisation (pwcnet.py:
 * 887-891) or loaded with vsc_flownet_set_params.
 *
 * Features are cached per frame in `slots`: vsc_flownet_extract runs the
 * extractor once per frame, vsc_flownet_flow runs the decoders for one or
 * two (first, second) slot pairs in one batch (a frame's forward and backward
 * flows).  downscale (FLOWDOWNSCALE, videostabilizer.cpp:74-94) resizes the
 * frames by nearest sampling to round(W/f) x round(H/f) before the network;
 * the flow is then produced at that size (flowmodel.cpp:126-133).
 * dtype: VSC_F32 (plain form) or VSC_BF16 (BF16 storage and MFMA operands,
 * FP32 accumulation).  The extractor and the decoders have scratch of their
 * own: an extraction and a decode may run at once on two streams when the
 * extraction writes none of the slots the decode reads (two extractions, or
 * two decodes, of one network may not).                                    */
typedef struct vsc_flownet vsc_flownet;

int vsc_flownet_create(int frame_width, int frame_height, int downscale, int dtype, unsigned long long seed, int slots,
                       vsc_flownet** out);
int vsc_flownet_destroy(vsc_flownet* net);
size_t vsc_flownet_footprint(int frame_width, int frame_height, int downscale, int dtype, int slots);
/* flow resolution (round(W/f), round(H/f)) */
int vsc_flownet_flow_size(const vsc_flownet* net, int* width, int* height);
/* Element type of the cost volume's products (no counterpart in the
 * reference, whose ONNX graph runs in FP32): VSC_BF16 in an FP32 network
 * rounds the features to BF16 as the cost volume stages them and multiplies
 * on the BF16 MFMA with FP32 accumulation (the north star's BF16-correlation
 * form); every other layer stays FP32.  VSC_F32 (default) keeps FP32
 * products.  A BF16 network accepts only VSC_BF16.  Errors: VSC_E_DTYPE. */
int vsc_flownet_set_corr_dtype(vsc_flownet* net, int dtype);
/* parameters as float32 in layer order (extractor 18 convs, upflow
 * ConvTranspose2d of levels 5, 4, 3, decoder convs of levels 6, 5, 4, 3 (6 each),
 * refiner 7 convs), each weight in torch layout then its bias. */
size_t vsc_flownet_param_count(const vsc_flownet* net);
int vsc_flownet_get_params(const vsc_flownet* net, float* host, size_t count);
int vsc_flownet_set_params(vsc_flownet* net, const float* host, size_t count);
/* Multiply the weights and biases of the flow-output layers (each decoder's
 * last conv and the refiner's last conv) by factor.  With seeded, untrained
 * weights the network's flows reach ~100 px; a small factor gives flows of
 * the magnitude of real motion, so that the warps' memory access pattern in a
 * timing run resembles a trained network's. */
int vsc_flownet_scale_flow_heads(vsc_flownet* net, float factor);
/* rgba: device RGBA8 frame_width x frame_height (4-byte aligned) */
int vsc_flownet_extract(vsc_flownet* net, const uint8_t* rgba, int slot, vsc_hip_stream stream);
/* npairs 1|2; out[i]: device HWC float flow_width x flow_height x out_channels (2|3) */
int vsc_flownet_flow(vsc_flownet* net, int npairs, const int* first_slots, const int* second_slots, float* const* out,
                     int out_channels, vsc_hip_stream stream);
/* Copy an FP32 flow of the last vsc_flownet_flow call (item < npairs) into
 * dst ((H/2^level) x (W/2^level) x 2 device floats of the 64-padded frame):
 * which 0 = upsampled flow entering level 5..3, 1 = decoder flow of level
 * 6..3, 2 = refined flow (level 3).  For tests and diagnostics. */
int vsc_flownet_intermediate(const vsc_flownet* net, int which, int level, int item, float* dst,
                             vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * Learned flow upsampler for FLOWDOWNSCALE (replaces the get_bilinear flow
 * upsampling of flowmodel.cpp:158-161 when enabled): input = bilinear x scale
 * of the low-resolution flow (the reference's kernel_bilinear mapping) plus
 * the luma of the full-resolution frame; conv3x3 3->16, ReLU, 16->16, ReLU,
 * 16->2, added to the bilinear flow.  Inference: FP16 storage, FP32
 * accumulation.  Training (FP32 master weights, forward, backward and Adam on
 * the device): Charbonnier end-point loss sqrt(|out - gt|^2 + 1e-6), mean
 * over pixels.  Parameters in torch layout: w1 [16][3][3][3], b1 [16],
 * w2 [16][16][3][3], b2 [16], w3 [2][16][3][3], b3 [2].                     */
typedef struct vsc_upsampler vsc_upsampler;

int vsc_upsampler_create(unsigned long long seed, vsc_upsampler** out);
int vsc_upsampler_destroy(vsc_upsampler* ups);
size_t vsc_upsampler_param_count(void);
int vsc_upsampler_get_params(const vsc_upsampler* ups, float* host);
int vsc_upsampler_set_params(vsc_upsampler* ups, const float* host);
/* device bytes of the inference workspace for a width x height output: 0 (the
 * fused inference kernel keeps its activations in LDS); kept so callers can
 * size a workspace without knowing the kernel. */
size_t vsc_upsampler_workspace_size(int width, int height);
/* flow_lo: device HWC lo_w x lo_h x lo_ch (2|3); rgba: device RGBA8 width x
 * height (4-byte aligned, else VSC_E_INVALID); flow_out: device HWC width x height x out_ch (2|3, third channel 0);
 * scale: factor applied to the bilinear flow vectors (FLOWDOWNSCALE f with
 * vector scaling, 1 for the reference's unscaled vectors); gate: NULL or a
 * device float multiplying the learned residual (see vsc_flow_fb_gate);
 * workspace: vsc_upsampler_workspace_size bytes (may be NULL when that is 0).
 * One fused kernel: the three layers of a 20 x 28 tile in LDS.             */
int vsc_upsampler_run(vsc_upsampler* ups, const float* flow_lo, int lo_w, int lo_h, int lo_ch, const uint8_t* rgba,
                      int width, int height, float scale, float* flow_out, int out_ch, const float* gate,
                      void* workspace, vsc_hip_stream stream);
/* Training on a batch: flow_lo [batch][lo_h][lo_w][2], rgba [batch][height][width][4],
 * flow_gt [batch][height][width][2] (device).  loss (nullable host pointer)
 * receives the batch loss (synchronises).  train_step applies one Adam step
 * (beta 0.9 / 0.999, eps 1e-8); loss_grad only computes the gradient, copied
 * to grad_host (nullable) in the parameter layout above. */
int vsc_upsampler_train_step(vsc_upsampler* ups, int batch, const float* flow_lo, int lo_w, int lo_h,
                             const uint8_t* rgba, const float* flow_gt, int width, int height, float scale, float lr,
                             float* loss, vsc_hip_stream stream);
int vsc_upsampler_loss_grad(vsc_upsampler* ups, int batch, const float* flow_lo, int lo_w, int lo_h,
                            const uint8_t* rgba, const float* flow_gt, int width, int height, float scale, float* loss,
                            float* grad_host, vsc_hip_stream stream);

/* Forward-backward flow consistency (Sundaram, Brox & Keutzer 2010) as the
 * per-frame trigger of the bilinear fallback: fwd maps frame A to B, bwd maps
 * B to A (device HWC, channels 2|3).  A pixel is inconsistent when x + fwd(x)
 * leaves the frame or |fwd(x) + bwd(x + fwd(x))|^2 >= 0.01 (|fwd|^2 + |bwd|^2)
 * + 0.5 (bwd sampled bilinearly).  Writes the count to *count and to *gate 1.0
 * when count <= max_inconsistent * width * height, else 0.0 (both device
 * scalars; no host synchronisation). */
int vsc_flow_fb_gate(const float* fwd, const float* bwd, int width, int height, int channels, float max_inconsistent,
                     uint32_t* count, float* gate, vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * Per-stream frame pipeline: VideoStabilizer::doOneStep (videostabilizer.cpp:
 * 167-265) including retrieveOpticalFlow (:267-279) through the flow network,
 * FLOWDOWNSCALE (videostabilizer.cpp:74-94, flowmodel.cpp:126-133,158-161)
 * and the learned upsampler.  Owns a vsc_flownet (flow_batch + 2 feature
 * slots; 2 flow_batch + 2 with overlap), the flow buffers, an upsampler and
 * a vsc_stream; a frame allocates nothing.                                   */
typedef struct vsc_pipeline vsc_pipeline;

typedef struct {
    int flow_dtype;             /* VSC_F32 (plain form) | VSC_BF16 */
    int downscale;              /* FLOWDOWNSCALE f: 1 | 2 | 4 */
    int upsample;               /* f > 1: 0 bilinear (get_bilinear), 1 learned upsampler */
    int scale_vectors;          /* bilinear path: multiply vectors by f (the reference does not: 0);
                                   the learned path always scales (trained on full-resolution vectors) */
    float fb_max_inconsistent;  /* learned path: bilinear fallback when the inconsistent fraction exceeds this (< 0: never) */
    int ref_bwd_index;          /* 1: backward flow = net(next -> cur) as videostabilizer.cpp:271-272; 0: cur -> prev */
    unsigned long long seed;    /* flow network (and untrained upsampler) initialisation */
    int solver;                 /* stabilization solver: 0 reference op order (VSC_FRAME_FUSED),
                                   1 FMA form (VSC_FRAME_FMA, <= 1 LSB from it) */
    int overlap;                /* 1: frame-pipelined execution on three internal HIP streams: the flows
                                   of a frame run while the previous frame is stabilized, and the new
                                   frames' features of a call while the previous call decodes.  Frame t's
                                   outputs are then ordered on the caller's stream by the call for
                                   frame t+1 (or vsc_pipeline_sync), and the window frames of call t
                                   must stay unchanged until then.  Same results as 0. */
    int flow_batch;             /* B, 1..4: the most frames whose flows vsc_pipeline_frames computes in
                                   one batched network decode (the reference's batchSize: FlowModel::run
                                   over B frame pairs every B frames, videostabilizer.cpp:65-71,177,
                                   267-279).  Each flow has the same bits in any batch. */
    int corr_dtype;             /* cost volume products: VSC_F32 (default) or VSC_BF16 (in an FP32 network:
                                   the BF16-correlation form, vsc_flownet_set_corr_dtype) */
} vsc_pipeline_cfg;

/* The k = 1 frame window (device RGBA8 frames, 4-byte aligned). */
typedef struct {
    const uint8_t* orig_prev;
    const uint8_t* orig_cur;
    const uint8_t* orig_next;
    const uint8_t* proc_prev;
    const uint8_t* proc_cur;
    const uint8_t* proc_next;
} vsc_window;

void vsc_pipeline_default_cfg(vsc_pipeline_cfg* cfg);
int vsc_pipeline_create(int width, int height, const vsc_pipeline_cfg* cfg, vsc_pipeline** out);
int vsc_pipeline_destroy(vsc_pipeline* p);
size_t vsc_pipeline_footprint(int width, int height, const vsc_pipeline_cfg* cfg);
int vsc_pipeline_set_upsampler_params(vsc_pipeline* p, const float* host);
/* borrowed handles of the parts (for `last` preloading and diagnostics) */
int vsc_pipeline_flownet(vsc_pipeline* p, vsc_flownet** net);
int vsc_pipeline_stream(vsc_pipeline* p, vsc_stream** s);
/* Flows of window `frame` (= index of orig_cur; frames extracted once, keyed
 * by index) at the flow resolution, HWC 2 channels (device). */
int vsc_pipeline_flows(vsc_pipeline* p, const vsc_window* w, long long frame, float* fwd, float* bwd,
                       vsc_hip_stream stream);
/* One stabilized frame of window `frame` (flows, upsampling, vsc_stab_frame). */
int vsc_pipeline_frame(vsc_pipeline* p, const vsc_window* w, long long frame, const vsc_params* params,
                       uint8_t* out_rgba, float* out_float, vsc_hip_stream stream);
/* n (1 .. flow_batch) consecutive stabilized frames: w[i] is the window of
 * frame first + i (the reference's doOneStep calls for currentFrame .. +B-1,
 * whose frame window holds 2k+B frames).  Consecutive windows share frames;
 * the network reads each original frame once, keyed by its index (prev of
 * w[0], cur of every window, next of w[n-1]), like vsc_pipeline_flows.  The flows of all n frames come from
 * one batched decode of 2n pairs, then the frames are stabilized in order;
 * the results equal n vsc_pipeline_frame calls.  out_rgba / out_float: n
 * pointers each (the arrays, or their entries, may be NULL).  With overlap =
 * 1 the outputs of the batch are ordered on `stream` by the next call (or
 * vsc_pipeline_sync).  Errors: VSC_E_INVALID for n outside [1, flow_batch]
 * or a NULL window, before any work. */
int vsc_pipeline_frames(vsc_pipeline* p, const vsc_window* w, int n, long long first, const vsc_params* params,
                        uint8_t* const* out_rgba, float* const* out_float, vsc_hip_stream stream);
/* The last frame's fallback gate (1 learned, 0 bilinear) and inconsistent count. */
int vsc_pipeline_gate(const vsc_pipeline* p, float* gate, uint32_t* count, vsc_hip_stream stream);
/* Order `stream` after every frame submitted so far (needed only with overlap = 1). */
int vsc_pipeline_sync(vsc_pipeline* p, vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * Multi-stream frame server: many concurrent video streams of different
 * resolutions on one GPU (the reference serves one stream per process:
 * VideoPlayer / stabilizefiles, videoplayer.cpp:185-190).  Streams are admitted
 * by device-memory footprint against hbm_budget; each owns a vsc_pipeline
 * (form: the pipeline configuration, by default the chosen form: BF16 flow
 * network at FLOWDOWNSCALE 2 with the learned upsampler), a 3-slot ring of
 * input frames and a 6-slot output ring.  vsc_server_submit copies frame t of
 * a stream into its ring (host or device memory) on the stream's HIP stream;
 * from frame 2 on, each submission makes the window centred at t-1 ready.
 * vsc_server_tick launches every ready window of every stream: the streams'
 * flow work spread over hip_streams HIP streams (default priority) so their
 * kernels overlap, and (batch = 1) the stabilization of all of the tick's
 * windows as one graph of persistent launches on the server's tick stream;
 * streams opened or frames submitted while earlier work runs join the next
 * tick.  vsc_server_poll reports finished frames in order per stream; a
 * reported frame's output stays in the stream's 6-slot output ring until the
 * next tick (read it with vsc_server_read before ticking again).  A window is
 * launched only when its output slot has been reported, and frame t is taken
 * only when window t-2 has been launched; otherwise submit returns
 * VSC_E_AGAIN.  Parameter changes (vsc_server_set_params) apply from the next
 * tick, as the GUI's sliders do (videostabilizer.cpp:192).                   */
typedef struct vsc_server vsc_server;

typedef struct {
    int max_streams;
    size_t hbm_budget;        /* bytes of device memory the server may commit */
    int hip_streams;          /* HIP streams in the pool (1..16) */
    vsc_pipeline_cfg form;
    int use_graphs;           /* 1: replay each stream's per-frame work from hipGraphs (6 per stream, one
                                 per frame phase, captured once, re-captured after a parameter change) */
    int batch;                /* 1 (default): a tick stabilizes the windows of all its streams together:
                                 each stream's flow work on its pool stream, then one graph on the
                                 server's tick stream (work-list copy + persistent launches that walk
                                 the frames of every stream); windows the batch cannot take (more than
                                 2 pyramid levels, odd sizes at 2 levels, num_iter > 304) run per
                                 stream.  0: every frame runs per stream on its pool stream. */
} vsc_server_cfg;

typedef struct {
    int stream;
    long long tag;            /* the tag given with the submission that completed the window */
    long long frame;          /* stabilized frame index (window centre) */
    const uint8_t* out_rgba;  /* device RGBA8 W x H */
} vsc_server_done;

void vsc_server_default_cfg(vsc_server_cfg* cfg);
/* upsampler_params: the learned upsampler's weights (vsc_upsampler layout),
 * required when the form uses it. */
int vsc_server_create(const vsc_server_cfg* cfg, const float* upsampler_params, vsc_server** out);
int vsc_server_destroy(vsc_server* sv);
size_t vsc_server_stream_footprint(const vsc_server* sv, int width, int height);
size_t vsc_server_used_bytes(const vsc_server* sv);
/* VSC_E_NOMEM when the budget (or the stream table) cannot admit the stream now */
int vsc_server_open(vsc_server* sv, int width, int height, int* stream_id);
int vsc_server_close(vsc_server* sv, int stream_id);
int vsc_server_set_params(vsc_server* sv, int stream_id, const vsc_params* params);
int vsc_server_submit(vsc_server* sv, int stream_id, const uint8_t* orig_rgba, const uint8_t* proc_rgba, long long tag);
int vsc_server_tick(vsc_server* sv, int* launched);
int vsc_server_poll(vsc_server* sv, vsc_server_done* done, int max, int* count);
/* wait for all launched work */
int vsc_server_drain(vsc_server* sv);
/* Copy stabilized frame `frame` (still in the output ring) of a stream into
 * dst (device or pinned host RGBA8), ordered after the stream's work and
 * before later work on `stream`. */
int vsc_server_read(vsc_server* sv, int stream_id, long long frame, uint8_t* dst, vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * Flow files (ReadFlowFile, flowIO.cpp:31-78): the Middlebury .flo format,
 * host memory.  vsc_flo_read with flow == NULL returns the size only; errors
 * (VSC_E_INVALID, message in vsc_last_error) for a wrong tag, a width or height
 * outside [1, 99999], a short file or trailing bytes, as the reference rejects
 * them.  Flows are HWC (u, v) float32, little-endian.                        */
int vsc_flo_read(const char* path, float* flow, size_t capacity_floats, int* width, int* height);
int vsc_flo_write(const char* path, const float* flow, int width, int height);
/* Flow-file mode of the file stabilizer (`-f <directory>`, main.cpp:60-63):
 * FileStabilizer::retrieveOpticalFlow (stabilizefiles.cpp:134-149) reads, for
 * window `frame` = i, the forward flow i -> i+1 from frame_<i+1>.flo and the
 * backward flow i -> i-1 from frame_<i>_bwd.flo (indices zero-padded to 6
 * digits, formatIndex, stabilizefiles.cpp:99-101).
 * vsc_flow_dir_path: that file's path (backward 0: forward flow).
 * vsc_flow_dir_read: both flows (host, HWC (u, v) float32; either pointer may
 * be NULL, both NULL = sizes only); VSC_E_INVALID when a file is missing or
 * malformed, or the two differ in size.                                     */
int vsc_flow_dir_path(const char* dir, long long frame, int backward, char* path, size_t capacity);
int vsc_flow_dir_read(const char* dir, long long frame, float* fwd, float* bwd, size_t capacity_floats, int* width,
                      int* height);
/* Receives each output frame in order (host RGBA8, valid during the call);
 * a nonzero return stops the loop. */
typedef int (*vsc_frame_sink)(void* user, long long index, const uint8_t* rgba, int width, int height);
/* The -f frame loop: FileStabilizer::stabilizeAll (stabilizefiles.cpp:108-132)
 * with preloadProcessedFrames / doOneStep / outputFinalFrames
 * (videostabilizer.cpp:136-238), k = 1, batch 1.  orig_rgba / proc_rgba:
 * nframes (>= 3) host RGBA8 frames of width x height (decoded by the caller).
 * Frame 0 and frame nframes-1 leave as the processed frames, converted to
 * float and back as the reference's gpuToImage does (alpha 1); frame i in
 * 1..nframes-2 is stabilized with the flows of vsc_flow_dir_read(flow_dir, i),
 * the first `last` being processed frame 2.  A flow whose size differs from
 * the frame is refused (VSC_E_INVALID), as initializeFlowImage throws
 * (imagehelpers.cpp:42-50).  params: NULL = defaults; mode VSC_FRAME_FUSED
 * or VSC_FRAME_FMA.  Allocates its device buffers once per call and
 * synchronises `stream` once per frame (before the sink sees the frame).  */
int vsc_stabilize_flow_dir(const uint8_t* const* orig_rgba, const uint8_t* const* proc_rgba, int nframes, int width,
                           int height, const char* flow_dir, const vsc_params* params, int mode, vsc_frame_sink sink,
                           void* user, vsc_hip_stream stream);

/* ---------------------------------------------------------------------------
 * Per-kernel timing (replaces the per-stage QElapsedTimer accounting of
 * VideoStabilizer, videostabilizer.h:72-76 / videostabilizer.cpp:175,240-262,
 * with device-side timing).  Between begin and end, every launch of the
 * instrumented kernels ("consist_blocked": the temporally blocked solver,
 * "front_pass": the fused front pass) is bracketed by HIP events recorded on
 * the launch stream, up to `capacity` launches.  vsc_profile_end synchronises
 * those events and returns, for one kernel name, the number of launches, their
 * summed duration and their summed algorithmic bytes (tensors crossing the
 * launch boundary, each counted once).  Host-thread-global; not reentrant.  */
int vsc_profile_begin(int capacity);
int vsc_profile_end(const char* kernel, int* launches, double* total_ms, double* algorithmic_bytes);

#ifdef __cplusplus
}
#endif

#endif /* VSC_H */
