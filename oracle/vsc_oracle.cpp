/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of the per-frame consistency hot path of
 * MaxReimann/video-stream-consistency, used as the correctness checker for the
 * HIP implementation.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load this library; the product never links or calls it.
 *
 * Two precisions of every function:
 *   *_f32 : "reference order" — the float arithmetic of the reference kernels,
 *           operation by operation, compiled with -ffp-contract=off so that no
 *           a*b+c is fused (the GPU build uses the same contraction policy).
 *   *_f64 : an FP64 twin (same formulas in double, libm exp) used as the error
 *           budget ("FP64 oracle") for PSNR / relative-error reporting.
 *
 * The solver also has an FMA form (stab_consist_out_fma, *_fma_f32): the same
 * update rearranged into fused multiply-adds, the arithmetic of the library's
 * VSC_FRAME_FMA mode; it is pinned against the reference-order solver and the
 * FP64 twin (tests/test_oracle_stab.py).
 *
 * Semantics that differ from the reference on purpose (documented deviations):
 *   - the consistency solver (kernel_consist_out) is evaluated with Jacobi
 *     (double-buffered) semantics: every sweep reads only the previous sweep.
 *     The reference launches it in place (flowconsistency.cu:370), which makes
 *     its GPU result schedule-dependent; Jacobi is the only well-defined order.
 *   - exp is vsco_expf below (Cody–Waite reduction + degree-7 Taylor
 *     polynomial), not CUDA's expf, so that weight clamps agree bit for bit
 *     between this oracle and the GPU build.
 *   - legacy correlation uses zero padding (the reference leaves its padded
 *     buffers uninitialised, correlation_cuda.cu:367-370; the cupy original
 *     zero-fills, model-conversion/pwc_pytorch/correlation/correlation.py:281).
 *
 * Parity pins (tests/test_oracle_*.py): FP64 brute force and closed forms for
 * correlation; identity/shift/half-pixel/grid_sample cross-check for warp;
 * 1x1 heavy-ball recurrence, gamma=0 identity and the sparse direct solve of the
 * fixed point for the solver; exact decimation / midpoint / linear-ramp for the
 * resampler; exhaustive u8 conversions; libm error bound for vsco_expf.
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

#define VSCO_API extern "C" __attribute__((visibility("default")))

/* Host threads used by the OpenMP loops (1 = the reference CPU ops' own
 * single-threaded execution). Returns the count now in effect. */
VSCO_API int vsco_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* exp                                                                       */
/* ------------------------------------------------------------------------- */

static inline float bits_to_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

/* exp(x) in binary32 by Cody & Waite (1980) argument reduction
 *   x = k*ln2 + r,  k = rint(x*log2(e)),  r = (x - k*ln2_hi) - k*ln2_lo
 * (each step one fused multiply-add; k*ln2_hi is exact) and a degree-7 Taylor
 * polynomial for e^r on |r| <= ~0.35 (truncation error r^8/8! < 6e-9),
 * evaluated by Horner with fused multiply-adds; the result is scaled by 2^k
 * through the exponent field.  Every step is a single correctly rounded IEEE
 * binary32 operation (std::fma; v_fma_f32 on the GPU), so the value is the same
 * on any IEEE machine. */
VSCO_API float vsco_expf(float x)
{
    if (x != x) return x;
    if (x > 88.72283935546875f) return INFINITY;
    if (x < -103.97208404541015625f) return 0.0f;
    const float log2e = 1.44269502162933349609375f;
    const float ln2_hi = 0.693145751953125f;          /* 0x3f317200, 15 significant bits */
    const float ln2_lo = 1.428606765330187045e-06f;   /* ln2 - ln2_hi */
    float kf = rintf(x * log2e);
    int k = (int)kf;
    float r = std::fma(-kf, ln2_hi, x);
    r = std::fma(-kf, ln2_lo, r);
    float p = 1.0f / 5040.0f;
    p = std::fma(p, r, 1.0f / 720.0f);
    p = std::fma(p, r, 1.0f / 120.0f);
    p = std::fma(p, r, 1.0f / 24.0f);
    p = std::fma(p, r, 1.0f / 6.0f);
    p = std::fma(p, r, 0.5f);
    p = std::fma(p, r, 1.0f);
    p = std::fma(p, r, 1.0f);
    if (k > 127) { p = p * 2.0f; k -= 1; }
    if (k < -126) { p = p * 5.42101086242752217003726400434970855712890625e-20f; k += 64; }
    return p * bits_to_float((uint32_t)(k + 127) << 23);
}

template <typename T> static inline T oexp(T x);
template <> inline float oexp<float>(float x) { return vsco_expf(x); }
template <> inline double oexp<double>(double x) { return std::exp(x); }

/* ------------------------------------------------------------------------- */
/* Correlation (ORT custom op "Correlation")                                 */
/* ------------------------------------------------------------------------- */

/* Non-legacy correlation, reference CPU op:
 *   correlation.cc:45-143 (ComputeCPU: k=1, stride 1, pad 0, patch 2md+1),
 *   correlation.cc:147-182 (correlate_patch: *dst = 0; *dst += v1*v2 over c),
 *   correlation.cc:203-275 (loop nest n, ph, pw, h, w).
 * out[n][ph][pw][h][w] = sum_c f1[n][c][h][w] * f2[n][c][h+ph-md][w+pw-md],
 * terms with the shifted position outside [0,H)x[0,W) are skipped (0). The sum
 * runs sequentially in c and accumulates directly in T. */
template <typename T>
static void correlation_ref(const T* f1, const T* f2, T* out, int N, int C, int H, int W, int md)
{
    const int P = 2 * md + 1;
    const size_t HW = (size_t)H * W;
#pragma omp parallel for collapse(3) schedule(static)
    for (int n = 0; n < N; ++n)
        for (int ph = 0; ph < P; ++ph)
            for (int pw = 0; pw < P; ++pw) {
                const T* a = f1 + (size_t)n * C * HW;
                const T* b = f2 + (size_t)n * C * HW;
                T* o = out + (((size_t)n * P + ph) * P + pw) * HW;
                const int dy = ph - md, dx = pw - md;
                for (int h = 0; h < H; ++h)
                    for (int w = 0; w < W; ++w) {
                        T acc = 0;
                        const int h2 = h + dy, w2 = w + dx;
                        if (h2 >= 0 && h2 < H && w2 >= 0 && w2 < W) {
                            for (int c = 0; c < C; ++c) {
                                T v1 = a[c * HW + (size_t)h * W + w];
                                T v2 = b[c * HW + (size_t)h2 * W + w2];
                                acc += v1 * v2;
                            }
                        }
                        o[(size_t)h * W + w] = acc;
                    }
            }
}

VSCO_API void vsco_correlation_f32(const float* f1, const float* f2, float* out, int N, int C, int H, int W, int md)
{ correlation_ref<float>(f1, f2, out, N, C, H, W, md); }
VSCO_API void vsco_correlation_f64(const double* f1, const double* f2, double* out, int N, int C, int H, int W, int md)
{ correlation_ref<double>(f1, f2, out, N, C, H, W, md); }

/* Legacy correlation (original PWC-Net kernel), reference CUDA op:
 *   correlation_cuda.cu:182-265 (correlation_old_kernel), :267-331 (launch),
 *   correlation_cuda.cu:359-360 (pad hard-coded to 4, so md <= 4).
 * top[n][tc][h][w] = (sum_c f1*f2(shifted)) / C with tc = (dy+md)*(2md+1)+(dx+md)
 * and zero padding.  Summation order of the reference: 32 lane partial sums
 * over c = t, t+32, ... (kernel :70-78), then lane 0 adds the 32 partials in
 * lane order (:84-88), then one division by k*k*C (:89-91). */
template <typename T>
static void correlation_legacy_ref(const T* f1, const T* f2, T* out, int N, int C, int H, int W, int md)
{
    const int P = 2 * md + 1;
    const size_t HW = (size_t)H * W;
#pragma omp parallel for collapse(2) schedule(static)
    for (int n = 0; n < N; ++n)
        for (int tc = 0; tc < P * P; ++tc) {
            const int dx = tc % P - md, dy = tc / P - md;
            const T* a = f1 + (size_t)n * C * HW;
            const T* b = f2 + (size_t)n * C * HW;
            T* o = out + ((size_t)n * P * P + tc) * HW;
            for (int h = 0; h < H; ++h)
                for (int w = 0; w < W; ++w) {
                    T part[32];
                    for (int t = 0; t < 32; ++t) part[t] = 0;
                    const int h2 = h + dy, w2 = w + dx;
                    const bool in = (h2 >= 0 && h2 < H && w2 >= 0 && w2 < W);
                    for (int t = 0; t < 32; ++t)
                        for (int c = t; c < C; c += 32) {
                            T v2 = in ? b[c * HW + (size_t)h2 * W + w2] : (T)0;
                            part[t] += a[c * HW + (size_t)h * W + w] * v2;
                        }
                    T total = 0;
                    for (int t = 0; t < 32; ++t) total += part[t];
                    o[(size_t)h * W + w] = total / (T)C;
                }
        }
}

VSCO_API void vsco_correlation_legacy_f32(const float* f1, const float* f2, float* out, int N, int C, int H, int W, int md)
{ correlation_legacy_ref<float>(f1, f2, out, N, C, H, W, md); }
VSCO_API void vsco_correlation_legacy_f64(const double* f1, const double* f2, double* out, int N, int C, int H, int W, int md)
{ correlation_legacy_ref<double>(f1, f2, out, N, C, H, W, md); }

/* ------------------------------------------------------------------------- */
/* Warp (ORT custom op "Warp")                                               */
/* ------------------------------------------------------------------------- */

/* Reference CPU op warp.cc:71-134 (twin of warp_cuda.cu:28-84).  For T=float the
 * reference mixes precisions through its double literals: the bilinear weights
 * (1.0-alpha), (1.0-beta) are double, alpha*beta is a float product, and the
 * float accumulators val/mask are updated as float = (double)acc + double.
 * This restatement keeps exactly that evaluation. */
static void warp_ref_f32(const float* input, const float* flow, float* dst, int N, int C, int H, int W)
{
    const size_t HW = (size_t)H * W;
#pragma omp parallel for collapse(2) schedule(static)
    for (int n = 0; n < N; ++n)
        for (int c = 0; c < C; ++c)
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    const float* in = input + ((size_t)n * C + c) * HW;
                    const float fx = flow[(size_t)n * 2 * HW + (size_t)y * W + x];
                    const float fy = flow[(size_t)n * 2 * HW + HW + (size_t)y * W + x];
                    const float xf = (float)x + fx;
                    const float yf = (float)y + fy;
                    const float alpha = xf - std::floor(xf);
                    const float beta = yf - std::floor(yf);
                    const float right_edge = (float)(W - 1);
                    const float bottom_edge = (float)(H - 1);
                    const float xL = std::floor(xf);
                    const float xR = (float)((double)std::floor(xf) + 1.0);
                    const float yT = std::floor(yf);
                    const float yB = (float)((double)std::floor(yf) + 1.0);
                    const int mL = (0 <= xL && xL <= right_edge) ? 1 : 0;
                    const int mR = (0 <= xR && xR <= right_edge) ? 1 : 0;
                    const int mT = (0 <= yT && yT <= bottom_edge) ? 1 : 0;
                    const int mB = (0 <= yB && yB <= bottom_edge) ? 1 : 0;
                    float val = 0.0f, mask = 0.0f;
                    const double wTL = (1.0 - (double)alpha) * (1.0 - (double)beta);
                    const double wTR = (double)alpha * (1.0 - (double)beta);
                    const double wBL = (1.0 - (double)alpha) * (double)beta;
                    const double wBR = (double)(alpha * beta);  /* float product */
                    mask = (float)((double)mask + wTL * (mT + mL == 2 ? 1.0 : 0.0));
                    mask = (float)((double)mask + wTR * (mT + mR == 2 ? 1.0 : 0.0));
                    mask = (float)((double)mask + wBL * (mB + mL == 2 ? 1.0 : 0.0));
                    mask = (float)((double)mask + wBR * (mB + mR == 2 ? 1.0 : 0.0));
                    if ((double)mask > 0.999) {
                        const int iL = (int)xL, iR = (int)xR, iT = (int)yT, iB = (int)yB;
                        val = (float)((double)val + wTL * (mT + mL == 2 ? (double)in[(size_t)iT * W + iL] : 0.0));
                        val = (float)((double)val + wTR * (mT + mR == 2 ? (double)in[(size_t)iT * W + iR] : 0.0));
                        val = (float)((double)val + wBL * (mB + mL == 2 ? (double)in[(size_t)iB * W + iL] : 0.0));
                        val = (float)((double)val + wBR * (mB + mR == 2 ? (double)in[(size_t)iB * W + iR] : 0.0));
                    }
                    dst[((size_t)n * C + c) * HW + (size_t)y * W + x] = val;
                }
}

static void warp_ref_f64(const double* input, const double* flow, double* dst, int N, int C, int H, int W)
{
    const size_t HW = (size_t)H * W;
#pragma omp parallel for collapse(2) schedule(static)
    for (int n = 0; n < N; ++n)
        for (int c = 0; c < C; ++c)
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    const double* in = input + ((size_t)n * C + c) * HW;
                    const double xf = (double)x + flow[(size_t)n * 2 * HW + (size_t)y * W + x];
                    const double yf = (double)y + flow[(size_t)n * 2 * HW + HW + (size_t)y * W + x];
                    const double alpha = xf - std::floor(xf), beta = yf - std::floor(yf);
                    const double xL = std::floor(xf), xR = xL + 1.0, yT = std::floor(yf), yB = yT + 1.0;
                    const int mL = (0 <= xL && xL <= W - 1), mR = (0 <= xR && xR <= W - 1);
                    const int mT = (0 <= yT && yT <= H - 1), mB = (0 <= yB && yB <= H - 1);
                    const double wTL = (1 - alpha) * (1 - beta), wTR = alpha * (1 - beta);
                    const double wBL = (1 - alpha) * beta, wBR = alpha * beta;
                    double mask = 0, val = 0;
                    mask += wTL * (mT && mL) + wTR * (mT && mR) + wBL * (mB && mL) + wBR * (mB && mR);
                    if (mask > 0.999) {
                        if (mT && mL) val += wTL * in[(size_t)yT * W + (size_t)xL];
                        if (mT && mR) val += wTR * in[(size_t)yT * W + (size_t)xR];
                        if (mB && mL) val += wBL * in[(size_t)yB * W + (size_t)xL];
                        if (mB && mR) val += wBR * in[(size_t)yB * W + (size_t)xR];
                    }
                    dst[((size_t)n * C + c) * HW + (size_t)y * W + x] = val;
                }
}

VSCO_API void vsco_warp_f32(const float* in, const float* flow, float* out, int N, int C, int H, int W)
{ warp_ref_f32(in, flow, out, N, C, H, W); }
VSCO_API void vsco_warp_f64(const double* in, const double* flow, double* out, int N, int C, int H, int W)
{ warp_ref_f64(in, flow, out, N, C, H, W); }

/* ------------------------------------------------------------------------- */
/* Stabilization stages (src/stabilization/flowconsistency.cu)               */
/* Images are HWC interleaved: element (x,y,c) at ((y*W+x)*ch + c)            */
/* (gpuimagewrapper.h:18-24).                                                */
/* ------------------------------------------------------------------------- */

/* kernel_warp, flowconsistency.cu:77-114: clamp the sample position to
 * [0, W-3] x [0, H-3], bilinear, no mask, 3 channels; flow stride = flow_ch. */
template <typename T>
static void stab_warp(const T* in, const T* flow, int flow_ch, T* out, int W, int H)
{
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < H; ++iy)
        for (int ix = 0; ix < W; ++ix) {
            const T flo_x = flow[((size_t)iy * W + ix) * flow_ch + 0];
            const T flo_y = flow[((size_t)iy * W + ix) * flow_ch + 1];
            const T map_fx = std::max((T)0, std::min((T)ix + flo_x, (T)(W - 3)));
            const T map_fy = std::max((T)0, std::min((T)iy + flo_y, (T)(H - 3)));
            const int map_ix = (int)std::floor(map_fx);
            const int map_iy = (int)std::floor(map_fy);
            const T fx = map_fx - (T)map_ix;
            const T fy = map_fy - (T)map_iy;
            for (int c = 0; c < 3; ++c) {
                const T i00 = in[((size_t)map_iy * W + map_ix) * 3 + c];
                const T i10 = in[((size_t)map_iy * W + map_ix + 1) * 3 + c];
                const T i01 = in[((size_t)(map_iy + 1) * W + map_ix) * 3 + c];
                const T i11 = in[((size_t)(map_iy + 1) * W + map_ix + 1) * 3 + c];
                const T t1 = i00 * ((T)1 - fx) + i10 * fx;
                const T t2 = i01 * ((T)1 - fx) + i11 * fx;
                out[((size_t)iy * W + ix) * 3 + c] = t1 * ((T)1 - fy) + t2 * fy;
            }
        }
}

/* kernel_adap_comb, flowconsistency.cu:116-166. */
template <typename T>
static void stab_adap_comb(const T* cIn, const T* cPr, const T* pIn, const T* pPr, const T* nIn, const T* nPr,
                           T* adIn, T* adPr, const T* last, int W, int H, T alpha)
{
    const size_t n = (size_t)W * H * 3;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        const T ci = cIn[i], pi = pIn[i], ni = nIn[i];
        const T cp = cPr[i], pp = pPr[i], np = nPr[i];
        const T ls = last[i];
        T wp = oexp<T>(((-alpha) * (ci - pi)) * (ci - pi));
        T wn = oexp<T>(((-alpha) * (ci - ni)) * (ci - ni));
        if (wp > (T)0.45f) wp = (T)0.45f;
        if (wn > (T)0.3f) wn = (T)0.3f;
        if (wp < (T)0.001f) wp = 0;
        if (wn < (T)0.001f) wn = 0;
        adIn[i] = (wp * pi + wn * ni) + ((T)1 - (wp + wn)) * ci;
        T ap = (wp * pp + wn * np) + ((T)1 - (wp + wn)) * cp;
        adPr[i] = wp * ls + ((T)1 - wp) * ap;
    }
}

/* kernel_consist_wt, flowconsistency.cu:168-191. */
template <typename T>
static void stab_consist_wt(const T* adIn, const T* cIn, T* wt, int W, int H, T beta, T gamma)
{
    const size_t n = (size_t)W * H * 3;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        const T d = cIn[i] - adIn[i];
        T w = gamma * oexp<T>(((-beta) * d) * d);
        if (w < (T)0.001f) w = 0;
        wt[i] = w;
    }
}

/* kernel_consist_out, flowconsistency.cu:193-258, driven by get_consist_out
 * :350-374, with Jacobi semantics (see file header).  Sweep k for each pixel
 * and channel:
 *   neighbours, added in this order to a sum that starts at 0:
 *     right if x+1 < W-1, left if x-1 >= 0, down if y+1 < H-1, up if y-1 >= 0
 *   lapP = sumP - cnt*P ; lapO = sumO - cnt*O
 *   g = (lapO - lapP) - w*(O - A)
 *   O' = O + s*g (+ m*U for k > 0) ;  U' = s*g
 * O holds the initial solution on entry and the result on exit. */
template <typename T>
static void stab_consist_out(const T* P, const T* A, const T* Wt, int iters, T step, T mom, T* O, int W, int H)
{
    const size_t n = (size_t)W * H * 3;
    std::vector<T> Ob(n), U(n, (T)0), Un(n);
    T* cur = O;
    T* nxt = Ob.data();
    for (int k = 0; k < iters; ++k) {
#pragma omp parallel for schedule(static)
        for (int iy = 0; iy < H; ++iy)
            for (int ix = 0; ix < W; ++ix)
                for (int c = 0; c < 3; ++c) {
                    const size_t i = ((size_t)iy * W + ix) * 3 + c;
                    int cnt = 0;
                    T lp = 0, lo = 0;
                    if ((ix + 1) < (W - 1)) { lp += P[i + 3]; lo += cur[i + 3]; cnt += 1; }
                    if ((ix - 1) >= 0) { lp += P[i - 3]; lo += cur[i - 3]; cnt += 1; }
                    if ((iy + 1) < (H - 1)) { lp += P[i + (size_t)W * 3]; lo += cur[i + (size_t)W * 3]; cnt += 1; }
                    if ((iy - 1) >= 0) { lp += P[i - (size_t)W * 3]; lo += cur[i - (size_t)W * 3]; cnt += 1; }
                    lp -= (T)cnt * P[i];
                    lo -= (T)cnt * cur[i];
                    const T t1 = Wt[i] * (cur[i] - A[i]);
                    const T t2 = lo - lp;
                    const T g = t2 - t1;
                    T o;
                    if (k > 0) o = (cur[i] + step * g) + mom * U[i];
                    else o = cur[i] + step * g;
                    Un[i] = step * g;
                    nxt[i] = o;
                }
        std::swap(cur, nxt);
        U.swap(Un);
    }
    if (cur != O) std::memcpy(O, cur, n * sizeof(T));
}

/* The same solver in FMA form (the library's VSC_FRAME_FMA mode): the
 * algebraically equal update
 *   g = sumO - (cnt + w)*O - (lapP - w*A)
 * with per-pixel constants q = cnt + w and k = s*fma(-w, A, lapP) (lapP as
 * above), and per sweep
 *   X = fma(-q, O, sumO); U' = fma(s, X, -k); O' = O + U' (then fma(m, U, O')
 *   for k > 0).
 * Each fma rounds once (std::fmaf), so the result differs from the reference
 * op order by rounding only. */
template <typename T>
static void stab_consist_out_fma(const T* P, const T* A, const T* Wt, int iters, T step, T mom, T* O, int W, int H)
{
    const size_t n = (size_t)W * H * 3;
    std::vector<T> Ob(n), U(n, (T)0), Un(n), Q(n), K(n);
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < H; ++iy)
        for (int ix = 0; ix < W; ++ix)
            for (int c = 0; c < 3; ++c) {
                const size_t i = ((size_t)iy * W + ix) * 3 + c;
                int cnt = 0;
                T lp = 0;
                if ((ix + 1) < (W - 1)) { lp += P[i + 3]; cnt += 1; }
                if ((ix - 1) >= 0) { lp += P[i - 3]; cnt += 1; }
                if ((iy + 1) < (H - 1)) { lp += P[i + (size_t)W * 3]; cnt += 1; }
                if ((iy - 1) >= 0) { lp += P[i - (size_t)W * 3]; cnt += 1; }
                lp -= (T)cnt * P[i];
                Q[i] = (T)cnt + Wt[i];
                K[i] = step * std::fma(-Wt[i], A[i], lp);
            }
    T* cur = O;
    T* nxt = Ob.data();
    for (int k = 0; k < iters; ++k) {
#pragma omp parallel for schedule(static)
        for (int iy = 0; iy < H; ++iy)
            for (int ix = 0; ix < W; ++ix)
                for (int c = 0; c < 3; ++c) {
                    const size_t i = ((size_t)iy * W + ix) * 3 + c;
                    T sum = 0;
                    if ((ix + 1) < (W - 1)) sum += cur[i + 3];
                    if ((ix - 1) >= 0) sum += cur[i - 3];
                    if ((iy + 1) < (H - 1)) sum += cur[i + (size_t)W * 3];
                    if ((iy - 1) >= 0) sum += cur[i - (size_t)W * 3];
                    const T X = std::fma(-Q[i], cur[i], sum);
                    const T sg = std::fma(step, X, -K[i]);
                    T o = cur[i] + sg;
                    if (k > 0) o = std::fma(mom, U[i], o);
                    Un[i] = sg;
                    nxt[i] = o;
                }
        std::swap(cur, nxt);
        U.swap(Un);
    }
    if (cur != O) std::memcpy(O, cur, n * sizeof(T));
}

/* kernel_bilinear, flowconsistency.cu:50-75: xx = (ox*inW)/outW (no half-pixel
 * offset), neighbours clamped to the last row/column, channels = out channels,
 * read with the input's channel stride. */
template <typename T>
static void stab_bilinear(const T* in, int inW, int inH, int inC, T* out, int outW, int outH, int outC)
{
#pragma omp parallel for schedule(static)
    for (int oy = 0; oy < outH; ++oy)
        for (int ox = 0; ox < outW; ++ox) {
            const float xx = ((float)ox * (float)inW) / (float)outW;
            const float yy = ((float)oy * (float)inH) / (float)outH;
            const int ix = (int)std::floor(xx);
            const int iy = (int)std::floor(yy);
            const T fx = (T)(xx - (float)ix);
            const T fy = (T)(yy - (float)iy);
            const int ix1 = std::min(ix + 1, inW - 1), iy1 = std::min(iy + 1, inH - 1);
            for (int c = 0; c < outC; ++c) {
                const T v00 = in[((size_t)iy * inW + ix) * inC + c];
                const T v10 = in[((size_t)iy * inW + ix1) * inC + c];
                const T v01 = in[((size_t)iy1 * inW + ix) * inC + c];
                const T v11 = in[((size_t)iy1 * inW + ix1) * inC + c];
                const T v = ((v00 * ((T)1 - fx)) * ((T)1 - fy) + (v10 * fx) * ((T)1 - fy)) +
                            (v01 * ((T)1 - fx)) * fy + (v11 * fx) * fy;
                out[((size_t)oy * outW + ox) * outC + c] = v;
            }
        }
}

VSCO_API void vsco_stab_warp_f32(const float* in, const float* flow, int flow_ch, float* out, int W, int H)
{ stab_warp<float>(in, flow, flow_ch, out, W, H); }
VSCO_API void vsco_stab_warp_f64(const double* in, const double* flow, int flow_ch, double* out, int W, int H)
{ stab_warp<double>(in, flow, flow_ch, out, W, H); }
VSCO_API void vsco_stab_adap_comb_f32(const float* cIn, const float* cPr, const float* pIn, const float* pPr,
                                      const float* nIn, const float* nPr, float* adIn, float* adPr,
                                      const float* last, int W, int H, float alpha)
{ stab_adap_comb<float>(cIn, cPr, pIn, pPr, nIn, nPr, adIn, adPr, last, W, H, alpha); }
VSCO_API void vsco_stab_adap_comb_f64(const double* cIn, const double* cPr, const double* pIn, const double* pPr,
                                      const double* nIn, const double* nPr, double* adIn, double* adPr,
                                      const double* last, int W, int H, double alpha)
{ stab_adap_comb<double>(cIn, cPr, pIn, pPr, nIn, nPr, adIn, adPr, last, W, H, alpha); }
VSCO_API void vsco_stab_consist_wt_f32(const float* adIn, const float* cIn, float* wt, int W, int H, float beta, float gamma)
{ stab_consist_wt<float>(adIn, cIn, wt, W, H, beta, gamma); }
VSCO_API void vsco_stab_consist_wt_f64(const double* adIn, const double* cIn, double* wt, int W, int H, double beta, double gamma)
{ stab_consist_wt<double>(adIn, cIn, wt, W, H, beta, gamma); }
VSCO_API void vsco_stab_consist_out_f32(const float* P, const float* A, const float* Wt, int iters, float step, float mom,
                                        float* O, int W, int H)
{ stab_consist_out<float>(P, A, Wt, iters, step, mom, O, W, H); }
VSCO_API void vsco_stab_consist_out_f64(const double* P, const double* A, const double* Wt, int iters, double step,
                                        double mom, double* O, int W, int H)
{ stab_consist_out<double>(P, A, Wt, iters, step, mom, O, W, H); }
VSCO_API void vsco_stab_consist_out_fma_f32(const float* P, const float* A, const float* Wt, int iters, float step,
                                            float mom, float* O, int W, int H)
{ stab_consist_out_fma<float>(P, A, Wt, iters, step, mom, O, W, H); }
VSCO_API void vsco_stab_bilinear_f32(const float* in, int inW, int inH, int inC, float* out, int outW, int outH, int outC)
{ stab_bilinear<float>(in, inW, inH, inC, out, outW, outH, outC); }
VSCO_API void vsco_stab_bilinear_f64(const double* in, int inW, int inH, int inC, double* out, int outW, int outH, int outC)
{ stab_bilinear<double>(in, inW, inH, inC, out, outW, outH, outC); }

/* kernel_average_frames, flowconsistency.cu:35-48 (0.5 is a double literal). */
VSCO_API void vsco_average_frames_f32(const float* a, const float* b, float* out, int W, int H, int ch)
{
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            for (int c = 0; c < 3; ++c) {
                const size_t i = ((size_t)y * W + x) * ch + c;
                out[i] = (float)(0.5 * (double)a[i] + 0.5 * (double)b[i]);
            }
}

/* ------------------------------------------------------------------------- */
/* u8 <-> float conversions (src/stabilization/gpuimage.cu)                  */
/* ------------------------------------------------------------------------- */

/* kernel_to_float_image, gpuimage.cu:39-51: RGBA8 -> RGB float, v/255.0 in
 * double then stored as float; alpha dropped. */
VSCO_API void vsco_rgba_to_float(const uint8_t* rgba, float* out, int W, int H)
{
    for (size_t p = 0; p < (size_t)W * H; ++p)
        for (int c = 0; c < 3; ++c) out[p * 3 + c] = (float)((double)rgba[p * 4 + c] / 255.0);
}

/* kernel_to_char_image, gpuimage.cu:54-67: u8 = (uchar)__float2uint_rd(fabs(v)*255):
 * round toward -inf to uint32 with saturation (NaN -> 0), then truncation
 * mod 256 by the unsigned char cast; alpha byte = 1. */
static inline uint8_t to_u8(float v)
{
    const float f = std::fabs(v) * 255.0f;
    uint32_t u;
    if (!(f == f)) u = 0;
    else if (f >= 4294967296.0f) u = 0xFFFFFFFFu;
    else u = (uint32_t)std::floor(f);
    return (uint8_t)(u & 0xFFu);
}

VSCO_API void vsco_float_to_rgba(const float* in, uint8_t* rgba, int W, int H)
{
    for (size_t p = 0; p < (size_t)W * H; ++p) {
        for (int c = 0; c < 3; ++c) rgba[p * 4 + c] = to_u8(in[p * 3 + c]);
        rgba[p * 4 + 3] = 1;
    }
}

/* FP64 twin of the output quantiser (same semantics, double input). */
VSCO_API void vsco_double_to_rgba(const double* in, uint8_t* rgba, int W, int H)
{
    for (size_t p = 0; p < (size_t)W * H; ++p) {
        for (int c = 0; c < 3; ++c) {
            const double f = std::fabs(in[p * 3 + c]) * 255.0;
            uint32_t u = (f >= 4294967296.0) ? 0xFFFFFFFFu : (uint32_t)std::floor(f);
            rgba[p * 4 + c] = (uint8_t)(u & 0xFFu);
        }
        rgba[p * 4 + 3] = 1;
    }
}

/* ------------------------------------------------------------------------- */
/* One stabilized frame: the stabilization body of VideoStabilizer::doOneStep */
/* (videostabilizer.cpp:177-238) for flows that are already at frame size.   */
/* ------------------------------------------------------------------------- */

struct FrameParams {
    float alpha, beta, gamma;
    int pyramid_levels, num_iter;
    float step_size, mom_fac;
    bool fma;  /* solver in FMA form (stab_consist_out_fma) */
    /* Planted schedule mistakes, used only by the tests that show the
     * schedule pins detect them (0 = the reference schedule):
     * 1 prev frames warped with the forward flow, 2 next frames warped with
     * the backward flow, 3 last output warped with the forward flow,
     * 4 every level runs num_iter sweeps, 5 level j runs num_iter/(j+1)+1,
     * 6 the finer level keeps P instead of the upsampled coarse solution. */
    int mutation = 0;
};

/* Inputs: the 3-frame window of originals and processed frames as RGB float
 * HWC images, fwd flow (cur->next) and bwd flow (cur->prev) HWC with flow_ch
 * channels, and the previous stabilized frame.  Output: the stabilized frame
 * (float HWC; it becomes the next call's `last`).
 * Schedule (videostabilizer.cpp:182-228):
 *   prevWarpIn/Pr = warp(orig/proc[prev], bwd); nextWarpIn/Pr = warp(.., fwd);
 *   lastWarp = warp(last, bwd); adap_comb(alpha); consist_wt(beta, gamma);
 *   level 0 = copies of P, adPr, w and O0 = P; level j = bilinear(level j-1)
 *   of each (dims halved by integer division);
 *   for j = L-1..0: if j < L-1: O_j = bilinear(O_{j+1});
 *                   consist_out(P_j, A_j, w_j, num_iter/(j+1)).            */
template <typename T>
static void stab_frame(const T* oPrev, const T* oCur, const T* oNext, const T* pPrev, const T* pCur, const T* pNext,
                       const T* fwd, const T* bwd, int flow_ch, const T* last, const FrameParams& prm, T* out, int W, int H)
{
    const size_t n = (size_t)W * H * 3;
    std::vector<T> pwi(n), pwp(n), nwi(n), nwp(n), lw(n), adIn(n), adPr(n), wt(n);
    const int mu = prm.mutation;
    const T* prevFlow = mu == 1 ? fwd : bwd;
    const T* nextFlow = mu == 2 ? bwd : fwd;
    const T* lastFlow = mu == 3 ? fwd : bwd;
    stab_warp<T>(oPrev, prevFlow, flow_ch, pwi.data(), W, H);
    stab_warp<T>(pPrev, prevFlow, flow_ch, pwp.data(), W, H);
    stab_warp<T>(oNext, nextFlow, flow_ch, nwi.data(), W, H);
    stab_warp<T>(pNext, nextFlow, flow_ch, nwp.data(), W, H);
    stab_warp<T>(last, lastFlow, flow_ch, lw.data(), W, H);
    stab_adap_comb<T>(oCur, pCur, pwi.data(), pwp.data(), nwi.data(), nwp.data(), adIn.data(), adPr.data(), lw.data(), W,
                      H, (T)prm.alpha);
    stab_consist_wt<T>(adIn.data(), oCur, wt.data(), W, H, (T)prm.beta, (T)prm.gamma);

    const int L = prm.pyramid_levels;
    std::vector<int> lw_(L), lh_(L);
    std::vector<std::vector<T>> pP(L), pA(L), pW(L), pO(L);
    int w = W, h = H;
    for (int j = 0; j < L; ++j) {
        lw_[j] = w; lh_[j] = h;
        const size_t m = (size_t)w * h * 3;
        pP[j].resize(m); pA[j].resize(m); pW[j].resize(m); pO[j].resize(m);
        w /= 2; h /= 2;
    }
    for (int j = 0; j < L; ++j) {
        if (j == 0) {
            std::memcpy(pP[0].data(), pCur, n * sizeof(T));
            std::memcpy(pA[0].data(), adPr.data(), n * sizeof(T));
            std::memcpy(pW[0].data(), wt.data(), n * sizeof(T));
            std::memcpy(pO[0].data(), pCur, n * sizeof(T));
        } else {
            stab_bilinear<T>(pP[j - 1].data(), lw_[j - 1], lh_[j - 1], 3, pP[j].data(), lw_[j], lh_[j], 3);
            stab_bilinear<T>(pA[j - 1].data(), lw_[j - 1], lh_[j - 1], 3, pA[j].data(), lw_[j], lh_[j], 3);
            stab_bilinear<T>(pW[j - 1].data(), lw_[j - 1], lh_[j - 1], 3, pW[j].data(), lw_[j], lh_[j], 3);
            stab_bilinear<T>(pO[j - 1].data(), lw_[j - 1], lh_[j - 1], 3, pO[j].data(), lw_[j], lh_[j], 3);
        }
    }
    for (int j = L - 1; j >= 0; --j) {
        if (j != L - 1 && mu != 6)
            stab_bilinear<T>(pO[j + 1].data(), lw_[j + 1], lh_[j + 1], 3, pO[j].data(), lw_[j], lh_[j], 3);
        const int iters = mu == 4 ? prm.num_iter : mu == 5 ? prm.num_iter / (j + 1) + 1 : prm.num_iter / (j + 1);
        if (prm.fma)
            stab_consist_out_fma<T>(pP[j].data(), pA[j].data(), pW[j].data(), iters, (T)prm.step_size,
                                    (T)prm.mom_fac, pO[j].data(), lw_[j], lh_[j]);
        else
            stab_consist_out<T>(pP[j].data(), pA[j].data(), pW[j].data(), iters, (T)prm.step_size, (T)prm.mom_fac,
                                pO[j].data(), lw_[j], lh_[j]);
    }
    std::memcpy(out, pO[0].data(), n * sizeof(T));
}

VSCO_API void vsco_stab_frame_f32(const float* oPrev, const float* oCur, const float* oNext, const float* pPrev,
                                  const float* pCur, const float* pNext, const float* fwd, const float* bwd, int flow_ch,
                                  const float* last, float alpha, float beta, float gamma, int levels, int iters,
                                  float step, float mom, float* out, int W, int H)
{
    FrameParams p{alpha, beta, gamma, levels, iters, step, mom, false};
    stab_frame<float>(oPrev, oCur, oNext, pPrev, pCur, pNext, fwd, bwd, flow_ch, last, p, out, W, H);
}

VSCO_API void vsco_stab_frame_fma_f32(const float* oPrev, const float* oCur, const float* oNext, const float* pPrev,
                                      const float* pCur, const float* pNext, const float* fwd, const float* bwd,
                                      int flow_ch, const float* last, float alpha, float beta, float gamma, int levels,
                                      int iters, float step, float mom, float* out, int W, int H)
{
    FrameParams p{alpha, beta, gamma, levels, iters, step, mom, true};
    stab_frame<float>(oPrev, oCur, oNext, pPrev, pCur, pNext, fwd, bwd, flow_ch, last, p, out, W, H);
}

VSCO_API void vsco_stab_frame_f64(const double* oPrev, const double* oCur, const double* oNext, const double* pPrev,
                                  const double* pCur, const double* pNext, const double* fwd, const double* bwd,
                                  int flow_ch, const double* last, float alpha, float beta, float gamma, int levels,
                                  int iters, float step, float mom, double* out, int W, int H)
{
    FrameParams p{alpha, beta, gamma, levels, iters, step, mom, false};
    stab_frame<double>(oPrev, oCur, oNext, pPrev, pCur, pNext, fwd, bwd, flow_ch, last, p, out, W, H);
}

/* stab_frame with a planted schedule mistake (FrameParams::mutation), float64:
 * only for the tests that show the schedule pins fail on it. */
VSCO_API void vsco_stab_frame_mutant_f64(const double* oPrev, const double* oCur, const double* oNext,
                                         const double* pPrev, const double* pCur, const double* pNext,
                                         const double* fwd, const double* bwd, int flow_ch, const double* last,
                                         float alpha, float beta, float gamma, int levels, int iters, float step,
                                         float mom, double* out, int W, int H, int mutation)
{
    FrameParams p{alpha, beta, gamma, levels, iters, step, mom, false};
    p.mutation = mutation;
    stab_frame<double>(oPrev, oCur, oNext, pPrev, pCur, pNext, fwd, bwd, flow_ch, last, p, out, W, H);
}

/* ------------------------------------------------------------------------- */
/* .flo files (src/stabilization/flowIO.cpp)                                  */
/* ------------------------------------------------------------------------- */

/* ReadFlowFile, flowIO.cpp:31-78: float tag 202021.25, int32 width, height in
 * [1, 99999], then W*H*2 interleaved floats (u, v) row by row, and nothing
 * after.  Returns 0 on success, a negative code naming the failed check:
 * -1 open, -2 short header, -3 tag, -4 width, -5 height, -6 short data,
 * -7 trailing bytes, -8 capacity (caller's buffer smaller than W*H*2). */
VSCO_API int vsco_read_flo(const char* path, float* flow, long long capacity, int* width, int* height)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return -1;
    float tag; int w, h;
    if (std::fread(&tag, 4, 1, f) != 1 || std::fread(&w, 4, 1, f) != 1 || std::fread(&h, 4, 1, f) != 1) { std::fclose(f); return -2; }
    if (tag != 202021.25f) { std::fclose(f); return -3; }
    if (w < 1 || w > 99999) { std::fclose(f); return -4; }
    if (h < 1 || h > 99999) { std::fclose(f); return -5; }
    *width = w; *height = h;
    const long long need = (long long)w * h * 2;
    if (flow == nullptr) { std::fclose(f); return 0; }
    if (need > capacity) { std::fclose(f); return -8; }
    for (int y = 0; y < h; ++y)
        if (std::fread(flow + (size_t)y * w * 2, 4, (size_t)w * 2, f) != (size_t)w * 2) { std::fclose(f); return -6; }
    if (std::fgetc(f) != EOF) { std::fclose(f); return -7; }
    std::fclose(f);
    return 0;
}
