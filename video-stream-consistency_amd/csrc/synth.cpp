// Deterministic synthetic video for the stabilization workloads: seed-generated
// texture frames with global sub-pixel motion, independently moving occluder
// patches, flickering "processed" frames and exact ground-truth flows.
// Host-only (g++), no GPU; the same seed gives the same bytes on every machine
// (SplitMix64 stream, double arithmetic, explicit rounding).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

typedef struct {
    int width, height, nframes;
    int nblobs;           // Gaussian blobs of the background texture
    float sigma_min, sigma_max;
    float motion;         // per-frame global motion ~ U[-motion, motion]^2 px
    int npatches;         // independently moving occluders
    int patch_size;
    float gain_min, gain_max;   // processed = orig*gain + offset + N(0, noise_sigma)
    float offset_max;
    float noise_sigma;
    unsigned long long seed;
} vsc_synth_cfg;

int vsc_synth_generate(const vsc_synth_cfg* cfg, uint8_t* orig, uint8_t* proc, float* fwd, float* bwd);

}

namespace {

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
    double uni(double a, double b) { return a + (b - a) * uni(); }
};

inline uint64_t mix(uint64_t a, uint64_t b)
{
    Rng r{a * 0x100000001B3ull ^ b};
    return r.next();
}

struct Blob {
    double cx, cy, sig, amp[3];
};

struct Patch {
    double x, y;  // top-left at frame 0
    double col[3], fx, fy, ph;
};

inline uint8_t q8(double v)
{
    if (v < 0.0) v = 0.0;
    if (v > 255.0) v = 255.0;
    return (uint8_t)std::floor(v + 0.5);
}

}  // namespace

extern "C" int vsc_synth_generate(const vsc_synth_cfg* cfg, uint8_t* orig, uint8_t* proc, float* fwd, float* bwd)
{
    if (!cfg || cfg->width < 4 || cfg->height < 4 || cfg->nframes < 1 || cfg->nblobs < 0 || cfg->npatches < 0)
        return -1;
    const int W = cfg->width, H = cfg->height, N = cfg->nframes;
    Rng rng{cfg->seed};

    // global motion and accumulated shift s_t = sum_{tau < t} d_tau
    std::vector<double> dx(N), dy(N), sx(N + 1, 0.0), sy(N + 1, 0.0);
    for (int t = 0; t < N; ++t) {
        dx[t] = rng.uni(-cfg->motion, cfg->motion);
        dy[t] = rng.uni(-cfg->motion, cfg->motion);
        sx[t + 1] = sx[t] + dx[t];
        sy[t + 1] = sy[t] + dy[t];
    }
    double minsx = 0, maxsx = 0, minsy = 0, maxsy = 0;
    for (int t = 0; t < N; ++t) {
        minsx = std::fmin(minsx, sx[t]); maxsx = std::fmax(maxsx, sx[t]);
        minsy = std::fmin(minsy, sy[t]); maxsy = std::fmax(maxsy, sy[t]);
    }
    // blob centres cover the canvas the view slides over
    const double m = 3.0 * cfg->sigma_max;
    std::vector<Blob> blobs(cfg->nblobs);
    for (auto& b : blobs) {
        b.cx = rng.uni(-maxsx - m, W - minsx + m);
        b.cy = rng.uni(-maxsy - m, H - minsy + m);
        b.sig = rng.uni(cfg->sigma_min, cfg->sigma_max);
        for (int c = 0; c < 3; ++c) b.amp[c] = rng.uni(-1.0, 1.0);
    }
    // texture scale: the blob sum has variance ~ nblobs * E[area]/canvas / 3
    const double canvas = (W + (maxsx - minsx) + 2 * m) * (H + (maxsy - minsy) + 2 * m);
    double area = 0;
    for (auto& b : blobs) area += 3.14159265358979 * b.sig * b.sig;
    const double dens = area / canvas;
    const double scale = 60.0 / std::sqrt(std::fmax(dens / 3.0, 1e-6));

    // occluders: position p_t = p_0 + sum of their own per-frame motion
    const int PS = cfg->patch_size;
    std::vector<Patch> patches(cfg->npatches);
    std::vector<double> pvx((size_t)cfg->npatches * N), pvy((size_t)cfg->npatches * N);
    for (int k = 0; k < cfg->npatches; ++k) {
        Patch& p = patches[k];
        p.x = rng.uni(0, std::fmax(1.0, W - PS));
        p.y = rng.uni(0, std::fmax(1.0, H - PS));
        for (int c = 0; c < 3; ++c) p.col[c] = rng.uni(40, 215);
        p.fx = rng.uni(0.2, 0.7);
        p.fy = rng.uni(0.2, 0.7);
        p.ph = rng.uni(0, 6.28318530718);
        for (int t = 0; t < N; ++t) {
            pvx[(size_t)k * N + t] = rng.uni(-cfg->motion, cfg->motion);
            pvy[(size_t)k * N + t] = rng.uni(-cfg->motion, cfg->motion);
        }
    }
    std::vector<double> ppx((size_t)cfg->npatches * (N + 1)), ppy((size_t)cfg->npatches * (N + 1));
    for (int k = 0; k < cfg->npatches; ++k) {
        ppx[(size_t)k * (N + 1)] = patches[k].x;
        ppy[(size_t)k * (N + 1)] = patches[k].y;
        for (int t = 0; t < N; ++t) {
            ppx[(size_t)k * (N + 1) + t + 1] = ppx[(size_t)k * (N + 1) + t] + pvx[(size_t)k * N + t];
            ppy[(size_t)k * (N + 1) + t + 1] = ppy[(size_t)k * (N + 1) + t] + pvy[(size_t)k * N + t];
        }
    }
    // processed-frame flicker per frame and channel
    std::vector<double> gain((size_t)N * 3), offs((size_t)N * 3);
    for (int t = 0; t < N; ++t)
        for (int c = 0; c < 3; ++c) {
            gain[t * 3 + c] = rng.uni(cfg->gain_min, cfg->gain_max);
            offs[t * 3 + c] = rng.uni(-cfg->offset_max, cfg->offset_max);
        }
    const uint64_t noise_seed = rng.next();

#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < N; ++t) {
        std::vector<double> img((size_t)W * H * 3, 0.0);
        std::vector<double> gx, gy;
        for (const Blob& b : blobs) {
            const double cx = b.cx + sx[t], cy = b.cy + sy[t];
            const double r = 3.0 * b.sig;
            const int x0 = std::max(0, (int)std::floor(cx - r)), x1 = std::min(W - 1, (int)std::ceil(cx + r));
            const int y0 = std::max(0, (int)std::floor(cy - r)), y1 = std::min(H - 1, (int)std::ceil(cy + r));
            if (x0 > x1 || y0 > y1) continue;
            const double k = -0.5 / (b.sig * b.sig);
            gx.resize(x1 - x0 + 1);
            gy.resize(y1 - y0 + 1);
            for (int x = x0; x <= x1; ++x) gx[x - x0] = std::exp(k * (x - cx) * (x - cx));
            for (int y = y0; y <= y1; ++y) gy[y - y0] = std::exp(k * (y - cy) * (y - cy));
            for (int y = y0; y <= y1; ++y) {
                double* row = &img[((size_t)y * W) * 3];
                const double vy = gy[y - y0];
                for (int x = x0; x <= x1; ++x) {
                    const double g = gx[x - x0] * vy;
                    row[x * 3 + 0] += g * b.amp[0];
                    row[x * 3 + 1] += g * b.amp[1];
                    row[x * 3 + 2] += g * b.amp[2];
                }
            }
        }
        const double fdx = (t < N - 1) ? dx[t] : 0.0, fdy = (t < N - 1) ? dy[t] : 0.0;
        const double bdx = (t > 0) ? -dx[t - 1] : 0.0, bdy = (t > 0) ? -dy[t - 1] : 0.0;
        for (size_t p = 0; p < (size_t)W * H; ++p) {
            for (int c = 0; c < 3; ++c) img[p * 3 + c] = 128.0 + scale * img[p * 3 + c];
            if (fwd) { fwd[((size_t)t * W * H + p) * 2] = (float)fdx; fwd[((size_t)t * W * H + p) * 2 + 1] = (float)fdy; }
            if (bwd) { bwd[((size_t)t * W * H + p) * 2] = (float)bdx; bwd[((size_t)t * W * H + p) * 2 + 1] = (float)bdy; }
        }
        // occluders, drawn in order (later ones on top); texture moves with the patch
        for (int k = 0; k < cfg->npatches; ++k) {
            const Patch& pa = patches[k];
            const double px = ppx[(size_t)k * (N + 1) + t], py = ppy[(size_t)k * (N + 1) + t];
            const double vfx = (t < N - 1) ? pvx[(size_t)k * N + t] : 0.0, vfy = (t < N - 1) ? pvy[(size_t)k * N + t] : 0.0;
            const double vbx = (t > 0) ? -pvx[(size_t)k * N + t - 1] : 0.0, vby = (t > 0) ? -pvy[(size_t)k * N + t - 1] : 0.0;
            const int x0 = std::max(0, (int)std::ceil(px)), x1 = std::min(W - 1, (int)std::floor(px + PS - 1e-9));
            const int y0 = std::max(0, (int)std::ceil(py)), y1 = std::min(H - 1, (int)std::floor(py + PS - 1e-9));
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    const double u = x - px, v = y - py;
                    const double tex = std::sin(pa.fx * u + pa.ph) * std::cos(pa.fy * v);
                    const size_t p = (size_t)y * W + x;
                    for (int c = 0; c < 3; ++c) img[p * 3 + c] = pa.col[c] + 35.0 * tex * (c == 1 ? -1.0 : 1.0);
                    if (fwd) { fwd[((size_t)t * W * H + p) * 2] = (float)vfx; fwd[((size_t)t * W * H + p) * 2 + 1] = (float)vfy; }
                    if (bwd) { bwd[((size_t)t * W * H + p) * 2] = (float)vbx; bwd[((size_t)t * W * H + p) * 2 + 1] = (float)vby; }
                }
        }
        for (size_t p = 0; p < (size_t)W * H; ++p) {
            uint8_t* o = orig ? &orig[((size_t)t * W * H + p) * 4] : nullptr;
            uint8_t* q = proc ? &proc[((size_t)t * W * H + p) * 4] : nullptr;
            for (int c = 0; c < 3; ++c) {
                const uint8_t ov = q8(img[p * 3 + c]);
                if (o) o[c] = ov;
                if (q) {
                    // Box-Muller from a counter-based stream: independent of thread order
                    const uint64_t h1 = mix(noise_seed ^ (uint64_t)t, p * 3 + c);
                    const uint64_t h2 = mix(h1, 0x5851F42D4C957F2Dull);
                    const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740993.0);
                    const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
                    const double nz = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                    q[c] = q8((double)ov * gain[t * 3 + c] + offs[t * 3 + c] + cfg->noise_sigma * nz);
                }
            }
            if (o) o[3] = 255;
            if (q) q[3] = 255;
        }
    }
    return 0;
}
