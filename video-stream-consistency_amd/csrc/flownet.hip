// PWCNet-light flow network ("pwcnet-my-light4dec": Extractor, DecoderLight
// at levels 6..3, Refiner; pwcnet.py:179-253, 458-611, 625-659, 787-1014)
// running on the library's own kernels, in FP32 or BF16 storage (FP32
// accumulation everywhere).  The trained weights of the reference are not
// available, so the weights are drawn from a seed with the reference's
// initialisation (pwcnet.py:887-891: Kaiming normal, fan_in, LeakyReLU a=0.1,
// zero biases) or loaded through vsc_flownet_set_params.
//
// Layout: channels-last tensors.  Each decoder level keeps two concatenation
// buffers whose channel slices are the reference's torch.cat operands, so no
// concatenation is ever materialised twice:
//   cat0 = [ leaky(corr) 81 | 0-pad to 88 | f1 (C, padded) | upflow 2 | pad ]
//   dec  = [ x1 128 | x2 128 | x3 96 | x4 64 | x5 32 | flow 2 | pad ]
// so cat([x1,x2]) = dec[0,256), cat([x2,x3]) = dec[128,352), cat([x3,x4]) =
// dec[256,416), cat([x4,x5]) = dec[352,448), refiner input cat([x5,flow]) =
// dec[416, 416+40).  Padding channels hold zeros and meet zero weights.
#include "vsc_internal.h"
#include "flownet_internal.h"

#include <cmath>
#include <new>
#include <vector>

namespace vsc {
namespace {

constexpr int kFeatC[7] = {3, 16, 32, 64, 96, 128, 196};   // Extractor channels per level (pwcnet.py:175)
constexpr float kDbl[7] = {0, 0, 0, 2.5f, 1.25f, 0.625f, 0};  // dblBackward per decoder level (pwcnet.py:474)
constexpr int kDecOut[6] = {128, 128, 96, 64, 32, 2};
constexpr int kDecIn0[6] = {0, 0, 0, 128, 256, 352};       // input slice offsets inside dec (conv 2..6)
constexpr int kDecInC[6] = {0, 128, 256, 224, 160, 96};
constexpr int kDecOut0[5] = {0, 128, 256, 352, 416};
constexpr int kDecC = 456;                                  // dec channels: 448 + flow slot (8)
constexpr int kRefIn0 = 416, kRefInC = 40;
constexpr int kRefOut[7] = {128, 128, 128, 96, 64, 32, 2};
constexpr int kRefDil[7] = {1, 2, 4, 8, 16, 1, 1};
constexpr size_t kPartialFloatsPerItem = (size_t)2 << 20;  // 8 MB of FP32 partial sums per batch item

inline int rup(int x, int m) { return (x + m - 1) / m * m; }

// deterministic normal draws: SplitMix64 uniforms, Box-Muller
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        const double u = uni(), v = uni();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
};

struct Layer {
    int cin, cout, k;        // reference shape: Conv2d [cout][cin][k][k]; ConvTranspose2d (k=4) [cin][cout][4][4]
    int stride, dil, act;
    std::vector<int> cmap;   // reference input channel -> padded position (conv layers)
    int cin_pad;             // padded input channels of the device weights
    std::vector<float> w, b; // reference layout (host)
    void* dw = nullptr;      // device weights: conv [cout][3][3][cin_pad] (element kind), upflow FP32 reference layout
    float* db = nullptr;     // device bias
};

}  // namespace
}  // namespace vsc

struct vsc_flownet {
    int W, H, down, Wd, Hd, Wp, Hp;
    int ek, eb, vec, slots, pairs;       // pairs: the most (first, second) pairs one decode batches
    int frames;                          // the most frames one extraction batches
    bool corr_bf16 = false;              // FP32 network: cost volume products on the BF16 MFMA
    std::vector<vsc::Layer> L;
    int ext0, up0, dec0, ref0;            // first layer index of each group
    void* img = nullptr;                  // [frames][Hp][Wp][vec]
    void* tmp[2] = {nullptr, nullptr};    // extractor scratch, [frames][level-1 size]
    void* feat[7] = {};                   // levels 3..6: [slots][h][w][CV], slot-major
    void* cat0[7] = {};                   // per level [pairs][h][w][ccat]
    void* dec[7] = {};                    // per level [pairs][h][w][kDecC]
    void* warped = nullptr;               // [pairs][h3][w3][CV max]
    float* upflow[7] = {};                // per level [pairs][h][w][2]
    float* flow[7] = {};                  // per level [pairs][h][w][2]
    float* refined = nullptr;             // [pairs][h3][w3][2]
    void* rbuf[2] = {nullptr, nullptr};   // refiner ping-pong [pairs][h3][w3][128]
    float* partial = nullptr;             // split-channel partial sums of the small decoder layers
    float* partial_x = nullptr;           // the same for the extractor ([frames] items): extraction and
                                          // decoding may run on two streams at once
    std::vector<void*> owned;
    int last_pairs = 0;
};

namespace vsc {
namespace {

int ccat_of(const vsc_flownet* n, int lvl) { return 88 + rup(kFeatC[lvl], n->vec) + n->vec; }
int cv_of(const vsc_flownet* n, int lvl) { return rup(kFeatC[lvl], n->vec); }

void add_conv(vsc_flownet* n, int cin, int cout, int stride, int dil, int act, const std::vector<int>& cmap, int cin_pad)
{
    Layer l;
    l.cin = cin; l.cout = cout; l.k = 3; l.stride = stride; l.dil = dil; l.act = act;
    l.cmap = cmap; l.cin_pad = cin_pad;
    n->L.push_back(l);
}

std::vector<int> ident(int c)
{
    std::vector<int> m(c);
    for (int i = 0; i < c; ++i) m[i] = i;
    return m;
}

void build_layers(vsc_flownet* n)
{
    const int v = n->vec;
    n->ext0 = 0;
    for (int k = 1; k <= 6; ++k) {
        const int ci = kFeatC[k - 1], co = kFeatC[k];
        add_conv(n, ci, co, 2, 1, 1, ident(ci), rup(ci, v));
        add_conv(n, co, co, 1, 1, 1, ident(co), rup(co, v));
        add_conv(n, co, co, 1, 1, 1, ident(co), rup(co, v));
    }
    n->up0 = (int)n->L.size();
    for (int lvl = 5; lvl >= 3; --lvl) {
        Layer l;
        l.cin = 2; l.cout = 2; l.k = 4; l.stride = 2; l.dil = 1; l.act = 0; l.cin_pad = 2;
        n->L.push_back(l);
    }
    n->dec0 = (int)n->L.size();
    for (int lvl = 6; lvl >= 3; --lvl) {
        const int C = kFeatC[lvl], CV = cv_of(n, lvl);
        // cat([corr 81, f1 C, flow 2]) -> padded positions
        std::vector<int> m0;
        for (int i = 0; i < 81; ++i) m0.push_back(i);
        for (int i = 0; i < C; ++i) m0.push_back(88 + i);
        m0.push_back(88 + CV);
        m0.push_back(88 + CV + 1);
        add_conv(n, 83 + C, 128, 1, 1, 1, m0, ccat_of(n, lvl));
        for (int j = 1; j < 6; ++j)
            add_conv(n, kDecInC[j], kDecOut[j], 1, 1, j < 5 ? 1 : 0, ident(kDecInC[j]), kDecInC[j]);
    }
    n->ref0 = (int)n->L.size();
    int ci = 34;
    for (int j = 0; j < 7; ++j) {
        add_conv(n, ci, kRefOut[j], 1, kRefDil[j], j < 6 ? 1 : 0, ident(ci), j == 0 ? kRefInC : ci);
        ci = kRefOut[j];
    }
}

void init_params(vsc_flownet* n, unsigned long long seed)
{
    const double gain = std::sqrt(2.0 / (1.0 + 0.1 * 0.1));
    for (size_t i = 0; i < n->L.size(); ++i) {
        Layer& l = n->L[i];
        const size_t nw = (size_t)l.cin * l.cout * l.k * l.k;
        const double fan_in = l.k == 4 ? (double)l.cout * 16 : (double)l.cin * 9;
        const double sd = gain / std::sqrt(fan_in);
        Rng r{seed * 0x2545F4914F6CDD1Dull + 0x1000 * (i + 1)};
        l.w.resize(nw);
        for (size_t j = 0; j < nw; ++j) l.w[j] = (float)(sd * r.normal());
        l.b.assign(l.cout, 0.0f);
    }
}

size_t param_count_of(const vsc_flownet* n)
{
    size_t c = 0;
    for (const Layer& l : n->L) c += l.w.size() + l.b.size();
    return c;
}

int upload_params(vsc_flownet* n)
{
    for (Layer& l : n->L) {
        if (l.k == 4) {
            if (check_hip(hipMemcpy(l.dw, l.w.data(), l.w.size() * 4, hipMemcpyHostToDevice), "flownet upload")) return VSC_E_HIP;
        } else {
            // [cout][cin][3][3] -> [cout][3][3][cin_pad] with the channel map
            std::vector<float> p((size_t)l.cout * 9 * l.cin_pad, 0.0f);
            for (int co = 0; co < l.cout; ++co)
                for (int ci = 0; ci < l.cin; ++ci)
                    for (int t = 0; t < 9; ++t)
                        p[((size_t)co * 9 + t) * l.cin_pad + l.cmap[ci]] = l.w[((size_t)co * l.cin + ci) * 9 + t];
            if (n->ek == EK_F32) {
                if (check_hip(hipMemcpy(l.dw, p.data(), p.size() * 4, hipMemcpyHostToDevice), "flownet upload")) return VSC_E_HIP;
            } else {
                std::vector<uint16_t> h(p.size());
                for (size_t j = 0; j < p.size(); ++j) h[j] = f32_to_bf16(p[j]);
                if (check_hip(hipMemcpy(l.dw, h.data(), h.size() * 2, hipMemcpyHostToDevice), "flownet upload")) return VSC_E_HIP;
            }
        }
        if (check_hip(hipMemcpy(l.db, l.b.data(), l.b.size() * 4, hipMemcpyHostToDevice), "flownet upload")) return VSC_E_HIP;
    }
    return VSC_OK;
}

// device bytes, computed without allocating (sizes as in alloc_all)
struct Sizes {
    size_t img, tmp, feat_slot, cat0[7], dec[7], warped, flow[7], refined, rbuf, weights;
};

Sizes sizes_of(const vsc_flownet* n)
{
    Sizes z{};
    const size_t eb = n->eb;
    z.img = (size_t)n->frames * n->Hp * n->Wp * n->vec * eb;
    z.tmp = (size_t)n->frames * (n->Hp / 2) * (n->Wp / 2) * rup(16, n->vec) * eb;
    z.feat_slot = 0;
    for (int lvl = 3; lvl <= 6; ++lvl) z.feat_slot += (size_t)(n->Hp >> lvl) * (n->Wp >> lvl) * cv_of(n, lvl) * eb;
    for (int lvl = 3; lvl <= 6; ++lvl) {
        const size_t px = (size_t)n->pairs * (n->Hp >> lvl) * (n->Wp >> lvl);
        z.cat0[lvl] = px * ccat_of(n, lvl) * eb;
        z.dec[lvl] = px * kDecC * eb;
        z.flow[lvl] = px * 2 * 4;
    }
    const size_t px3 = (size_t)n->pairs * (n->Hp >> 3) * (n->Wp >> 3);
    z.warped = px3 * cv_of(n, 6) * eb;  // the widest level, sized at level 3's pixel count
    z.refined = px3 * 2 * 4;
    z.rbuf = px3 * 128 * eb;
    z.weights = 0;
    for (const Layer& l : n->L)
        z.weights += (l.k == 4 ? l.w.size() * 4 : (size_t)l.cout * 9 * l.cin_pad * eb) + l.b.size() * 4 + 512;
    z.weights += kPartialFloatsPerItem * (n->pairs + n->frames) * 4;
    return z;
}

size_t footprint_of(const vsc_flownet* n)
{
    const Sizes z = sizes_of(n);
    size_t t = z.img + 2 * z.tmp + z.feat_slot * n->slots + z.warped + z.refined + 2 * z.rbuf + z.weights;
    for (int lvl = 3; lvl <= 6; ++lvl) t += z.cat0[lvl] + z.dec[lvl] + 2 * z.flow[lvl];
    return t;
}

void* dalloc(vsc_flownet* n, size_t bytes, int& rc)
{
    if (rc != VSC_OK) return nullptr;
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
        rc = set_error(VSC_E_NOMEM, "vsc_flownet_create: device allocation failed");
        return nullptr;
    }
    if (hipMemset(p, 0, bytes ? bytes : 16) != hipSuccess) rc = set_error(VSC_E_HIP, "vsc_flownet_create: memset");
    n->owned.push_back(p);
    return p;
}

int alloc_all(vsc_flownet* n)
{
    int rc = VSC_OK;
    const Sizes z = sizes_of(n);
    n->img = dalloc(n, z.img, rc);
    n->tmp[0] = dalloc(n, z.tmp, rc);
    n->tmp[1] = dalloc(n, z.tmp, rc);
    for (int lvl = 3; lvl <= 6; ++lvl)
        n->feat[lvl] = dalloc(n, (size_t)n->slots * (n->Hp >> lvl) * (n->Wp >> lvl) * cv_of(n, lvl) * n->eb, rc);
    for (int lvl = 3; lvl <= 6; ++lvl) {
        n->cat0[lvl] = dalloc(n, z.cat0[lvl], rc);
        n->dec[lvl] = dalloc(n, z.dec[lvl], rc);
        n->flow[lvl] = static_cast<float*>(dalloc(n, z.flow[lvl], rc));
        n->upflow[lvl] = static_cast<float*>(dalloc(n, z.flow[lvl], rc));
    }
    n->warped = dalloc(n, z.warped, rc);
    n->refined = static_cast<float*>(dalloc(n, z.refined, rc));
    n->rbuf[0] = dalloc(n, z.rbuf, rc);
    n->rbuf[1] = dalloc(n, z.rbuf, rc);
    n->partial = static_cast<float*>(dalloc(n, kPartialFloatsPerItem * n->pairs * 4, rc));
    n->partial_x = static_cast<float*>(dalloc(n, kPartialFloatsPerItem * n->frames * 4, rc));
    for (Layer& l : n->L) {
        l.dw = dalloc(n, l.k == 4 ? l.w.size() * 4 : (size_t)l.cout * 9 * l.cin_pad * n->eb, rc);
        l.db = static_cast<float*>(dalloc(n, l.b.size() * 4, rc));
    }
    return rc;
}

void* feat_ptr(vsc_flownet* n, int slot, int lvl)
{
    return static_cast<char*>(n->feat[lvl]) + (size_t)slot * (n->Hp >> lvl) * (n->Wp >> lvl) * cv_of(n, lvl) * n->eb;
}

int conv(vsc_flownet* n, const Layer& l, Nhwc in, Nhwc out, Nhwc add, int ek_out, hipStream_t s, Nhwc out2 = Nhwc{},
         float* partial = nullptr)
{
    // the device weights are [cout][3][3][cin_pad]: the slices must match them exactly
    if (in.c != l.cin_pad || out.c != l.cout || (add.p && add.c != l.cout))
        return set_error(VSC_E_INVALID, "flownet: layer slice does not match its weights");
    ConvDesc d{};
    d.in = in; d.out = out; d.add = add; d.out2 = out2;
    d.wt = l.dw; d.bias = l.db;
    d.stride = l.stride; d.dil = l.dil; d.pad = l.dil; d.act = l.act;
    d.ek_in = n->ek; d.ek_out = ek_out;
    d.partial = partial ? partial : n->partial;
    d.partial_cap = kPartialFloatsPerItem;
    return launch_conv3x3(d, s);
}

Nhwc T(void* p, int N, int h, int w, int c, int cs, int c0) { return Nhwc{p, N, h, w, c, cs, c0}; }

// features of m frames into slots slot .. slot+m-1 (one batch through every
// extractor layer; each frame's features have the same bits as alone)
int extract(vsc_flownet* n, const uint8_t* const* rgba, int m, int slot, hipStream_t s)
{
    int rc;
    const size_t img_item = (size_t)n->Hp * n->Wp * n->vec * n->eb;
    for (int i = 0; i < m; ++i)
        if ((rc = launch_preprocess(n->ek, rgba[i], n->W, n->H, n->Wd, n->Hd, n->Wp, n->Hp,
                                    static_cast<char*>(n->img) + i * img_item, n->vec, s)))
            return rc;
    Nhwc x = T(n->img, m, n->Hp, n->Wp, n->vec, n->vec, 0);
    for (int k = 1; k <= 6; ++k) {
        const int h = n->Hp >> k, w = n->Wp >> k, cv = cv_of(n, k);
        const Layer* l = &n->L[n->ext0 + 3 * (k - 1)];
        // scratch order keeps each conv's input and output apart (level 1 ends in tmp[0], level 2 in tmp[1])
        void* a = (k == 2) ? n->tmp[1] : n->tmp[0];
        void* b = (k == 2) ? n->tmp[0] : n->tmp[1];
        void* c = k >= 3 ? feat_ptr(n, slot, k) : (k == 1 ? n->tmp[0] : n->tmp[1]);
        if (k == 1) { a = n->tmp[0]; b = n->tmp[1]; }
        // outputs carry the layer's channel count; as inputs the slices include
        // the zero padding channels up to the 16-byte multiple
        const int co = kFeatC[k];
        float* px = n->partial_x;
        if ((rc = conv(n, l[0], x, T(a, m, h, w, co, cv, 0), Nhwc{}, n->ek, s, Nhwc{}, px))) return rc;
        if ((rc = conv(n, l[1], T(a, m, h, w, cv, cv, 0), T(b, m, h, w, co, cv, 0), Nhwc{}, n->ek, s, Nhwc{}, px)))
            return rc;
        if ((rc = conv(n, l[2], T(b, m, h, w, cv, cv, 0), T(c, m, h, w, co, cv, 0), Nhwc{}, n->ek, s, Nhwc{}, px)))
            return rc;
        x = T(c, m, h, w, cv, cv, 0);
    }
    return VSC_OK;
}

int decode(vsc_flownet* n, int np, const int* first, const int* second, hipStream_t s)
{
    int rc;
    const int ek = n->ek, v = n->vec;
    for (int lvl = 6; lvl >= 3; --lvl) {
        const int h = n->Hp >> lvl, w = n->Wp >> lvl, C = kFeatC[lvl], CV = cv_of(n, lvl), cc = ccat_of(n, lvl);
        const long long px = (long long)h * w;
        const void* f1[FLOWNET_MAX_PAIRS] = {};
        const void* f2[FLOWNET_MAX_PAIRS] = {};
        for (int i = 0; i < np; ++i) { f1[i] = feat_ptr(n, first[i], lvl); f2[i] = feat_ptr(n, second[i], lvl); }
        void* cat = n->cat0[lvl];
        if (lvl == 6) {
            // level 6 has no upflow: its channels of cat0 stay as zeroed at create
            if ((rc = launch_corr_nhwc(ek, f1, f2, CV, CV, CV, np, h, w, cat, cc, 88, s, n->corr_bf16))) return rc;
        } else {
            // upflow of the coarser level's flow, computed inside the warp of f2 by it
            const Layer& up = n->L[n->up0 + (5 - lvl)];
            const UpflowArgs ua{n->flow[lvl + 1], static_cast<const float*>(up.dw), up.db, n->upflow[lvl], cat, cc, 88 + CV};
            if ((rc = launch_warp_nhwc(ek, f2, CV, nullptr, 2, kDbl[lvl], n->warped, CV, CV, np, h, w, s, &ua))) return rc;
            const void* wp[FLOWNET_MAX_PAIRS] = {};
            for (int i = 0; i < np; ++i) wp[i] = static_cast<const char*>(n->warped) + i * px * CV * n->eb;
            if ((rc = launch_corr_nhwc(ek, f1, wp, CV, CV, CV, np, h, w, cat, cc, 88, s, n->corr_bf16))) return rc;
        }
        // cat0 = [leaky(corr) | pad | f1 | upflow | pad]: the cost volume kernel wrote f1 at channel 88
        const Layer* l = &n->L[n->dec0 + 6 * (6 - lvl)];
        void* d = n->dec[lvl];
        if ((rc = conv(n, l[0], T(cat, np, h, w, cc, cc, 0), T(d, np, h, w, 128, kDecC, 0), Nhwc{}, ek, s))) return rc;
        for (int j = 1; j < 5; ++j)
            if ((rc = conv(n, l[j], T(d, np, h, w, kDecInC[j], kDecC, kDecIn0[j]),
                           T(d, np, h, w, kDecOut[j], kDecC, kDecOut0[j]), Nhwc{}, ek, s)))
                return rc;
        // level 3: the flow also goes into dec (channel 448) as the refiner's input
        if ((rc = conv(n, l[5], T(d, np, h, w, kDecInC[5], kDecC, kDecIn0[5]), T(n->flow[lvl], np, h, w, 2, 2, 0),
                       Nhwc{}, EK_F32, s, lvl == 3 ? T(d, np, h, w, 2, kDecC, 448) : Nhwc{})))
            return rc;
        (void)C;
    }
    // refiner on cat([x5, flow]) of level 3; flow + refinement (pwcnet.py:981-986)
    const int h = n->Hp >> 3, w = n->Wp >> 3;
    void* d = n->dec[3];
    const Layer* r = &n->L[n->ref0];
    Nhwc x = T(d, np, h, w, kRefInC, kDecC, kRefIn0);
    for (int j = 0; j < 6; ++j) {
        Nhwc y = T(n->rbuf[j & 1], np, h, w, kRefOut[j], 128, 0);
        if ((rc = conv(n, r[j], x, y, Nhwc{}, ek, s))) return rc;
        x = T(n->rbuf[j & 1], np, h, w, rup(kRefOut[j], v), 128, 0);
    }
    return conv(n, r[6], x, T(n->refined, np, h, w, 2, 2, 0), T(n->flow[3], np, h, w, 2, 2, 0), EK_F32, s);
}

}  // namespace
}  // namespace vsc

using namespace vsc;

static int flownet_dims(int W, int H, int down, int& Wd, int& Hd, int& Wp, int& Hp)
{
    if (W <= 0 || H <= 0 || (down != 1 && down != 2 && down != 4))
        return set_error(VSC_E_INVALID, "vsc_flownet: frame size must be positive and downscale 1, 2 or 4");
    Wd = (int)std::lround((double)W / down);
    Hd = (int)std::lround((double)H / down);
    Wp = (Wd + 63) / 64 * 64;
    Hp = (Hd + 63) / 64 * 64;
    return VSC_OK;
}

static vsc_flownet* flownet_shell(int W, int H, int down, int dtype, int slots, int pairs, int& rc)
{
    const int frames = (pairs + 1) / 2;
    int Wd, Hd, Wp, Hp;
    rc = flownet_dims(W, H, down, Wd, Hd, Wp, Hp);
    if (rc) return nullptr;
    if (dtype != VSC_F32 && dtype != VSC_BF16) {
        rc = set_error(VSC_E_DTYPE, "vsc_flownet: dtype must be VSC_F32 or VSC_BF16");
        return nullptr;
    }
    if (slots < 2 || slots > 64) {
        rc = set_error(VSC_E_INVALID, "vsc_flownet: slots must be in [2, 64]");
        return nullptr;
    }
    if (pairs < 1 || pairs > FLOWNET_MAX_PAIRS) {
        rc = set_error(VSC_E_INVALID, "vsc_flownet: pair capacity must be in [1, 8]");
        return nullptr;
    }
    vsc_flownet* n = new (std::nothrow) vsc_flownet();
    if (!n) { rc = set_error(VSC_E_NOMEM, "vsc_flownet: out of host memory"); return nullptr; }
    n->W = W; n->H = H; n->down = down; n->Wd = Wd; n->Hd = Hd; n->Wp = Wp; n->Hp = Hp;
    n->ek = elem_kind(dtype); n->eb = elem_bytes(n->ek); n->vec = n->ek == EK_F32 ? 4 : 8; n->slots = slots;
    n->pairs = pairs;
    n->frames = frames;
    build_layers(n);
    return n;
}

namespace vsc {

size_t flownet_footprint_pairs(int width, int height, int downscale, int dtype, int slots, int pairs)
{
    int rc;
    vsc_flownet* n = flownet_shell(width, height, downscale, dtype, slots, pairs, rc);
    if (!n) return 0;
    for (Layer& l : n->L) { l.w.resize((size_t)l.cin * l.cout * l.k * l.k); l.b.resize(l.cout); }
    const size_t f = footprint_of(n);
    delete n;
    return f;
}

int flownet_create_pairs(int width, int height, int downscale, int dtype, unsigned long long seed, int slots,
                         int pairs, vsc_flownet** out)
{
    if (!out) return set_error(VSC_E_INVALID, "vsc_flownet_create: out is NULL");
    *out = nullptr;
    int rc;
    vsc_flownet* n = flownet_shell(width, height, downscale, dtype, slots, pairs, rc);
    if (!n) return rc;
    init_params(n, seed);
    if ((rc = alloc_all(n)) || (rc = upload_params(n))) {
        vsc_flownet_destroy(n);
        return rc;
    }
    *out = n;
    return VSC_OK;
}

}  // namespace vsc

extern "C" size_t vsc_flownet_footprint(int width, int height, int downscale, int dtype, int slots)
{
    return flownet_footprint_pairs(width, height, downscale, dtype, slots, 2);
}

extern "C" int vsc_flownet_create(int width, int height, int downscale, int dtype, unsigned long long seed, int slots,
                                  vsc_flownet** out)
{
    return flownet_create_pairs(width, height, downscale, dtype, seed, slots, 2, out);
}

extern "C" int vsc_flownet_destroy(vsc_flownet* n)
{
    if (!n) return VSC_OK;
    for (void* p : n->owned) (void)hipFree(p);
    delete n;
    return VSC_OK;
}

extern "C" int vsc_flownet_flow_size(const vsc_flownet* n, int* width, int* height)
{
    if (!n || !width || !height) return set_error(VSC_E_INVALID, "vsc_flownet_flow_size: NULL argument");
    *width = n->Wd;
    *height = n->Hd;
    return VSC_OK;
}

extern "C" size_t vsc_flownet_param_count(const vsc_flownet* n) { return n ? param_count_of(n) : 0; }

extern "C" int vsc_flownet_get_params(const vsc_flownet* n, float* host, size_t count)
{
    if (!n || !host) return set_error(VSC_E_INVALID, "vsc_flownet_get_params: NULL argument");
    if (count != param_count_of(n)) return set_error(VSC_E_INVALID, "vsc_flownet_get_params: wrong count");
    for (const Layer& l : n->L) {
        for (float x : l.w) *host++ = x;
        for (float x : l.b) *host++ = x;
    }
    return VSC_OK;
}

extern "C" int vsc_flownet_set_params(vsc_flownet* n, const float* host, size_t count)
{
    if (!n || !host) return set_error(VSC_E_INVALID, "vsc_flownet_set_params: NULL argument");
    if (count != param_count_of(n)) return set_error(VSC_E_INVALID, "vsc_flownet_set_params: wrong count");
    for (Layer& l : n->L) {
        for (float& x : l.w) x = *host++;
        for (float& x : l.b) x = *host++;
    }
    if (hipDeviceSynchronize() != hipSuccess) return set_error(VSC_E_HIP, "vsc_flownet_set_params: sync");
    return upload_params(n);
}

extern "C" int vsc_flownet_scale_flow_heads(vsc_flownet* n, float factor)
{
    if (!n || !(factor >= 0.0f)) return set_error(VSC_E_INVALID, "vsc_flownet_scale_flow_heads: bad argument");
    // the layers whose outputs are flows: each decoder's last conv and the refiner's last conv
    for (int lvl = 6; lvl >= 3; --lvl) {
        Layer& l = n->L[n->dec0 + 6 * (6 - lvl) + 5];
        for (float& x : l.w) x *= factor;
        for (float& x : l.b) x *= factor;
    }
    Layer& r = n->L[n->ref0 + 6];
    for (float& x : r.w) x *= factor;
    for (float& x : r.b) x *= factor;
    if (hipDeviceSynchronize() != hipSuccess) return set_error(VSC_E_HIP, "vsc_flownet_scale_flow_heads: sync");
    return upload_params(n);
}

namespace vsc {

int flownet_extract_frames(vsc_flownet* n, const uint8_t* const* rgba, int m, int slot, hipStream_t s)
{
    if (!n || !rgba) return set_error(VSC_E_INVALID, "vsc_flownet_extract: NULL argument");
    if (m < 1 || m > n->frames) return set_error(VSC_E_INVALID, "vsc_flownet_extract: batch beyond the network's");
    if (slot < 0 || slot + m > n->slots) return set_error(VSC_E_INVALID, "vsc_flownet_extract: slot out of range");
    for (int i = 0; i < m; ++i)
        if (!rgba[i]) return set_error(VSC_E_INVALID, "vsc_flownet_extract: NULL argument");
    return extract(n, rgba, m, slot, s);
}

}  // namespace vsc

extern "C" int vsc_flownet_set_corr_dtype(vsc_flownet* n, int dtype)
{
    if (!n) return set_error(VSC_E_INVALID, "vsc_flownet_set_corr_dtype: NULL network");
    if (dtype != VSC_F32 && dtype != VSC_BF16)
        return set_error(VSC_E_DTYPE, "vsc_flownet_set_corr_dtype: dtype must be VSC_F32 or VSC_BF16");
    if (n->ek == EK_BF16 && dtype != VSC_BF16)
        return set_error(VSC_E_DTYPE, "vsc_flownet_set_corr_dtype: a BF16 network has a BF16 cost volume");
    n->corr_bf16 = n->ek == EK_F32 && dtype == VSC_BF16;
    return VSC_OK;
}

extern "C" int vsc_flownet_extract(vsc_flownet* n, const uint8_t* rgba, int slot, vsc_hip_stream stream)
{
    return flownet_extract_frames(n, &rgba, 1, slot, as_stream(stream));
}

extern "C" int vsc_flownet_flow(vsc_flownet* n, int npairs, const int* first, const int* second, float* const* out,
                                int out_channels, vsc_hip_stream stream)
{
    if (!n || !first || !second || !out) return set_error(VSC_E_INVALID, "vsc_flownet_flow: NULL argument");
    if (npairs < 1 || npairs > n->pairs)
        return set_error(VSC_E_INVALID, "vsc_flownet_flow: npairs must be in [1, the network's pair capacity]");
    if (out_channels != 2 && out_channels != 3) return set_error(VSC_E_INVALID, "vsc_flownet_flow: out_channels 2|3");
    for (int i = 0; i < npairs; ++i) {
        if (first[i] < 0 || first[i] >= n->slots || second[i] < 0 || second[i] >= n->slots)
            return set_error(VSC_E_INVALID, "vsc_flownet_flow: slot out of range");
        if (!out[i]) return set_error(VSC_E_INVALID, "vsc_flownet_flow: NULL output");
    }
    hipStream_t s = as_stream(stream);
    int rc = decode(n, npairs, first, second, s);
    if (rc) return rc;
    const int h = n->Hp >> 3, w = n->Wp >> 3;
    if ((rc = launch_flow_final(n->refined, npairs, h, w, n->Hp, n->Wp, out, n->Wd, n->Hd, out_channels, 0, s)))
        return rc;
    n->last_pairs = npairs;
    return VSC_OK;
}

extern "C" int vsc_flownet_intermediate(const vsc_flownet* n, int which, int level, int item, float* dst,
                                        vsc_hip_stream stream)
{
    if (!n || !dst) return set_error(VSC_E_INVALID, "vsc_flownet_intermediate: NULL argument");
    if (item < 0 || item >= n->last_pairs) return set_error(VSC_E_INVALID, "vsc_flownet_intermediate: bad item");
    const float* src = nullptr;
    if (which == 0 && level >= 3 && level <= 5) src = n->upflow[level];
    else if (which == 1 && level >= 3 && level <= 6) src = n->flow[level];
    else if (which == 2 && level == 3) src = n->refined;
    else return set_error(VSC_E_INVALID, "vsc_flownet_intermediate: bad which/level");
    const size_t px = (size_t)(n->Hp >> level) * (n->Wp >> level);
    return check_hip(hipMemcpyAsync(dst, src + px * 2 * item, px * 2 * 4, hipMemcpyDeviceToDevice, as_stream(stream)),
                     "vsc_flownet_intermediate");
}
