// Middlebury .flo files (ReadFlowFile, flowIO.cpp:11-20,31-78): the float tag
// 202021.25, int32 width and height in [1, 99999], then width*height
// interleaved (u, v) float32 pairs, row-major, little-endian, nothing after.
// Host-side file IO for the flow-file mode of the stabilizer
// (FileStabilizer::retrieveOpticalFlow, stabilizefiles.cpp:134-149).
#include "vsc_internal.h"

#include <cstdio>
#include <cstring>
#include <string>

namespace {

constexpr float kFloTag = 202021.25f;

struct File {
    FILE* f;
    explicit File(FILE* p) : f(p) {}
    ~File() { if (f) std::fclose(f); }
};

}  // namespace

extern "C" int vsc_flo_read(const char* path, float* flow, size_t capacity, int* width, int* height)
{
    using vsc::set_error;
    if (!path || !width || !height) return set_error(VSC_E_INVALID, "vsc_flo_read: NULL argument");
    File fh(std::fopen(path, "rb"));
    if (!fh.f) return set_error(VSC_E_INVALID, std::string("vsc_flo_read: cannot open ") + path);
    float tag = 0.0f;
    int32_t wh[2] = {0, 0};
    if (std::fread(&tag, sizeof(float), 1, fh.f) != 1 || std::fread(wh, sizeof(int32_t), 2, fh.f) != 2)
        return set_error(VSC_E_INVALID, "vsc_flo_read: file too short for the header");
    if (tag != kFloTag) return set_error(VSC_E_INVALID, "vsc_flo_read: wrong tag (not a .flo file)");
    if (wh[0] < 1 || wh[0] > 99999) return set_error(VSC_E_INVALID, "vsc_flo_read: illegal width");
    if (wh[1] < 1 || wh[1] > 99999) return set_error(VSC_E_INVALID, "vsc_flo_read: illegal height");
    *width = wh[0];
    *height = wh[1];
    if (!flow) return VSC_OK;
    const size_t n = (size_t)wh[0] * wh[1] * 2;
    if (capacity < n) return set_error(VSC_E_INVALID, "vsc_flo_read: destination too small");
    if (std::fread(flow, sizeof(float), n, fh.f) != n) return set_error(VSC_E_INVALID, "vsc_flo_read: file too short");
    if (std::fgetc(fh.f) != EOF) return set_error(VSC_E_INVALID, "vsc_flo_read: file too long");
    return VSC_OK;
}

extern "C" int vsc_flo_write(const char* path, const float* flow, int width, int height)
{
    using vsc::set_error;
    if (!path || !flow) return set_error(VSC_E_INVALID, "vsc_flo_write: NULL argument");
    if (width < 1 || width > 99999 || height < 1 || height > 99999)
        return set_error(VSC_E_INVALID, "vsc_flo_write: width and height must be in [1, 99999]");
    File fh(std::fopen(path, "wb"));
    if (!fh.f) return set_error(VSC_E_INVALID, std::string("vsc_flo_write: cannot create ") + path);
    const int32_t wh[2] = {width, height};
    const size_t n = (size_t)width * height * 2;
    if (std::fwrite(&kFloTag, sizeof(float), 1, fh.f) != 1 || std::fwrite(wh, sizeof(int32_t), 2, fh.f) != 2 ||
        std::fwrite(flow, sizeof(float), n, fh.f) != n)
        return set_error(VSC_E_INVALID, "vsc_flo_write: write failed");
    return VSC_OK;
}
