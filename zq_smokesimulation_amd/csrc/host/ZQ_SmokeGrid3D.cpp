// ZQ_SmokeGrid3D (reference: ZQ_SmokeGrid3D.cpp:5-131)
#include "ZQ_SmokeGrid3D.h"

#include <cstdio>
#include <cstring>
#include <strings.h>

namespace {
template <class T>
T* alloc_zero(size_t n) {
    T* p = new T[n];
    std::memset(p, 0, sizeof(T) * n);
    return p;
}
template <class T>
void release(T*& p) {
    delete[] p;
    p = nullptr;
}
bool ends_with_ci(const char* s, const char* suffix) {
    const size_t n = std::strlen(s), m = std::strlen(suffix);
    return n >= m && strcasecmp(s + n - m, suffix) == 0;
}
}  // namespace

ZQ_SmokeGrid3D::ZQ_SmokeGrid3D(const unsigned int width_, const unsigned int height_, const unsigned int depth_)
    : width(width_), height(height_), depth(depth_),
      unOccupyRatioU(nullptr), unOccupyRatioV(nullptr), unOccupyRatioW(nullptr), unOccupyVolume(nullptr),
      c_occupy(nullptr), c_pressure(nullptr), c_temperature(nullptr), c_density(nullptr), c_u(nullptr),
      c_v(nullptr), c_w(nullptr), c_ratioU(nullptr), c_ratioV(nullptr), c_ratioW(nullptr), c_volume(nullptr) {
    const size_t N = (size_t)width * height * depth;
    occupy = alloc_zero<bool>(N);
    pressure = alloc_zero<float>(N);
    temperature = alloc_zero<float>(N);
    density = alloc_zero<float>(N);
    u = alloc_zero<float>((size_t)(width + 1) * height * depth);
    v = alloc_zero<float>((size_t)width * (height + 1) * depth);
    w = alloc_zero<float>((size_t)width * height * (depth + 1));
}

ZQ_SmokeGrid3D::~ZQ_SmokeGrid3D() {
    release(occupy);
    release(pressure);
    release(temperature);
    release(density);
    release(u);
    release(v);
    release(w);
    ClearFaceUnOccupyRatio();
    ClearVolumeUnOccupyRatio();
}

void ZQ_SmokeGrid3D::BuildFaceUnOccupyRatio() {
    ClearFaceUnOccupyRatio();
    unOccupyRatioU = alloc_zero<float>((size_t)(width + 1) * height * depth);
    unOccupyRatioV = alloc_zero<float>((size_t)width * (height + 1) * depth);
    unOccupyRatioW = alloc_zero<float>((size_t)width * height * (depth + 1));
}

void ZQ_SmokeGrid3D::ClearFaceUnOccupyRatio() {
    release(unOccupyRatioU);
    release(unOccupyRatioV);
    release(unOccupyRatioW);
}

void ZQ_SmokeGrid3D::BuildVolumeUnOccupyRatio() {
    ClearVolumeUnOccupyRatio();
    unOccupyVolume = alloc_zero<float>((size_t)width * height * depth);
}

void ZQ_SmokeGrid3D::ClearVolumeUnOccupyRatio() { release(unOccupyVolume); }

ZQ_SmokeGrid3D* ZQ_SmokeGrid3D::Clone() const {
    ZQ_SmokeGrid3D* t = new ZQ_SmokeGrid3D(width, height, depth);
    const size_t N = (size_t)width * height * depth;
    const size_t Nu = (size_t)(width + 1) * height * depth, Nv = (size_t)width * (height + 1) * depth,
                 Nw = (size_t)width * height * (depth + 1);
    std::memcpy(t->occupy, occupy, N * sizeof(bool));
    std::memcpy(t->pressure, pressure, N * 4);
    std::memcpy(t->temperature, temperature, N * 4);
    std::memcpy(t->density, density, N * 4);
    std::memcpy(t->u, u, Nu * 4);
    std::memcpy(t->v, v, Nv * 4);
    std::memcpy(t->w, w, Nw * 4);
    if (unOccupyRatioU) {
        t->BuildFaceUnOccupyRatio();
        std::memcpy(t->unOccupyRatioU, unOccupyRatioU, Nu * 4);
        std::memcpy(t->unOccupyRatioV, unOccupyRatioV, Nv * 4);
        std::memcpy(t->unOccupyRatioW, unOccupyRatioW, Nw * 4);
    }
    if (unOccupyVolume) {
        t->BuildVolumeUnOccupyRatio();
        std::memcpy(t->unOccupyVolume, unOccupyVolume, N * 4);
    }
    return t;
}

// .di3: 16-byte type tag "float", int32 width, height, depth, channels, one flag byte, then the
// FP32 samples x fastest (the container ZQlib writes; layout read off the .di3 files under
// ZQ_OpticalFlow3D/par3D1 in the reference tree).
bool ZQ_SmokeGrid3D::ExportDensity(const char* file, float quality) const {
    (void)quality;
    if (!file || !ends_with_ci(file, ".di3")) return false;
    FILE* f = std::fopen(file, "wb");
    if (!f) return false;
    char tag[16] = {0};
    std::strcpy(tag, "float");
    const int hdr[4] = {(int)width, (int)height, (int)depth, 1};
    const unsigned char flag = 0;
    const size_t N = (size_t)width * height * depth;
    bool ok = std::fwrite(tag, 1, 16, f) == 16 && std::fwrite(hdr, 4, 4, f) == 4 && std::fwrite(&flag, 1, 1, f) == 1 &&
              std::fwrite(density, 4, N, f) == N;
    ok = (std::fclose(f) == 0) && ok;
    return ok;
}

// raw u, v, w, density, temperature (ZQ_SmokeGrid3D.cpp:83-121), the .guidedat layout
bool ZQ_SmokeGrid3D::ExportGuide(const char* file) const {
    if (!file) return false;
    FILE* f = std::fopen(file, "wb");
    if (!f) return false;
    const size_t N = (size_t)width * height * depth;
    const size_t Nu = (size_t)(width + 1) * height * depth, Nv = (size_t)width * (height + 1) * depth,
                 Nw = (size_t)width * height * (depth + 1);
    bool ok = std::fwrite(u, 4, Nu, f) == Nu && std::fwrite(v, 4, Nv, f) == Nv && std::fwrite(w, 4, Nw, f) == Nw &&
              std::fwrite(density, 4, N, f) == N && std::fwrite(temperature, 4, N, f) == N;
    ok = (std::fclose(f) == 0) && ok;
    return ok;
}

// the reference's occupancy/velocity exporters write nothing and report success (ZQ_SmokeGrid3D.cpp:123-131)
bool ZQ_SmokeGrid3D::ExportOccupy(const char* file) const {
    (void)file;
    return true;
}

bool ZQ_SmokeGrid3D::ExportVelocity(const char* file) const {
    (void)file;
    return true;
}
