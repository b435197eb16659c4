// ZQ_SmokeGuideGrid3D (reference: ZQ_SmokeGuideGrid3D.cpp:3-56).
#include "ZQ_SmokeGuideGrid3D.h"

#include <cstdio>

ZQ_SmokeGuideGrid3D::ZQ_SmokeGuideGrid3D(const unsigned int width_, const unsigned int height_,
                                         const unsigned int depth_)
    : width(width_),
      height(height_),
      depth(depth_),
      u((size_t)(width_ + 1) * height_ * depth_, 0.0f),
      v((size_t)width_ * (height_ + 1) * depth_, 0.0f),
      w((size_t)width_ * height_ * (depth_ + 1), 0.0f),
      density((size_t)width_ * height_ * depth_, 0.0f),
      temperature((size_t)width_ * height_ * depth_, 0.0f),
      c_u(nullptr),
      c_v(nullptr),
      c_w(nullptr),
      c_density(nullptr),
      c_temperature(nullptr) {}

ZQ_SmokeGuideGrid3D::~ZQ_SmokeGuideGrid3D() {}

// u, v, w, density, temperature as raw FP32 arrays; false when the file is missing or short
bool ZQ_SmokeGuideGrid3D::LoadFromFile(const char* file) {
    if (!file) return false;
    FILE* in = std::fopen(file, "rb");
    if (!in) return false;
    bool ok = true;
    for (std::vector<float>* a : {&u, &v, &w, &density, &temperature})
        if (ok && std::fread(a->data(), sizeof(float), a->size(), in) != a->size()) ok = false;
    std::fclose(in);
    return ok;
}
