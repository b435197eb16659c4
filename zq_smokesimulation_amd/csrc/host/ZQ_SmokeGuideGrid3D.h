// ZQ_SmokeGuideGrid3D -- coarse guide fields with the reference's interface
// (ZQ_SmokeGuideGrid3D.h:7-32): u, v, w (MAC faces), density and temperature, read from a .guidedat
// file (raw FP32 arrays in that order, the layout ZQ_SmokeGrid3D::ExportGuide writes).
#pragma once

#include <vector>

class ZQ_SmokeGuideGrid3D {
public:
    ZQ_SmokeGuideGrid3D(const unsigned int width, const unsigned int height, const unsigned int depth);
    ~ZQ_SmokeGuideGrid3D();

private:
    unsigned int width, height, depth;
    std::vector<float> u, v, w, density, temperature;
    mutable const float *c_u, *c_v, *c_w, *c_density, *c_temperature;

public:
    unsigned int GetWidth() const { return width; }
    unsigned int GetHeight() const { return height; }
    unsigned int GetDepth() const { return depth; }

    // ZQ_SmokeGuideGrid3D.cpp:20-56
    bool LoadFromFile(const char* file);
    const float*& GetUptr() const { return c_u = u.data(); }
    const float*& GetVptr() const { return c_v = v.data(); }
    const float*& GetWptr() const { return c_w = w.data(); }
    const float*& GetDensityPtr() const { return c_density = density.data(); }
    const float*& GetTemperaturePtr() const { return c_temperature = temperature.data(); }
};
