// ZQ_SmokeGrid3D -- field container with the reference's interface (ZQ_SmokeGrid3D.h:7-70).
// Storage is plain zero-initialised host arrays (the reference holds ZQlib ZQ_DImage3D images);
// the getters return references to the array pointers exactly as the reference's do.
#pragma once

#include <cstddef>

class ZQ_SmokeGrid3D {
public:
    ZQ_SmokeGrid3D(const unsigned int width, const unsigned int height, const unsigned int depth);
    ~ZQ_SmokeGrid3D();
    ZQ_SmokeGrid3D(const ZQ_SmokeGrid3D&) = delete;
    ZQ_SmokeGrid3D& operator=(const ZQ_SmokeGrid3D&) = delete;

private:
    unsigned int width, height, depth;
    bool* occupy;
    float* pressure;
    float* temperature;
    float* density;
    float* u;
    float* v;
    float* w;
    float* unOccupyRatioU;
    float* unOccupyRatioV;
    float* unOccupyRatioW;
    float* unOccupyVolume;
    // const views handed out by the const getters (same storage)
    mutable const bool* c_occupy;
    mutable const float *c_pressure, *c_temperature, *c_density, *c_u, *c_v, *c_w;
    mutable const float *c_ratioU, *c_ratioV, *c_ratioW, *c_volume;

public:
    int GetWidth() const { return width; }
    int GetHeight() const { return height; }
    int GetDepth() const { return depth; }

    bool*& GetOccupyPtr() { return occupy; }
    const bool*& GetOccupyPtr() const { return c_occupy = occupy; }
    float*& GetPressurePtr() { return pressure; }
    const float*& GetPressurePtr() const { return c_pressure = pressure; }
    float*& GetTemperaturePtr() { return temperature; }
    const float*& GetTemperaturePtr() const { return c_temperature = temperature; }
    float*& GetDensityPtr() { return density; }
    const float*& GetDensityPtr() const { return c_density = density; }
    float*& GetVelocityUptr() { return u; }
    const float*& GetVelocityUptr() const { return c_u = u; }
    float*& GetVelocityVptr() { return v; }
    const float*& GetVelocityVptr() const { return c_v = v; }
    float*& GetVelocityWptr() { return w; }
    const float*& GetVelocityWptr() const { return c_w = w; }
    float*& GetFaceUnOccupyRatioUPtr() { return unOccupyRatioU; }
    const float*& GetFaceUnOccupyRatioUPtr() const { return c_ratioU = unOccupyRatioU; }
    float*& GetFaceUnOccupyRatioVPtr() { return unOccupyRatioV; }
    const float*& GetFaceUnOccupyRatioVPtr() const { return c_ratioV = unOccupyRatioV; }
    float*& GetFaceUnOccupyRatioWPtr() { return unOccupyRatioW; }
    const float*& GetFaceUnOccupyRatioWPtr() const { return c_ratioW = unOccupyRatioW; }
    float*& GetVolumeUnOccupyRatioPtr() { return unOccupyVolume; }
    const float*& GetVolumeUnOccupyRatioPtr() const { return c_volume = unOccupyVolume; }

    void BuildFaceUnOccupyRatio();
    void ClearFaceUnOccupyRatio();
    void BuildVolumeUnOccupyRatio();
    void ClearVolumeUnOccupyRatio();

    ZQ_SmokeGrid3D* Clone() const;

    // .di3 is written (ZQ_SmokeGrid3D.cpp:55-81); the wavelet/compressed exporters of the reference
    // (.zqwv/.zqci, ZQlib) are not built and return false.
    bool ExportDensity(const char* file, float quality = 0.9999) const;
    bool ExportOccupy(const char* file) const;
    bool ExportVelocity(const char* file) const;
    bool ExportGuide(const char* file) const;
};
