// zq3d_device.h -- device-side helpers shared by the gfx950 kernels.
//
// Layout is the reference's (ZQ_SmokeGrid3D.cpp:5-18), x fastest:
//   cell (i,j,k)   : (k*H + j)*W + i
//   u face (i,j,k) : (k*H + j)*(W+1) + i      i in [0,W]
//   v face (i,j,k) : (k*(H+1) + j)*W + i      j in [0,H]
//   w face (i,j,k) : (k*H + j)*W + i          k in [0,D]
//   padded pressure: ((k+1)*(H+2) + (j+1))*(W+2) + (i+1), ghost layer held at 0.
// Every kernel maps x to the 64 lanes of a wavefront, so each global access of a
// row is one coalesced 256-byte transaction per wave.
//
// Arithmetic is written operation by operation in a fixed order and the library
// is compiled with -ffp-contract=off, so results are reproducible bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zq3d {

struct Dims {
    int W, H, D;
};

__device__ __forceinline__ int cidx(const Dims& d, int i, int j, int k) { return (k * d.H + j) * d.W + i; }
__device__ __forceinline__ int uidx(const Dims& d, int i, int j, int k) { return (k * d.H + j) * (d.W + 1) + i; }
__device__ __forceinline__ int vidx(const Dims& d, int i, int j, int k) { return (k * (d.H + 1) + j) * d.W + i; }
__device__ __forceinline__ int widx(const Dims& d, int i, int j, int k) { return (k * d.H + j) * d.W + i; }
__device__ __forceinline__ int pidx(const Dims& d, int i, int j, int k) {
    return ((k + 1) * (d.H + 2) + (j + 1)) * (d.W + 2) + (i + 1);
}

__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// One axis of a clamp-to-edge linear filter in texel space: texels i0, i1 and weight a.
// The coordinate is first limited to [-1, n], which leaves the result unchanged and keeps the
// float->int conversion defined for any input (also NaN, which fminf/fmaxf map to a bound).
__device__ __forceinline__ void axis_setup(float x, int n, int& i0, int& i1, float& a) {
    x = fminf(fmaxf(x, -1.0f), (float)n);
    float fx = floorf(x);
    int k = (int)fx;
    a = x - fx;
    i1 = clampi(k + 1, 0, n - 1);
    i0 = clampi(k, 0, n - 1);
}

// Trilinear sample of a scalar grid nx*ny*nz at texel coordinates (x,y,z).
__device__ __forceinline__ float sample3(const float* __restrict__ g, int nx, int ny, int nz, float x, float y,
                                         float z) {
    int i0, i1, j0, j1, k0, k1;
    float ax, ay, az;
    axis_setup(x, nx, i0, i1, ax);
    axis_setup(y, ny, j0, j1, ay);
    axis_setup(z, nz, k0, k1, az);
    const int r00 = (k0 * ny + j0) * nx, r10 = (k0 * ny + j1) * nx;
    const int r01 = (k1 * ny + j0) * nx, r11 = (k1 * ny + j1) * nx;
    float c00 = lerpf(g[r00 + i0], g[r00 + i1], ax);
    float c10 = lerpf(g[r10 + i0], g[r10 + i1], ax);
    float c01 = lerpf(g[r01 + i0], g[r01 + i1], ax);
    float c11 = lerpf(g[r11 + i0], g[r11 + i1], ax);
    float c0 = lerpf(c00, c10, ay);
    float c1 = lerpf(c01, c11, ay);
    return lerpf(c0, c1, az);
}

// Trilinear sample of the first three components of a float4 cell grid (one 16-byte load per tap).
__device__ __forceinline__ float3 sample4(const float4* __restrict__ g, int nx, int ny, int nz, float x, float y,
                                          float z) {
    int i0, i1, j0, j1, k0, k1;
    float ax, ay, az;
    axis_setup(x, nx, i0, i1, ax);
    axis_setup(y, ny, j0, j1, ay);
    axis_setup(z, nz, k0, k1, az);
    const int r00 = (k0 * ny + j0) * nx, r10 = (k0 * ny + j1) * nx;
    const int r01 = (k1 * ny + j0) * nx, r11 = (k1 * ny + j1) * nx;
    const float4 t000 = g[r00 + i0], t100 = g[r00 + i1], t010 = g[r10 + i0], t110 = g[r10 + i1];
    const float4 t001 = g[r01 + i0], t101 = g[r01 + i1], t011 = g[r11 + i0], t111 = g[r11 + i1];
    float3 out;
    {
        float c00 = lerpf(t000.x, t100.x, ax), c10 = lerpf(t010.x, t110.x, ax);
        float c01 = lerpf(t001.x, t101.x, ax), c11 = lerpf(t011.x, t111.x, ax);
        out.x = lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
    }
    {
        float c00 = lerpf(t000.y, t100.y, ax), c10 = lerpf(t010.y, t110.y, ax);
        float c01 = lerpf(t001.y, t101.y, ax), c11 = lerpf(t011.y, t111.y, ax);
        out.y = lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
    }
    {
        float c00 = lerpf(t000.z, t100.z, ax), c10 = lerpf(t010.z, t110.z, ax);
        float c01 = lerpf(t001.z, t101.z, ax), c11 = lerpf(t011.z, t111.z, ax);
        out.z = lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az);
    }
    return out;
}

// Velocity at a position for the backtrace: cell-centred float4 grid sampled at pos-0.5
// (the reference's normalized coordinate pos/N with a linear filter), or the three staggered
// face grids (u at (x, y-0.5, z-0.5) on (W+1)*H*D, ...).
struct RegularVel {
    const float4* vel4;
    Dims d;
    __device__ __forceinline__ float3 operator()(float px, float py, float pz) const {
        return sample4(vel4, d.W, d.H, d.D, px - 0.5f, py - 0.5f, pz - 0.5f);
    }
};

struct MacVel {
    const float* u;
    const float* v;
    const float* w;
    Dims d;
    __device__ __forceinline__ float3 operator()(float px, float py, float pz) const {
        float3 r;
        r.x = sample3(u, d.W + 1, d.H, d.D, px, py - 0.5f, pz - 0.5f);
        r.y = sample3(v, d.W, d.H + 1, d.D, px - 0.5f, py, pz - 0.5f);
        r.z = sample3(w, d.W, d.H, d.D + 1, px - 0.5f, py - 0.5f, pz);
        return r;
    }
};

// Semi-Lagrangian backtrace with the reference's loop (ZQ_CUDA_Advection3D.cu:87-109): stop when
// the point leaves [0,W]x[0,H]x[0,D] or lands in a solid cell (point sample, clamp); otherwise
// remember it and step back by vel*dtt/h.  The end point is the last accepted position.
template <class Vel>
__device__ __forceinline__ float3 backtrace(const Vel& V, const uint8_t* __restrict__ occ, const Dims& d, float h,
                                            int steps, float dtt, float px, float py, float pz) {
    float3 last = make_float3(px, py, pz);
    float3 vel = V(px, py, pz);
    const float fW = (float)d.W, fH = (float)d.H, fD = (float)d.D;
    int istep = 0;
    do {
        if (!(px >= 0.0f && px <= fW && py >= 0.0f && py <= fH && pz >= 0.0f && pz <= fD)) break;
        const int oi = clampi((int)floorf(px), 0, d.W - 1);
        const int oj = clampi((int)floorf(py), 0, d.H - 1);
        const int ok = clampi((int)floorf(pz), 0, d.D - 1);
        if (occ[cidx(d, oi, oj, ok)]) break;
        last = make_float3(px, py, pz);
        px = px - (vel.x * dtt) / h;
        py = py - (vel.y * dtt) / h;
        pz = pz - (vel.z * dtt) / h;
        vel = V(px, py, pz);
        istep++;
    } while (istep < steps);
    return last;
}

}  // namespace zq3d
