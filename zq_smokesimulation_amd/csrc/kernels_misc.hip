// kernels_misc.hip -- smoke sources, the analytic moving-sphere obstacle and fill kernels (gfx950).
#include "zq3d_device.h"
#include "zq3d_kernels.h"
#include "zq3d_stencils.h"

namespace zq3d {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid_for(int nx, int ny, int nz) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz); }

// Reinjection bodies: zq3d_stencils.h.  The adjoint runs the same kernels with zero values: a
// reinjected cell forgets its past.
template <class R>
__global__ __launch_bounds__(BX* BY) void k_reinject_box(R* __restrict__ T, R* __restrict__ rho, Dims d, float cx,
                                                         float cy, float cz, float hx, float hy, float hz, float rd,
                                                         float rt) {
    st_reinject_box<R>(T, rho, d, cx, cy, cz, hx, hy, hz, rd, rt, blockIdx.x * BX + threadIdx.x,
                       blockIdx.y * BY + threadIdx.y, blockIdx.z);
}

template <class R>
__global__ __launch_bounds__(BX* BY) void k_reinject_sphere(R* __restrict__ T, R* __restrict__ rho, Dims d, float x0,
                                                            float y0, float z0, float r2, float rd, float rt) {
    st_reinject_sphere<R>(T, rho, d, x0, y0, z0, r2, rd, rt, blockIdx.x * BX + threadIdx.x,
                          blockIdx.y * BY + threadIdx.y, blockIdx.z);
}

// Occupancy of an analytic solid sphere plus the wall shell, with the six faces of every solid
// voxel set to the solid's velocity (semantics of ZQ_ApplyMovingObject3D,
// ZQ_SmokeSimulationUtils3D.cpp:1933-2046).  Neighbouring solid voxels write equal values to their
// shared face.
__global__ __launch_bounds__(BX* BY) void k_stamp_sphere(uint8_t* __restrict__ occ, float* __restrict__ u,
                                                         float* __restrict__ v, float* __restrict__ w, Dims d,
                                                         float x0, float y0, float z0, float r2, float su, float sv,
                                                         float sw, int shell) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const float dx = i + 0.5f - x0, dy = j + 0.5f - y0, dz = k + 0.5f - z0;
    const bool in_sphere = dx * dx + dy * dy + dz * dz <= r2;
    if (in_sphere) {
        u[uidx(d, i, j, k)] = su;
        u[uidx(d, i + 1, j, k)] = su;
        v[vidx(d, i, j, k)] = sv;
        v[vidx(d, i, j + 1, k)] = sv;
        w[widx(d, i, j, k)] = sw;
        w[widx(d, i, j, k + 1)] = sw;
    }
    const bool wall = ((shell & 1) && i == 0) || ((shell & 2) && i == d.W - 1) || ((shell & 4) && j == 0) ||
                      ((shell & 8) && j == d.H - 1) || ((shell & 16) && k == 0) || ((shell & 32) && k == d.D - 1);
    occ[cidx(d, i, j, k)] = (in_sphere || wall) ? 1 : 0;
}

template <class T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ p, size_t n, T val) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t x = i; x < n; x += stride) p[x] = val;
}

template <class T>
void fill(T* p, size_t n, T val, hipStream_t s) {
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, n, val);
}

}  // namespace

template <class R>
void launch_reinject_box(const Dims& d, const float* b, R* T, R* rho, hipStream_t s) {
    const float cx = b[0] * d.W, cy = b[1] * d.H, cz = b[2] * d.D;
    const float hx = b[3] * d.W, hy = b[4] * d.H, hz = b[5] * d.D;
    hipLaunchKernelGGL(k_reinject_box<R>, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, T, rho, d, cx, cy, cz, hx, hy,
                       hz, b[6], b[7]);
}

template <class R>
void launch_reinject_sphere(const Dims& d, const float* sp, R* T, R* rho, hipStream_t s) {
    const float x0 = sp[0] * d.W, y0 = sp[1] * d.H, z0 = sp[2] * d.D, rr = sp[3] * d.W;
    hipLaunchKernelGGL(k_reinject_sphere<R>, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, T, rho, d, x0, y0, z0,
                       rr * rr, sp[4], sp[5]);
}

template void launch_reinject_box<float>(const Dims&, const float*, float*, float*, hipStream_t);
template void launch_reinject_box<double>(const Dims&, const float*, double*, double*, hipStream_t);
template void launch_reinject_sphere<float>(const Dims&, const float*, float*, float*, hipStream_t);
template void launch_reinject_sphere<double>(const Dims&, const float*, double*, double*, hipStream_t);

void launch_stamp_sphere(const Dims& d, float cx, float cy, float cz, float r, float su, float sv, float sw,
                         int shell_mask, uint8_t* occ, float* u, float* v, float* w, hipStream_t s) {
    const float x0 = cx * d.W, y0 = cy * d.H, z0 = cz * d.D, rr = r * d.W;
    hipLaunchKernelGGL(k_stamp_sphere, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, occ, u, v, w, d, x0, y0, z0,
                       rr * rr, su, sv, sw, shell_mask);
}

void launch_fill_f32(float* p, size_t n, float val, hipStream_t s) { fill<float>(p, n, val, s); }
void launch_fill_f64(double* p, size_t n, double val, hipStream_t s) { fill<double>(p, n, val, s); }
void launch_fill_u64(unsigned long long* p, size_t n, unsigned long long val, hipStream_t s) {
    fill<unsigned long long>(p, n, val, s);
}
void launch_fill_u8(uint8_t* p, size_t n, uint8_t val, hipStream_t s) { fill<uint8_t>(p, n, val, s); }

}  // namespace zq3d
