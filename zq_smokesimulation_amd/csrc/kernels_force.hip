// kernels_force.hip -- buoyancy + vorticity confinement, fused with attenuation (gfx950).
//
// Reference: cu_AddForce3D (ZQ_CUDA_AddForce3D.cu:510-556) runs six kernels over scratch
// arrays; cu_Attenuation3D (ZQ_CUDA_Attenuation3D.cu:120-134) four more.  Here:
//   k_vorticity     faces -> (wx,wy,wz,|w|) per cell, cell velocities formed on the fly
//   k_force_cells   |w| stencil -> normalized gradient -> f = eps*h*(N x w) + buoyancy,
//                   T and rho attenuated in the same pass
//   k_force_faces   face += dt*avg(f) and *= aV on one mask, one pass per face grid
// Threads: 64 x 4 per block, x across the wave's lanes (coalesced rows), one cell per thread.
#include "zq3d_device.h"
#include "zq3d_kernels.h"

namespace zq3d {

namespace {

constexpr int BX = 64, BY = 4;

__device__ __forceinline__ float cell_u(const float* __restrict__ u, const Dims& d, int i, int j, int k) {
    return 0.5f * (u[uidx(d, i, j, k)] + u[uidx(d, i + 1, j, k)]);
}
__device__ __forceinline__ float cell_v(const float* __restrict__ v, const Dims& d, int i, int j, int k) {
    return 0.5f * (v[vidx(d, i, j, k)] + v[vidx(d, i, j + 1, k)]);
}
__device__ __forceinline__ float cell_w(const float* __restrict__ w, const Dims& d, int i, int j, int k) {
    return 0.5f * (w[widx(d, i, j, k)] + w[widx(d, i, j, k + 1)]);
}

// Compute_Vorticity_Vector_Scale_Kernel (ZQ_CUDA_AddForce3D.cu:11-34): central differences of the
// cell velocity over 2h on interior cells, zero on the border (the reference memsets the arrays).
__global__ __launch_bounds__(BX* BY) void k_vorticity(const float* __restrict__ u, const float* __restrict__ v,
                                                      const float* __restrict__ w, float4* __restrict__ vort,
                                                      Dims d, float two_h) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i >= 1 && i < d.W - 1 && j >= 1 && j < d.H - 1 && k >= 1 && k < d.D - 1) {
        const float wx = (cell_w(w, d, i, j + 1, k) - cell_w(w, d, i, j - 1, k) - cell_v(v, d, i, j, k + 1) +
                          cell_v(v, d, i, j, k - 1)) / two_h;
        const float wy = (cell_u(u, d, i, j, k + 1) - cell_u(u, d, i, j, k - 1) - cell_w(w, d, i + 1, j, k) +
                          cell_w(w, d, i - 1, j, k)) / two_h;
        const float wz = (cell_v(v, d, i + 1, j, k) - cell_v(v, d, i - 1, j, k) - cell_u(u, d, i, j + 1, k) +
                          cell_u(u, d, i, j - 1, k)) / two_h;
        out = make_float4(wx, wy, wz, sqrtf(wx * wx + wy * wy + wz * wz));
    }
    vort[cidx(d, i, j, k)] = out;
}

// Compute_Gradient_of_VorticityScale_Kernel (:64-111) + Compute_Force_Kernel (:161-184).
// Output is SoA: fcell[0..N) = fx, [N..2N) = fy, [2N..3N) = fz.
template <bool ATTEN>
__global__ __launch_bounds__(BX* BY) void k_force_cells(const float4* __restrict__ vort, float* __restrict__ T,
                                                        float* __restrict__ rho, const float* __restrict__ extra,
                                                        float* __restrict__ fcell, Dims d, float ch, float beta,
                                                        float alpha, float Tamb, float aT, float aR) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const int o = cidx(d, i, j, k);
    const int HW = d.H * d.W;
    const int N = HW * d.D;
    const float s0 = vort[o].w;
    float gx, gy, gz;
    if (i == 0) gx = vort[o + 1].w - s0;
    else if (i == d.W - 1) gx = s0 - vort[o - 1].w;
    else gx = 0.5f * (vort[o + 1].w - vort[o - 1].w);
    if (j == 0) gy = vort[o + d.W].w - s0;
    else if (j == d.H - 1) gy = s0 - vort[o - d.W].w;
    else gy = 0.5f * (vort[o + d.W].w - vort[o - d.W].w);
    if (k == 0) gz = vort[o + HW].w - s0;
    else if (k == d.D - 1) gz = s0 - vort[o - HW].w;
    else gz = 0.5f * (vort[o + HW].w - vort[o - HW].w);
    const float len = sqrtf(gx * gx + gy * gy + gz * gz);
    if (len != 0.0f) {
        gx = gx / len;
        gy = gy / len;
        gz = gz / len;
    }
    const float4 wv = vort[o];
    float fx = ch * (gy * wv.z - gz * wv.y);
    float fy = ch * (gz * wv.x - gx * wv.z);
    float fz = ch * (gx * wv.y - gy * wv.x);
    const float t = T[o];
    const float r = (ATTEN || alpha != 0.0f) ? rho[o] : 0.0f;  // AddForce3D passes no density
    fy += beta * (t - Tamb);
    if (alpha != 0.0f) fy -= alpha * r;
    if (extra) {
        fx += extra[o * 3 + 0];
        fy += extra[o * 3 + 1];
        fz += extra[o * 3 + 2];
    }
    fcell[o] = fx;
    fcell[N + o] = fy;
    fcell[2 * N + o] = fz;
    if (ATTEN) {
        T[o] = t * aT;
        rho[o] = r * aR;
    }
}

// AddForce_{u,v,w}_Kernel (:211-420) fused with Atten_{u,v,w}_Kernel (ZQ_CUDA_Attenuation3D.cu:8-96):
// both stages use the same face mask (both cells fluid; a domain face needs its inside cell fluid),
// so the face becomes (face + dt*avg f) * aV in one read-modify-write.  AXIS 0/1/2 = u/v/w.
// The reference's AddForce_w_Kernel also runs its interior formula at z = 0 reading cell z-1
// (:380-384, out of bounds); this follows the slice variant (:400-419).
template <int AXIS, bool ATTEN>
__global__ __launch_bounds__(BX* BY) void k_force_faces(float* __restrict__ face, const uint8_t* __restrict__ occ,
                                                        const float* __restrict__ fc, Dims d, float dt, float dth,
                                                        float aV) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
    const int step = AXIS == 0 ? 1 : (AXIS == 1 ? d.W : d.W * d.H);
    const int ci = cidx(d, AXIS == 0 ? (i < d.W ? i : d.W - 1) : i, AXIS == 1 ? (j < d.H ? j : d.H - 1) : j,
                        AXIS == 2 ? (k < d.D ? k : d.D - 1) : k);
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    const float* f = fc + (size_t)AXIS * d.W * d.H * d.D;
    bool ok;
    float add;
    if (pos == 0) {
        ok = !occ[ci];
        add = dt * f[ci];
    } else if (pos == lim) {
        ok = !occ[ci];  // ci already points at the last cell
        add = dt * f[ci];
    } else {
        ok = !occ[ci - step] && !occ[ci];
        add = dth * (f[ci] + f[ci - step]);
    }
    if (ok) {
        float x = face[fi] + add;
        if (ATTEN) x = x * aV;
        face[fi] = x;
    }
}

// Atten_{u,v,w}_Kernel without a force (:8-96)
template <int AXIS>
__global__ __launch_bounds__(BX* BY) void k_atten_faces(float* __restrict__ face, const uint8_t* __restrict__ occ,
                                                        Dims d, float aV) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
    const int step = AXIS == 0 ? 1 : (AXIS == 1 ? d.W : d.W * d.H);
    const int ci = cidx(d, AXIS == 0 ? (i < d.W ? i : d.W - 1) : i, AXIS == 1 ? (j < d.H ? j : d.H - 1) : j,
                        AXIS == 2 ? (k < d.D ? k : d.D - 1) : k);
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    const bool ok = (pos == 0 || pos == lim) ? !occ[ci] : (!occ[ci - step] && !occ[ci]);
    if (ok) face[fi] = face[fi] * aV;
}

// Atten_temperature_density_Kernel (ZQ_CUDA_Attenuation3D.cu:98-116)
__global__ __launch_bounds__(256) void k_atten_cells(float* __restrict__ T, float* __restrict__ rho, int n, float aT,
                                                     float aR) {
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= n) return;
    T[o] = T[o] * aT;
    rho[o] = rho[o] * aR;
}

inline dim3 grid_for(int nx, int ny, int nz) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz); }

}  // namespace

void launch_force_atten(const Dims& d, float* u, float* v, float* w, float* T, float* rho, const uint8_t* occ,
                        const float* extra, const ForceParams& fp, int atten, float aV, float aT, float aR,
                        float4* vort, float* fcell, hipStream_t s) {
    const dim3 blk(BX, BY, 1);
    const float two_h = 2 * fp.h;
    hipLaunchKernelGGL(k_vorticity, grid_for(d.W, d.H, d.D), blk, 0, s, u, v, w, vort, d, two_h);
    const float ch = fp.eps * fp.h;
    const float dth = fp.dt * 0.5f;
    if (atten) {
        hipLaunchKernelGGL(k_force_cells<true>, grid_for(d.W, d.H, d.D), blk, 0, s, vort, T, rho, extra, fcell, d,
                           ch, fp.beta, fp.alpha, fp.Tamb, aT, aR);
        hipLaunchKernelGGL((k_force_faces<0, true>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, u, occ, fcell, d, fp.dt,
                           dth, aV);
        hipLaunchKernelGGL((k_force_faces<1, true>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, v, occ, fcell, d, fp.dt,
                           dth, aV);
        hipLaunchKernelGGL((k_force_faces<2, true>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, w, occ, fcell, d, fp.dt,
                           dth, aV);
    } else {
        hipLaunchKernelGGL(k_force_cells<false>, grid_for(d.W, d.H, d.D), blk, 0, s, vort, T, rho, extra, fcell, d,
                           ch, fp.beta, fp.alpha, fp.Tamb, 1.0f, 1.0f);
        hipLaunchKernelGGL((k_force_faces<0, false>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, u, occ, fcell, d,
                           fp.dt, dth, 1.0f);
        hipLaunchKernelGGL((k_force_faces<1, false>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, v, occ, fcell, d,
                           fp.dt, dth, 1.0f);
        hipLaunchKernelGGL((k_force_faces<2, false>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, w, occ, fcell, d,
                           fp.dt, dth, 1.0f);
    }
}

void launch_attenuation(const Dims& d, float* u, float* v, float* w, float* T, float* rho, const uint8_t* occ,
                        float aV, float aT, float aR, hipStream_t s) {
    const dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL(k_atten_faces<0>, grid_for(d.W + 1, d.H, d.D), blk, 0, s, u, occ, d, aV);
    hipLaunchKernelGGL(k_atten_faces<1>, grid_for(d.W, d.H + 1, d.D), blk, 0, s, v, occ, d, aV);
    hipLaunchKernelGGL(k_atten_faces<2>, grid_for(d.W, d.H, d.D + 1), blk, 0, s, w, occ, d, aV);
    const int n = d.W * d.H * d.D;
    hipLaunchKernelGGL(k_atten_cells, dim3((n + 255) / 256), dim3(256), 0, s, T, rho, n, aT, aR);
}

}  // namespace zq3d
