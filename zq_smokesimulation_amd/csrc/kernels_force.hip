// kernels_force.hip -- buoyancy + vorticity confinement, fused with attenuation (gfx950).
//
// Reference: cu_AddForce3D (ZQ_CUDA_AddForce3D.cu:510-556) runs six kernels over scratch
// arrays; cu_Attenuation3D (ZQ_CUDA_Attenuation3D.cu:120-134) four more.  Here:
//   k_vorticity     faces -> (wx,wy,wz,|w|) per cell, cell velocities formed on the fly
//   k_force_cells   |w| stencil -> normalized gradient -> f = eps*h*(N x w) + buoyancy,
//                   T and rho attenuated in the same pass
//   k_force_faces   face += dt*avg(f) and *= aV on one mask, the three face grids in one pass
// Threads: 64 x 4 per block, x across the wave's lanes (coalesced rows), one cell per thread.
#include "zq3d_device.h"
#include "zq3d_kernels.h"
#include "zq3d_stencils.h"

namespace zq3d {

namespace {

constexpr int BX = 64, BY = 4;

// The bodies live in zq3d_stencils.h (shared with the ensemble kernels).  VO / VI: where the
// vorticity records go and are read from (one 4-wide record per cell, or four planes).
template <class R, class VO>
__global__ __launch_bounds__(BX* BY) void k_vorticity(const R* __restrict__ u, const R* __restrict__ v,
                                                      const R* __restrict__ w, VO vort, Dims d, R two_h) {
    st_vorticity_to<R>(u, v, w, vort, d, two_h, blockIdx.x * BX + threadIdx.x, blockIdx.y * BY + threadIdx.y,
                       blockIdx.z);
}

template <class R, bool ATTEN, class VI>
__global__ __launch_bounds__(BX* BY) void k_force_cells(VI vort, R* __restrict__ T, R* __restrict__ rho,
                                                        const R* __restrict__ extra, R* __restrict__ fcell, Dims d,
                                                        R ch, R beta, R alpha, R Tamb, R aT, R aR) {
    st_force_cell_from<R, ATTEN>(vort, T, rho, extra, fcell, d, ch, beta, alpha, Tamb, aT, aR,
                                 blockIdx.x * BX + threadIdx.x, blockIdx.y * BY + threadIdx.y, blockIdx.z);
}

// All three face grids in one launch: thread (i, j) walks a chunk of FACE_RZ planes and updates each
// cell's low u, v and w faces (the last column / row / plane also the domain's high face) with
// st_force_face.  The w face below plane k reads the force and occupancy of plane k-1, which the
// same thread read one iteration earlier.
constexpr int FACE_RZ = 16;
template <class R, bool ATTEN>
__global__ __launch_bounds__(BX* BY) void k_force_faces(R* __restrict__ u, R* __restrict__ v, R* __restrict__ w,
                                                        const uint8_t* __restrict__ occ, const R* __restrict__ fc,
                                                        Dims d, R dt, R dth, R aV) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    if (i >= d.W || j >= d.H) return;
    const int k1 = min((int)(blockIdx.z + 1) * FACE_RZ, d.D);
    for (int k = blockIdx.z * FACE_RZ; k < k1; k++) {
        st_force_face<R, 0, ATTEN>(u, occ, fc, d, dt, dth, aV, i, j, k);
        if (i == d.W - 1) st_force_face<R, 0, ATTEN>(u, occ, fc, d, dt, dth, aV, i + 1, j, k);
        st_force_face<R, 1, ATTEN>(v, occ, fc, d, dt, dth, aV, i, j, k);
        if (j == d.H - 1) st_force_face<R, 1, ATTEN>(v, occ, fc, d, dt, dth, aV, i, j + 1, k);
        st_force_face<R, 2, ATTEN>(w, occ, fc, d, dt, dth, aV, i, j, k);
        if (k == d.D - 1) st_force_face<R, 2, ATTEN>(w, occ, fc, d, dt, dth, aV, i, j, k + 1);
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
    const int oh = occ[ci], ol = occ[(pos > 0 && pos < lim) ? ci - step : ci];  // both load together
    const bool ok = (pos == 0 || pos == lim) ? !oh : (!ol && !oh);
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

template <class R, bool ATTEN>
void force_faces(const Dims& d, R* u, R* v, R* w, const uint8_t* occ, const R* fcell, R dt, R dth, R aV,
                 hipStream_t s) {
    hipLaunchKernelGGL((k_force_faces<R, ATTEN>), grid_for(d.W, d.H, (d.D + FACE_RZ - 1) / FACE_RZ), dim3(BX, BY, 1),
                       0, s, u, v, w, occ, fcell, d, dt, dth, aV);
}

}  // namespace

template <class R>
void launch_vorticity(const Dims& d, const R* u, const R* v, const R* w, vec4_t<R>* vort, float h, hipStream_t s) {
    const R two_h = R(2) * (R)h;
    hipLaunchKernelGGL((k_vorticity<R, VortCellsOut<R>>), grid_for(d.W, d.H, d.D), dim3(BX, BY, 1), 0, s, u, v, w,
                       VortCellsOut<R>{vort}, d, two_h);
}

namespace {
template <class R, class VO, class VI>
void force_atten(const Dims& d, R* u, R* v, R* w, R* T, R* rho, const uint8_t* occ, const R* extra,
                 const ForceParams& fp, int atten, float aV, float aT, float aR, VO vo, VI vi, R* fcell,
                 hipStream_t s) {
    const dim3 blk(BX, BY, 1);
    hipLaunchKernelGGL((k_vorticity<R, VO>), grid_for(d.W, d.H, d.D), blk, 0, s, u, v, w, vo, d, R(2) * (R)fp.h);
    // the products are formed in float, as the reference's host code passes them
    const R ch = (R)(fp.eps * fp.h);
    const R dth = (R)(fp.dt * 0.5f);
    const R dt = (R)fp.dt;
    if (atten) {
        hipLaunchKernelGGL((k_force_cells<R, true, VI>), grid_for(d.W, d.H, d.D), blk, 0, s, vi, T, rho, extra, fcell,
                           d, ch, (R)fp.beta, (R)fp.alpha, (R)fp.Tamb, (R)aT, (R)aR);
        force_faces<R, true>(d, u, v, w, occ, fcell, dt, dth, (R)aV, s);
    } else {
        hipLaunchKernelGGL((k_force_cells<R, false, VI>), grid_for(d.W, d.H, d.D), blk, 0, s, vi, T, rho, extra,
                           fcell, d, ch, (R)fp.beta, (R)fp.alpha, (R)fp.Tamb, R(1), R(1));
        force_faces<R, false>(d, u, v, w, occ, fcell, dt, dth, R(1), s);
    }
}
}  // namespace

template <class R>
void launch_force_atten(const Dims& d, R* u, R* v, R* w, R* T, R* rho, const uint8_t* occ, const R* extra,
                        const ForceParams& fp, int atten, float aV, float aT, float aR, vec4_t<R>* vort, R* fcell,
                        hipStream_t s, bool vort_planes) {
    if (vort_planes) {
        const int n = d.W * d.H * d.D;
        force_atten<R>(d, u, v, w, T, rho, occ, extra, fp, atten, aV, aT, aR, VortPlanesOut<R>{(R*)vort, n},
                       VortPlanesIn<R>{(const R*)vort, n}, fcell, s);
    } else {
        force_atten<R>(d, u, v, w, T, rho, occ, extra, fp, atten, aV, aT, aR, VortCellsOut<R>{vort},
                       VortCellsIn<R>{vort}, fcell, s);
    }
}

template void launch_vorticity<float>(const Dims&, const float*, const float*, const float*, float4*, float,
                                      hipStream_t);
template void launch_vorticity<double>(const Dims&, const double*, const double*, const double*, double4*, float,
                                       hipStream_t);
template void launch_force_atten<float>(const Dims&, float*, float*, float*, float*, float*, const uint8_t*,
                                        const float*, const ForceParams&, int, float, float, float, float4*, float*,
                                        hipStream_t, bool);
template void launch_force_atten<double>(const Dims&, double*, double*, double*, double*, double*, const uint8_t*,
                                         const double*, const ForceParams&, int, float, float, float, double4*,
                                         double*, hipStream_t, bool);

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
