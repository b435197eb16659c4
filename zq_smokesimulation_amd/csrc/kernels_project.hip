// kernels_project.hip -- pressure projection: divergence, red-black SOR, gradient correction,
// FP64 residual and iterative refinement (gfx950).
//
// The reference's solver (ZQ_CUDA_PoissonSolver3D::Solve{Open,Closed}PoissonRedBlackwithOccupy3D_MAC,
// called at ZQ_SmokeSimulationUtils3D.cpp:2354,2424) lives in the un-vendored ZQlibCuda.  This file
// defines it from scratch:
//   * per-cell neighbour count n: in-domain fluid neighbours, plus outside-domain neighbours across
//     open faces (p = 0 there); solid cells and walls are Neumann and not counted;
//   * colour (i+j+k)&1, colour 0 first; p = (1-omega) p + omega (sum_nb p - b)/n with the sum taken
//     over all six neighbours of a padded pressure whose ghost layer and solid cells hold 0;
//   * u_face -= p_r - p_l on faces between two fluid cells and on open domain faces.
#include "zq3d_device.h"
#include "zq3d_kernels.h"

namespace zq3d {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid_for(int nx, int ny, int nz) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz); }

__device__ __forceinline__ double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ double wave_max(double x) {
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    return x;
}
__device__ __forceinline__ bool lane0() { return ((threadIdx.y * blockDim.x + threadIdx.x) & 63) == 0; }

// 1/n for n = 1..6, each correctly rounded (the same values as 1.0f/(float)n)
__device__ __forceinline__ float rcp_count(int n) {
    return n == 1 ? 1.0f : n == 2 ? 0.5f : n == 3 ? (1.0f / 3.0f) : n == 4 ? 0.25f : n == 5 ? (1.0f / 5.0f)
                                                                                            : (1.0f / 6.0f);
}

// Neighbour counts, divergence (the reference's unscaled sum, ZQ_SmokeSimulationUtils3D.cpp:2403-2405,
// evaluated left to right), zero pressure where the count is 0, and the FP64 sum/count of the
// divergence over fluid cells for the closed-domain mean.
template <bool MP>
__global__ __launch_bounds__(BX* BY) void k_proj_setup(const float* __restrict__ u, const float* __restrict__ v,
                                                       const float* __restrict__ w, const uint8_t* __restrict__ occ,
                                                       int open_mask, uint8_t* __restrict__ nc,
                                                       float* __restrict__ b, double* __restrict__ b64,
                                                       float* __restrict__ p, double* __restrict__ p64,
                                                       double* __restrict__ sum_cnt, Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double dsum = 0.0, dcnt = 0.0;
    if (i < d.W && j < d.H) {
        const int o = cidx(d, i, j, k);
        const int HW = d.W * d.H;
        int n = 0;
        const bool fluid = !occ[o];
        if (fluid) {
            n += (i > 0) ? !occ[o - 1] : ((open_mask & 1) != 0);
            n += (i < d.W - 1) ? !occ[o + 1] : ((open_mask & 2) != 0);
            n += (j > 0) ? !occ[o - d.W] : ((open_mask & 4) != 0);
            n += (j < d.H - 1) ? !occ[o + d.W] : ((open_mask & 8) != 0);
            n += (k > 0) ? !occ[o - HW] : ((open_mask & 16) != 0);
            n += (k < d.D - 1) ? !occ[o + HW] : ((open_mask & 32) != 0);
        }
        nc[o] = (uint8_t)n;
        const float u0 = u[uidx(d, i, j, k)], u1 = u[uidx(d, i + 1, j, k)];
        const float v0 = v[vidx(d, i, j, k)], v1 = v[vidx(d, i, j + 1, k)];
        const float w0 = w[widx(d, i, j, k)], w1 = w[widx(d, i, j, k + 1)];
        if (MP) {
            const double dd = (((double)u1 - (double)u0) + ((double)v1 - (double)v0)) + ((double)w1 - (double)w0);
            b64[o] = n ? dd : 0.0;
            if (!n) p64[pidx(d, i, j, k)] = 0.0;
            if (fluid) {
                dsum = dd;
                dcnt = 1.0;
            }
        } else {
            const float dv = u1 - u0 + v1 - v0 + w1 - w0;
            b[o] = fluid ? dv : 0.0f;
            if (!n) p[pidx(d, i, j, k)] = 0.0f;
            if (fluid) {
                dsum = (double)dv;
                dcnt = 1.0;
            }
        }
    }
    if (sum_cnt) {
        dsum = wave_sum(dsum);
        dcnt = wave_sum(dcnt);
        if (lane0()) {
            atomicAdd(&sum_cnt[0], dsum);
            atomicAdd(&sum_cnt[1], dcnt);
        }
    }
}

__global__ void k_mean_finalize(const double* __restrict__ sum_cnt, float* __restrict__ mean32,
                                double* __restrict__ mean64, int want) {
    double m = 0.0;
    if (want && sum_cnt[1] > 0.0) m = sum_cnt[0] / sum_cnt[1];
    if (mean32) *mean32 = (float)m;
    if (mean64) *mean64 = m;
}

// One colour of a red-black SOR iteration.  Thread (t, j, k) owns cell x = 2t + ((colour+j+k)&1);
// the 64 lanes of a wave cover 128 consecutive cells of a row.
__global__ __launch_bounds__(BX* BY) void k_sor_half(const uint8_t* __restrict__ nc, const float* __restrict__ b,
                                                     const float* __restrict__ mean32, float* __restrict__ p, Dims d,
                                                     float omega, float omw, int colour) {
    const int t = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int i = 2 * t + ((colour + j + k) & 1);
    if (i >= d.W || j >= d.H) return;
    const int o = cidx(d, i, j, k);
    const int n = nc[o];
    if (!n) return;
    const int sy = d.W + 2, sz = (d.W + 2) * (d.H + 2);
    const int q = pidx(d, i, j, k);
    const float s = p[q - 1] + p[q + 1] + p[q - sy] + p[q + sy] + p[q - sz] + p[q + sz];
    const float bb = b[o] - *mean32;
    const float pg = (s - bb) * rcp_count(n);
    p[q] = omw * p[q] + omega * pg;
}

// Face correction u -= p_r - p_l (ZQ_SmokeSimulationUtils3D.cpp:2815-2861 open form, :2888-2922
// closed form).  P = float (plain) or double (mixed precision; the face is rounded once).
template <int AXIS, class P>
__global__ __launch_bounds__(BX* BY) void k_correct(const uint8_t* __restrict__ occ, int open_mask,
                                                    const P* __restrict__ p, float* __restrict__ face, Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
    const int step = AXIS == 0 ? 1 : (AXIS == 1 ? d.W : d.W * d.H);
    const int pstep = AXIS == 0 ? 1 : (AXIS == 1 ? d.W + 2 : (d.W + 2) * (d.H + 2));
    const int lo_bit = AXIS == 0 ? 1 : (AXIS == 1 ? 4 : 16);
    const int ci = cidx(d, AXIS == 0 ? (i < d.W ? i : d.W - 1) : i, AXIS == 1 ? (j < d.H ? j : d.H - 1) : j,
                        AXIS == 2 ? (k < d.D ? k : d.D - 1) : k);
    bool ok;
    if (pos == 0) ok = (open_mask & lo_bit) && !occ[ci];
    else if (pos == lim) ok = (open_mask & (lo_bit << 1)) && !occ[ci];
    else ok = !occ[ci - step] && !occ[ci];
    if (!ok) return;
    // padded index of the cell on the high side of the face (ghost when pos == lim)
    const int qh = pidx(d, i, j, k);
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    face[fi] = (float)((P)face[fi] - (p[qh] - p[qh - pstep]));
}

// FP64 residual r = (b64 - mean) - (sum_nb p - n p); r32 feeds the FP32 inner solve.
__global__ __launch_bounds__(BX* BY) void k_residual64(const uint8_t* __restrict__ nc, const double* __restrict__ b64,
                                                       const double* __restrict__ mean64,
                                                       const double* __restrict__ p, float* __restrict__ r32,
                                                       unsigned long long* __restrict__ rmax, Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double m = 0.0;
    if (i < d.W && j < d.H) {
        const int o = cidx(d, i, j, k);
        const int n = nc[o];
        float r = 0.0f;
        if (n) {
            const int sy = d.W + 2, sz = (d.W + 2) * (d.H + 2);
            const int q = pidx(d, i, j, k);
            const double s = p[q - 1] + p[q + 1] + p[q - sy] + p[q + sy] + p[q - sz] + p[q + sz];
            const double bb = b64[o] - *mean64;
            const double rr = bb - (s - (double)n * p[q]);
            r = (float)rr;
            m = fabs(rr);
        }
        r32[o] = r;
    }
    m = wave_max(m);
    if (lane0()) atomicMax(rmax, (unsigned long long)__double_as_longlong(m));
}

__global__ __launch_bounds__(BX* BY) void k_update_p64(const uint8_t* __restrict__ nc, float* __restrict__ e,
                                                       double* __restrict__ p, Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    if (!nc[cidx(d, i, j, k)]) return;
    const int q = pidx(d, i, j, k);
    p[q] = p[q] + (double)e[q];
    e[q] = 0.0f;
}

__global__ __launch_bounds__(BX* BY) void k_max_div64(const float* __restrict__ u, const float* __restrict__ v,
                                                      const float* __restrict__ w, const uint8_t* __restrict__ nc,
                                                      const double* __restrict__ mean64,
                                                      unsigned long long* __restrict__ dmax, Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double m = 0.0;
    if (i < d.W && j < d.H && nc[cidx(d, i, j, k)]) {
        const double du = (double)u[uidx(d, i + 1, j, k)] - (double)u[uidx(d, i, j, k)];
        const double dv = (double)v[vidx(d, i, j + 1, k)] - (double)v[vidx(d, i, j, k)];
        const double dw = (double)w[widx(d, i, j, k + 1)] - (double)w[widx(d, i, j, k)];
        m = fabs(((du + dv) + dw) - *mean64);
    }
    m = wave_max(m);
    if (lane0()) atomicMax(dmax, (unsigned long long)__double_as_longlong(m));
}

}  // namespace

void launch_proj_setup(const Dims& d, const float* u, const float* v, const float* w, const uint8_t* occ,
                       int open_mask, uint8_t* nc, float* b, double* b64, float* p_pad, double* p64_pad,
                       double* sum_cnt, hipStream_t s) {
    if (b64)
        hipLaunchKernelGGL(k_proj_setup<true>, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, u, v, w, occ, open_mask,
                           nc, b, b64, p_pad, p64_pad, sum_cnt, d);
    else
        hipLaunchKernelGGL(k_proj_setup<false>, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, u, v, w, occ, open_mask,
                           nc, b, b64, p_pad, p64_pad, sum_cnt, d);
}

void launch_mean_finalize(const double* sum_cnt, float* mean32, double* mean64, int want, hipStream_t s) {
    hipLaunchKernelGGL(k_mean_finalize, dim3(1), dim3(1), 0, s, sum_cnt, mean32, mean64, want);
}

void launch_sor(const Dims& d, const uint8_t* nc, const float* b, const float* mean32, float* p_pad, float omega,
                int iters, hipStream_t s) {
    const dim3 g = grid_for((d.W + 1) / 2, d.H, d.D);
    const float omw = 1.0f - omega;
    for (int it = 0; it < iters; it++) {
        hipLaunchKernelGGL(k_sor_half, g, dim3(BX, BY), 0, s, nc, b, mean32, p_pad, d, omega, omw, 0);
        hipLaunchKernelGGL(k_sor_half, g, dim3(BX, BY), 0, s, nc, b, mean32, p_pad, d, omega, omw, 1);
    }
}

void launch_correct(const Dims& d, const uint8_t* occ, int open_mask, const float* p_pad, float* u, float* v,
                    float* w, hipStream_t s) {
    hipLaunchKernelGGL((k_correct<0, float>), grid_for(d.W + 1, d.H, d.D), dim3(BX, BY), 0, s, occ, open_mask, p_pad,
                       u, d);
    hipLaunchKernelGGL((k_correct<1, float>), grid_for(d.W, d.H + 1, d.D), dim3(BX, BY), 0, s, occ, open_mask, p_pad,
                       v, d);
    hipLaunchKernelGGL((k_correct<2, float>), grid_for(d.W, d.H, d.D + 1), dim3(BX, BY), 0, s, occ, open_mask, p_pad,
                       w, d);
}

void launch_correct64(const Dims& d, const uint8_t* occ, int open_mask, const double* p64_pad, float* u, float* v,
                      float* w, hipStream_t s) {
    hipLaunchKernelGGL((k_correct<0, double>), grid_for(d.W + 1, d.H, d.D), dim3(BX, BY), 0, s, occ, open_mask,
                       p64_pad, u, d);
    hipLaunchKernelGGL((k_correct<1, double>), grid_for(d.W, d.H + 1, d.D), dim3(BX, BY), 0, s, occ, open_mask,
                       p64_pad, v, d);
    hipLaunchKernelGGL((k_correct<2, double>), grid_for(d.W, d.H, d.D + 1), dim3(BX, BY), 0, s, occ, open_mask,
                       p64_pad, w, d);
}

void launch_residual64(const Dims& d, const uint8_t* nc, const double* b64, const double* mean64,
                       const double* p64_pad, float* r32, unsigned long long* rmax, hipStream_t s) {
    hipLaunchKernelGGL(k_residual64, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, nc, b64, mean64, p64_pad, r32, rmax,
                       d);
}

void launch_update_p64(const Dims& d, const uint8_t* nc, float* e_pad, double* p64_pad, hipStream_t s) {
    hipLaunchKernelGGL(k_update_p64, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, nc, e_pad, p64_pad, d);
}

void launch_max_div64(const Dims& d, const float* u, const float* v, const float* w, const uint8_t* nc,
                      const double* mean64, unsigned long long* dmax, hipStream_t s) {
    hipLaunchKernelGGL(k_max_div64, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, u, v, w, nc, mean64, dmax, d);
}

}  // namespace zq3d
