// zq3d_stencils.h -- the per-cell / per-face work of the smoke step as device functions.
//
// Each function is the body of one kernel for one cell (i, j, k) or face of a grid d.  The
// single-simulation kernels (kernels_force/advect/project/misc.hip) map blockIdx/threadIdx to
// (i, j, k) and call them; the ensemble kernels (kernels_ensemble.hip) map a tile of a packed
// multi-grid launch to (simulation, i, j, k) and call the same functions, so both compute the same
// values bit for bit.  Every function returns at once for indices outside its grid.
#pragma once

#include "zq3d_device.h"

namespace zq3d {

// index of the cell on the high side of face (i,j,k) of axis AXIS, clamped into the grid, and the
// face's position along the axis
template <int AXIS>
__device__ __forceinline__ int face_cell(const Dims& d, int i, int j, int k) {
    return cidx(d, AXIS == 0 ? (i < d.W ? i : d.W - 1) : i, AXIS == 1 ? (j < d.H ? j : d.H - 1) : j,
                AXIS == 2 ? (k < d.D ? k : d.D - 1) : k);
}
template <int AXIS>
__device__ __forceinline__ int face_index(const Dims& d, int i, int j, int k) {
    return AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
}
template <int AXIS>
__device__ __forceinline__ bool face_in(const Dims& d, int i, int j, int k) {
    return i < d.W + (AXIS == 0) && j < d.H + (AXIS == 1) && k < d.D + (AXIS == 2);
}

// ---- force + attenuation (ZQ_CUDA_AddForce3D.cu, ZQ_CUDA_Attenuation3D.cu) ----

template <class R>
__device__ __forceinline__ R st_cell_u(const R* __restrict__ u, const Dims& d, int i, int j, int k) {
    return R(0.5) * (u[uidx(d, i, j, k)] + u[uidx(d, i + 1, j, k)]);
}
template <class R>
__device__ __forceinline__ R st_cell_v(const R* __restrict__ v, const Dims& d, int i, int j, int k) {
    return R(0.5) * (v[vidx(d, i, j, k)] + v[vidx(d, i, j + 1, k)]);
}
template <class R>
__device__ __forceinline__ R st_cell_w(const R* __restrict__ w, const Dims& d, int i, int j, int k) {
    return R(0.5) * (w[widx(d, i, j, k)] + w[widx(d, i, j, k + 1)]);
}

// Compute_Vorticity_Vector_Scale_Kernel (ZQ_CUDA_AddForce3D.cu:11-34): central differences of the
// cell velocity over 2h on interior cells, zero on the border (the reference memsets the arrays).
// Where the vorticity stage keeps (wx, wy, wz, |w|) per cell.  Cells: one 4-wide record per cell.
// Planes: four planes of N values (wx, wy, wz, |w|) in the same 4N-value buffer, so the force
// stage's |w| gradient reads six dense neighbours instead of one lane of 16-byte records.  The
// writers are the vorticity stage's, the readers the force stage's.
template <class R>
struct VortCellsOut {
    vec4_t<R>* p;
    __device__ __forceinline__ void store(int o, const vec4_t<R>& v) const { p[o] = v; }
};
template <class R>
struct VortCellsIn {
    const vec4_t<R>* p;
    __device__ __forceinline__ R mag(int o) const { return p[o].w; }
    __device__ __forceinline__ vec4_t<R> get(int o) const { return p[o]; }
};
template <class R>
struct VortPlanesOut {
    R* p;
    int n;
    __device__ __forceinline__ void store(int o, const vec4_t<R>& v) const {
        p[o] = v.x;
        p[n + o] = v.y;
        p[2 * n + o] = v.z;
        p[3 * n + o] = v.w;
    }
};
template <class R>
struct VortPlanesIn {
    const R* p;
    int n;
    __device__ __forceinline__ R mag(int o) const { return p[3 * n + o]; }
    __device__ __forceinline__ vec4_t<R> get(int o) const {
        vec4_t<R> v;
        v.x = p[o];
        v.y = p[n + o];
        v.z = p[2 * n + o];
        v.w = p[3 * n + o];
        return v;
    }
};

template <class R, class VS>
__device__ __forceinline__ void st_vorticity_to(const R* __restrict__ u, const R* __restrict__ v,
                                                const R* __restrict__ w, const VS& vort, const Dims& d, R two_h,
                                                int i, int j, int k) {
    if (i >= d.W || j >= d.H || k >= d.D) return;
    vec4_t<R> out;
    out.x = out.y = out.z = out.w = R(0);
    if (i >= 1 && i < d.W - 1 && j >= 1 && j < d.H - 1 && k >= 1 && k < d.D - 1) {
        const R wx = (st_cell_w(w, d, i, j + 1, k) - st_cell_w(w, d, i, j - 1, k) - st_cell_v(v, d, i, j, k + 1) +
                      st_cell_v(v, d, i, j, k - 1)) / two_h;
        const R wy = (st_cell_u(u, d, i, j, k + 1) - st_cell_u(u, d, i, j, k - 1) - st_cell_w(w, d, i + 1, j, k) +
                      st_cell_w(w, d, i - 1, j, k)) / two_h;
        const R wz = (st_cell_v(v, d, i + 1, j, k) - st_cell_v(v, d, i - 1, j, k) - st_cell_u(u, d, i, j + 1, k) +
                      st_cell_u(u, d, i, j - 1, k)) / two_h;
        out.x = wx;
        out.y = wy;
        out.z = wz;
        out.w = zsqrt(wx * wx + wy * wy + wz * wz);
    }
    vort.store(cidx(d, i, j, k), out);
}

template <class R>
__device__ __forceinline__ void st_vorticity(const R* __restrict__ u, const R* __restrict__ v,
                                             const R* __restrict__ w, vec4_t<R>* __restrict__ vort, const Dims& d,
                                             R two_h, int i, int j, int k) {
    st_vorticity_to<R>(u, v, w, VortCellsOut<R>{vort}, d, two_h, i, j, k);
}

// Compute_Gradient_of_VorticityScale_Kernel (:64-111) + Compute_Force_Kernel (:161-184).
// Output is SoA: fcell[0..N) = fx, [N..2N) = fy, [2N..3N) = fz.
template <class R, bool ATTEN, class VS>
__device__ __forceinline__ void st_force_cell_from(const VS& vort, R* __restrict__ T, R* __restrict__ rho,
                                                   const R* __restrict__ extra, R* __restrict__ fcell, const Dims& d,
                                                   R ch, R beta, R alpha, R Tamb, R aT, R aR, int i, int j, int k) {
    if (i >= d.W || j >= d.H || k >= d.D) return;
    const int o = cidx(d, i, j, k);
    const int HW = d.H * d.W;
    const int N = HW * d.D;
    const R s0 = vort.mag(o);
    R gx, gy, gz;
    if (i == 0) gx = vort.mag(o + 1) - s0;
    else if (i == d.W - 1) gx = s0 - vort.mag(o - 1);
    else gx = R(0.5) * (vort.mag(o + 1) - vort.mag(o - 1));
    if (j == 0) gy = vort.mag(o + d.W) - s0;
    else if (j == d.H - 1) gy = s0 - vort.mag(o - d.W);
    else gy = R(0.5) * (vort.mag(o + d.W) - vort.mag(o - d.W));
    if (k == 0) gz = vort.mag(o + HW) - s0;
    else if (k == d.D - 1) gz = s0 - vort.mag(o - HW);
    else gz = R(0.5) * (vort.mag(o + HW) - vort.mag(o - HW));
    const R len = zsqrt(gx * gx + gy * gy + gz * gz);
    if (len != R(0)) {
        gx = gx / len;
        gy = gy / len;
        gz = gz / len;
    }
    const vec4_t<R> wv = vort.get(o);
    R fx = ch * (gy * wv.z - gz * wv.y);
    R fy = ch * (gz * wv.x - gx * wv.z);
    R fz = ch * (gx * wv.y - gy * wv.x);
    const R t = T[o];
    const R r = (ATTEN || alpha != R(0)) ? rho[o] : R(0);  // AddForce3D passes no density
    fy += beta * (t - Tamb);
    if (alpha != R(0)) fy -= alpha * r;
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

template <class R, bool ATTEN>
__device__ __forceinline__ void st_force_cell(const vec4_t<R>* __restrict__ vort, R* __restrict__ T,
                                              R* __restrict__ rho, const R* __restrict__ extra, R* __restrict__ fcell,
                                              const Dims& d, R ch, R beta, R alpha, R Tamb, R aT, R aR, int i, int j,
                                              int k) {
    st_force_cell_from<R, ATTEN>(VortCellsIn<R>{vort}, T, rho, extra, fcell, d, ch, beta, alpha, Tamb, aT, aR, i,
                                 j, k);
}

// AddForce_{u,v,w}_Kernel (:211-420) fused with Atten_{u,v,w}_Kernel (ZQ_CUDA_Attenuation3D.cu:8-96):
// both stages use the same face mask (both cells fluid; a domain face needs its inside cell fluid),
// so the face becomes (face + dt*avg f) * aV in one read-modify-write.  AXIS 0/1/2 = u/v/w.
// The reference's AddForce_w_Kernel also runs its interior formula at z = 0 reading cell z-1
// (:380-384, out of bounds); this follows the slice variant (:400-419).
template <class R, int AXIS, bool ATTEN>
__device__ __forceinline__ void st_force_face(R* __restrict__ face, const uint8_t* __restrict__ occ,
                                              const R* __restrict__ fc, const Dims& d, R dt, R dth, R aV, int i, int j,
                                              int k) {
    if (!face_in<AXIS>(d, i, j, k)) return;
    const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
    const int step = AXIS == 0 ? 1 : (AXIS == 1 ? d.W : d.W * d.H);
    const int ci = face_cell<AXIS>(d, i, j, k);
    const int fi = face_index<AXIS>(d, i, j, k);
    const R* f = fc + (size_t)AXIS * d.W * d.H * d.D;
    // both cells' occupancy and force load together (on a domain face the low index is the face's
    // own cell and its values are not used)
    const int cl = (pos > 0 && pos < lim) ? ci - step : ci;
    const int oh = occ[ci], ol = occ[cl];
    const R fh = f[ci], fl = f[cl];
    bool ok;
    R add;
    if (pos == 0) {
        ok = !oh;
        add = dt * fh;
    } else if (pos == lim) {
        ok = !oh;  // ci already points at the last cell
        add = dt * fh;
    } else {
        ok = !ol && !oh;
        add = dth * (fh + fl);
    }
    const R f0 = face[fi];  // loaded with the occupancy, not after it
    if (ok) {
        R x = f0 + add;
        if (ATTEN) x = x * aV;
        face[fi] = x;
    }
}

// ---- advection (ZQ_CUDA_Advection3D.cu, ZQ_CUDA_MACtoRegular.cu) ----

// MAC_to_Regular4_Kernel (ZQ_CUDA_MACtoRegular.cu:9-30)
// (the fourth component is 0, or with occ the cell's occupancy, which the forward backtrace reads
// from the taps it samples)
template <class R>
__device__ __forceinline__ void st_mac_to_cell4(const R* __restrict__ u, const R* __restrict__ v,
                                                const R* __restrict__ w, vec4_t<R>* __restrict__ vel4, const Dims& d,
                                                int i, int j, int k, const uint8_t* __restrict__ occ = nullptr) {
    if (i >= d.W || j >= d.H || k >= d.D) return;
    vec4_t<R> c;
    c.x = R(0.5) * (u[uidx(d, i, j, k)] + u[uidx(d, i + 1, j, k)]);
    c.y = R(0.5) * (v[vidx(d, i, j, k)] + v[vidx(d, i, j + 1, k)]);
    c.z = R(0.5) * (w[widx(d, i, j, k)] + w[widx(d, i, j, k + 1)]);
    c.w = (occ && occ[cidx(d, i, j, k)]) ? R(1) : R(0);
    vel4[cidx(d, i, j, k)] = c;
}

// Advect_Velocity_inRegular_outRegular_Kernel (ZQ_CUDA_Advection3D.cu:76-113): backtrace from the
// cell centre through the cell grid and sample the same grid at the end point.  Output SoA
// (adv[0..N) x, [N..2N) y, [2N..3N) z) so each face pass reads one component plane.
template <class R>
__device__ __forceinline__ void st_advect_cell(const vec4_t<R>* __restrict__ vel4, const uint8_t* __restrict__ occ,
                                               R* __restrict__ adv, const Dims& d, const StepScale<R>& h, int steps,
                                               R dtt, int i, int j,
                                               int k) {
    if (i >= d.W || j >= d.H || k >= d.D) return;
    const RegularVelCached<R> V{vel4, d, {}};
    const R3<R> last = backtrace<R>(V, occ, d, h, steps, dtt, i + R(0.5), j + R(0.5), k + R(0.5));
    const R3<R> s = V(last.x, last.y, last.z);
    const int o = cidx(d, i, j, k);
    const int N = d.W * d.H * d.D;
    adv[o] = s.x;
    adv[N + o] = s.y;
    adv[2 * N + o] = s.z;
}

// Regular4_to_MAC_{u,v,w}_Kernel (ZQ_CUDA_MACtoRegular.cu:32-99) fused with
// Apply_Advect_Velocity_Result_{Open,Closed}_* (ZQ_CUDA_Advection3D.cu:471-650): the face value is
// the average of the two advected cells (a domain face copies its inside cell) and is stored only
// where the face's cells are fluid.
template <class R, int AXIS>
__device__ __forceinline__ void st_cells_to_face(const R* __restrict__ adv, const uint8_t* __restrict__ occ,
                                                 R* __restrict__ face, const Dims& d, int is_open, int i, int j,
                                                 int k) {
    if (!face_in<AXIS>(d, i, j, k)) return;
    const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
    const int step = AXIS == 0 ? 1 : (AXIS == 1 ? d.W : d.W * d.H);
    const int ci = face_cell<AXIS>(d, i, j, k);
    const R* a = adv + (size_t)AXIS * d.W * d.H * d.D;
    const int cl = (pos > 0 && pos < lim) ? ci - step : ci;  // as in st_force_face
    const int oh = occ[ci], ol = occ[cl];
    const R ah = a[ci], al = a[cl];
    bool ok;
    R val;
    if (pos == 0 || pos == lim) {
        ok = is_open && !oh;
        val = ah;
    } else {
        ok = !oh && !ol;
        val = R(0.5) * (ah + al);
    }
    if (ok) face[face_index<AXIS>(d, i, j, k)] = val;
}

// Advect_Scalar_{Regular,MAC}_Velocity_Kernel (:379-469): backtrace from the cell centre, sample T
// and rho (linear, clamp) at the end point.  Every cell is written, solid ones included.
template <class R, class Vel>
__device__ __forceinline__ void st_advect_scalar(const Vel& V, const uint8_t* __restrict__ occ,
                                                 const R* __restrict__ Tin, const R* __restrict__ Rin,
                                                 R* __restrict__ Tout, R* __restrict__ Rout, const Dims& d,
                                                 const StepScale<R>& h,
                                                 int steps, R dtt, int i, int j, int k) {
    if (i >= d.W || j >= d.H || k >= d.D) return;
    const R3<R> last = backtrace<R>(V, occ, d, h, steps, dtt, i + R(0.5), j + R(0.5), k + R(0.5));
    const R x = last.x - R(0.5), y = last.y - R(0.5), z = last.z - R(0.5);
    const int o = cidx(d, i, j, k);
    Tout[o] = sample3<R>(Tin, d.W, d.H, d.D, x, y, z);
    Rout[o] = sample3<R>(Rin, d.W, d.H, d.D, x, y, z);
}

// ---- projection (red-black SOR, FP32) ----

__device__ __forceinline__ double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ double wave_max(double x) {
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    return x;
}

// Sum (or max) over a 256-thread block; the result is valid in thread 0.
template <bool MAX>
__device__ __forceinline__ double block_reduce(double x, double* lds4) {
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    x = MAX ? wave_max(x) : wave_sum(x);
    if ((t & 63) == 0) lds4[t >> 6] = x;
    __syncthreads();
    double r = 0.0;
    if (t == 0) r = MAX ? fmax(fmax(lds4[0], lds4[1]), fmax(lds4[2], lds4[3])) : ((lds4[0] + lds4[1]) + lds4[2]) + lds4[3];
    return r;
}

// Fold per-block (sum, count) pairs in index order with 1024 threads; mean = sum/count when
// want != 0 (0 otherwise).  ss, sc: 1024 doubles of LDS each.
__device__ __forceinline__ void st_fold_mean(const double* __restrict__ part, int nblocks, float* __restrict__ mean32,
                                             double* __restrict__ mean64, int want, double* ss, double* sc) {
    const int t = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (int i = t; i < nblocks; i += 1024) {
        s += part[2 * i];
        c += part[2 * i + 1];
    }
    ss[t] = s;
    sc[t] = c;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h) {
            ss[t] += ss[t + h];
            sc[t] += sc[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double m = (want && sc[0] > 0.0) ? ss[0] / sc[0] : 0.0;
        if (mean32) *mean32 = (float)m;
        if (mean64) *mean64 = m;
    }
}


// 1/n for n = 1..6, each correctly rounded (the same values as 1.0f/(float)n)
__device__ __forceinline__ float rcp_count(int n) {
    return n == 1 ? 1.0f : n == 2 ? 0.5f : n == 3 ? (1.0f / 3.0f) : n == 4 ? 0.25f : n == 5 ? (1.0f / 5.0f)
                                                                                            : (1.0f / 6.0f);
}

// Plain (FP32) projection set-up of one cell: neighbour count, divergence (the reference's unscaled
// sum, ZQ_SmokeSimulationUtils3D.cpp:2403-2405, left to right), p = 0 at uncounted cells; the
// cell's divergence and count for the closed-domain mean go to dsum/dcnt.
__device__ __forceinline__ void st_proj_setup_plain(const float* __restrict__ u, const float* __restrict__ v,
                                                    const float* __restrict__ w, const uint8_t* __restrict__ occ,
                                                    int open_mask, uint8_t* __restrict__ nc, float* __restrict__ b,
                                                    float* __restrict__ p, const Dims& d, int i, int j, int k,
                                                    double& dsum, double& dcnt) {
    if (i >= d.W || j >= d.H || k >= d.D) return;
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
    const float dv = u1 - u0 + v1 - v0 + w1 - w0;
    b[o] = fluid ? dv : 0.0f;
    if (!n) p[pidx(d, i, j, k)] = 0.0f;
    if (fluid) {
        dsum = (double)dv;
        dcnt = 1.0;
    }
}

// One colour of a red-black SOR iteration, in place, for the cell x = 2t + ((colour+j+k)&1) of
// row (j, k).
__device__ __forceinline__ void st_sor_half(const uint8_t* __restrict__ nc, const float* __restrict__ b,
                                            const float* __restrict__ mean32, float* __restrict__ p, const Dims& d,
                                            float omega, float omw, int colour, int t, int j, int k) {
    const int i = 2 * t + ((colour + j + k) & 1);
    if (i >= d.W || j >= d.H || k >= d.D) return;
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
// st_corrected_face returns the value face (AXIS, i, j, k) holds after the correction (its own value
// when the face is not corrected): st_correct stores it, and the mixed projection's divergence check
// predicts the corrected field's divergence from it before anything is stored.
template <int AXIS, class P, class F>
__device__ __forceinline__ F st_corrected_face(const uint8_t* __restrict__ occ, int open_mask, const P* __restrict__ p,
                                               const F* __restrict__ face, const Dims& d, int i, int j, int k) {
    const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
    const int step = AXIS == 0 ? 1 : (AXIS == 1 ? d.W : d.W * d.H);
    const int pstep = AXIS == 0 ? 1 : (AXIS == 1 ? d.W + 2 : (d.W + 2) * (d.H + 2));
    const int lo_bit = AXIS == 0 ? 1 : (AXIS == 1 ? 4 : 16);
    const int ci = face_cell<AXIS>(d, i, j, k);
    const int oh = occ[ci], ol = occ[(pos > 0 && pos < lim) ? ci - step : ci];  // as in st_force_face
    bool ok;
    if (pos == 0) ok = (open_mask & lo_bit) && !oh;
    else if (pos == lim) ok = (open_mask & (lo_bit << 1)) && !oh;
    else ok = !ol && !oh;
    const F f = face[face_index<AXIS>(d, i, j, k)];
    // padded index of the cell on the high side of the face (ghost when pos == lim); both pressures
    // are loaded with the occupancy (every padded index is in range), not after the test
    const int qh = pidx(d, i, j, k);
    const P ph = p[qh], pl = p[qh - pstep];
    if (!ok) return f;
    return (F)((P)f - (ph - pl));
}

template <int AXIS, class P, class F>
__device__ __forceinline__ void st_correct(const uint8_t* __restrict__ occ, int open_mask, const P* __restrict__ p,
                                           F* __restrict__ face, const Dims& d, int i, int j, int k) {
    if (!face_in<AXIS>(d, i, j, k)) return;
    face[face_index<AXIS>(d, i, j, k)] = st_corrected_face<AXIS, P, F>(occ, open_mask, p, face, d, i, j, k);
}

// ---- sources ----

// ZQ_ReinjectTemperatureDensity3D (ZQ_SmokeSimulationUtils3D.cpp:3842-3882), box sources,
// randDensity off: cells whose centre is within the half extents get rho_src and T_src.
template <class R>
__device__ __forceinline__ void st_reinject_box(R* __restrict__ T, R* __restrict__ rho, const Dims& d, float cx,
                                                float cy, float cz, float hx, float hy, float hz, float rd, float rt,
                                                int i, int j, int k) {
    if (i >= d.W || j >= d.H || k >= d.D) return;
    if (fabsf(i + 0.5f - cx) <= hx && fabsf(j + 0.5f - cy) <= hy && fabsf(k + 0.5f - cz) <= hz) {
        const int o = cidx(d, i, j, k);
        rho[o] = (R)(rd * 1.0f);
        T[o] = (R)rt;
    }
}

// Sphere source (not in the reference): centre and squared radius in cells.
template <class R>
__device__ __forceinline__ void st_reinject_sphere(R* __restrict__ T, R* __restrict__ rho, const Dims& d, float x0,
                                                   float y0, float z0, float r2, float rd, float rt, int i, int j,
                                                   int k) {
    if (i >= d.W || j >= d.H || k >= d.D) return;
    const float dx = i + 0.5f - x0, dy = j + 0.5f - y0, dz = k + 0.5f - z0;
    if (dx * dx + dy * dy + dz * dz <= r2) {
        const int o = cidx(d, i, j, k);
        rho[o] = (R)rd;
        T[o] = (R)rt;
    }
}

}  // namespace zq3d
