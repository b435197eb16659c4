// kernels_advect.hip -- semi-Lagrangian advection of velocity, temperature and density (gfx950).
//
// Reference: ZQ_CUDA_Advection3D.cu binds every field to a CUDA texture (a cudaArray copy per call)
// and samples with the hardware filter.  Here fields stay in place in HBM and the kernels do the
// trilinear filter in FP32 (zq3d_device.h).  The cell-centred velocity is a float4 grid so each
// of the 8 taps of a sample is one 16-byte load; the backtrace's taps are neighbours of each
// other and of the neighbouring lanes', so they are served by L1/L2, and HBM sees roughly one
// read of the float4 grid per kernel.
#include <algorithm>

#include "zq3d_device.h"
#include "zq3d_kernels.h"
#include "zq3d_stencils.h"

namespace zq3d {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid_for(int nx, int ny, int nz) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz); }

// The cell / face bodies live in zq3d_stencils.h (shared with the ensemble kernels).
template <class R>
__global__ __launch_bounds__(BX* BY) void k_mac_to_cell4(const R* __restrict__ u, const R* __restrict__ v,
                                                         const R* __restrict__ w, vec4_t<R>* __restrict__ vel4,
                                                         Dims d, const uint8_t* __restrict__ occ) {
    st_mac_to_cell4<R>(u, v, w, vel4, d, blockIdx.x * BX + threadIdx.x, blockIdx.y * BY + threadIdx.y, blockIdx.z,
                       occ);
}

template <class R>
__global__ __launch_bounds__(BX* BY) void k_advect_cells(const vec4_t<R>* __restrict__ vel4,
                                                         const uint8_t* __restrict__ occ, R* __restrict__ adv,
                                                         Dims d, StepScale<R> h, int steps, R dtt) {
    st_advect_cell<R>(vel4, occ, adv, d, h, steps, dtt, blockIdx.x * BX + threadIdx.x, blockIdx.y * BY + threadIdx.y,
                      blockIdx.z);
}

// The three face grids from the advected cells in one launch: thread (i, j) walks a chunk of
// APPLY_RZ planes and writes each cell's low u, v and w faces (the last column / row / plane also the
// domain's high face), every face with st_cells_to_face's value.  The w face below plane k reads the
// cell of plane k-1 that the same thread read one iteration earlier.
constexpr int APPLY_RZ = 16;
template <class R>
__global__ __launch_bounds__(BX* BY) void k_cells_to_faces_apply(const R* __restrict__ adv,
                                                                 const uint8_t* __restrict__ occ, R* __restrict__ u,
                                                                 R* __restrict__ v, R* __restrict__ w, Dims d,
                                                                 int is_open) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    if (i >= d.W || j >= d.H) return;
    const int k1 = min((int)(blockIdx.z + 1) * APPLY_RZ, d.D);
    for (int k = blockIdx.z * APPLY_RZ; k < k1; k++) {
        st_cells_to_face<R, 0>(adv, occ, u, d, is_open, i, j, k);
        if (i == d.W - 1) st_cells_to_face<R, 0>(adv, occ, u, d, is_open, i + 1, j, k);
        st_cells_to_face<R, 1>(adv, occ, v, d, is_open, i, j, k);
        if (j == d.H - 1) st_cells_to_face<R, 1>(adv, occ, v, d, is_open, i, j + 1, k);
        st_cells_to_face<R, 2>(adv, occ, w, d, is_open, i, j, k);
        if (k == d.D - 1) st_cells_to_face<R, 2>(adv, occ, w, d, is_open, i, j, k + 1);
    }
}

template <class R>
void launch_cells_to_faces_apply(const Dims& d, const R* adv, const uint8_t* occ, R* u, R* v, R* w, int is_open,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_cells_to_faces_apply<R>, grid_for(d.W, d.H, (d.D + APPLY_RZ - 1) / APPLY_RZ), dim3(BX, BY),
                       0, s, adv, occ, u, v, w, d, is_open);
}

// Advect_Velocity_inRegular_outMAC_* (:115-230) and inMAC_outMAC_* (:232-377): backtrace from the
// face centre; the face's own component sampled at the end point goes to a scratch face grid.
template <class R, int AXIS, class Vel>
__global__ __launch_bounds__(BX* BY) void k_advect_faces(Vel V, const uint8_t* __restrict__ occ, R* __restrict__ out,
                                                         Dims d, StepScale<R> h, int steps, R dtt) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    const R px = AXIS == 0 ? (R)i : i + R(0.5);
    const R py = AXIS == 1 ? (R)j : j + R(0.5);
    const R pz = AXIS == 2 ? (R)k : k + R(0.5);
    const R3<R> last = backtrace<R>(V, occ, d, h, steps, dtt, px, py, pz);
    const R3<R> s = V(last.x, last.y, last.z);
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    out[fi] = AXIS == 0 ? s.x : (AXIS == 1 ? s.y : s.z);
}

// BFECC passes: backtrace with V, sample the input field I (another grid than the velocity) at the
// end point.  Faces take I's matching component; cells take all of it, as float4 (w = 0) or as the
// SoA scratch the face apply reads.
template <class R, int AXIS, class Vel, class In>
__global__ __launch_bounds__(BX* BY) void k_advect_faces_in(Vel V, In I, const uint8_t* __restrict__ occ,
                                                            R* __restrict__ out, Dims d, StepScale<R> h, int steps, R dtt) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    const R px = AXIS == 0 ? (R)i : i + R(0.5);
    const R py = AXIS == 1 ? (R)j : j + R(0.5);
    const R pz = AXIS == 2 ? (R)k : k + R(0.5);
    const R3<R> last = backtrace<R>(V, occ, d, h, steps, dtt, px, py, pz);
    const R3<R> s = I(last.x, last.y, last.z);
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    out[fi] = AXIS == 0 ? s.x : (AXIS == 1 ? s.y : s.z);
}

template <class R, class Vel, class In, bool SOA>
__global__ __launch_bounds__(BX* BY) void k_advect_cells_in(Vel V, In I, const uint8_t* __restrict__ occ,
                                                            void* __restrict__ out, Dims d, StepScale<R> h, int steps, R dtt) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const R3<R> last = backtrace<R>(V, occ, d, h, steps, dtt, i + R(0.5), j + R(0.5), k + R(0.5));
    const R3<R> s = I(last.x, last.y, last.z);
    const int o = cidx(d, i, j, k);
    if (SOA) {
        R* a = (R*)out;
        const int N = d.W * d.H * d.D;
        a[o] = s.x;
        a[N + o] = s.y;
        a[2 * N + o] = s.z;
    } else {
        vec4_t<R> r;
        r.x = s.x, r.y = s.y, r.z = s.z, r.w = R(0);
        ((vec4_t<R>*)out)[o] = r;
    }
}

// Input_Increment (ZQ_CUDA_Advection3D.cu:56-74): phi~ = phi*1.5 - phi^*0.5 evaluated in double
// (the literals are double), rounded once.  keep (nullable) receives phi first.
template <class R>
__global__ __launch_bounds__(256) void k_bfecc_increment(R* __restrict__ phi, const R* __restrict__ back,
                                                         R* __restrict__ keep, R* __restrict__ out, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const R a = phi[t];
    if (keep) keep[t] = a;
    out[t] = (R)((double)a * 1.5 - (double)back[t] * 0.5);
}

// Masked copy scratch -> face (Apply_Advect_Velocity_Result_*, :471-650)
template <class R, int AXIS>
__global__ __launch_bounds__(BX* BY) void k_apply_faces(const R* __restrict__ src, const uint8_t* __restrict__ occ,
                                                        R* __restrict__ face, Dims d, int is_open) {
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
    const int oh = occ[ci], ol = occ[(pos > 0 && pos < lim) ? ci - step : ci];  // both load together
    const bool ok = (pos == 0 || pos == lim) ? (is_open && !oh) : (!oh && !ol);
    if (ok) {
        const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
        face[fi] = src[fi];
    }
}

template <class R, class Vel>
__global__ __launch_bounds__(BX* BY) void k_advect_scalar(Vel V, const uint8_t* __restrict__ occ,
                                                          const R* __restrict__ Tin, const R* __restrict__ Rin,
                                                          R* __restrict__ Tout, R* __restrict__ Rout, Dims d, StepScale<R> h,
                                                          int steps, R dtt) {
    st_advect_scalar<R, Vel>(V, occ, Tin, Rin, Tout, Rout, d, h, steps, dtt, blockIdx.x * BX + threadIdx.x,
                             blockIdx.y * BY + threadIdx.y, blockIdx.z);
}

}  // namespace

template <class R>
void launch_mac_to_cell4(const Dims& d, const R* u, const R* v, const R* w, vec4_t<R>* vel4, hipStream_t s,
                         const uint8_t* occ) {
    hipLaunchKernelGGL(k_mac_to_cell4<R>, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, u, v, w, vel4, d, occ);
}

size_t bfecc_scratch_values(const Dims& d) {
    const size_t N = (size_t)d.W * d.H * d.D;
    const size_t NF = (size_t)(d.W + 1) * d.H * d.D + (size_t)d.W * (d.H + 1) * d.D + (size_t)d.W * d.H * (d.D + 1);
    return std::max(8 * N, 2 * NF);
}

namespace {

template <class R, class Vel, class In>
void faces_in(const Dims& d, const Vel& V, const In& I, const uint8_t* occ, R* ou, R* ov, R* ow, const AdvParams& ap,
              hipStream_t s) {
    const dim3 blk(BX, BY);
    const StepScale<R> h = step_scale<R>(ap.h);
    const R dtt = (R)ap.dtt;
    hipLaunchKernelGGL((k_advect_faces_in<R, 0, Vel, In>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, V, I, occ, ou, d, h,
                       ap.steps, dtt);
    hipLaunchKernelGGL((k_advect_faces_in<R, 1, Vel, In>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, V, I, occ, ov, d, h,
                       ap.steps, dtt);
    hipLaunchKernelGGL((k_advect_faces_in<R, 2, Vel, In>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, V, I, occ, ow, d, h,
                       ap.steps, dtt);
}

template <class R, bool SOA, class Vel, class In>
void cells_in(const Dims& d, const Vel& V, const In& I, const uint8_t* occ, void* out, const AdvParams& ap,
              hipStream_t s) {
    hipLaunchKernelGGL((k_advect_cells_in<R, Vel, In, SOA>), grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, V, I, occ,
                       out, d, step_scale<R>(ap.h), ap.steps, (R)ap.dtt);
}

template <class R>
void increment(R* phi, const R* back, R* keep, R* out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_bfecc_increment<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, phi, back, keep, out,
                       n);
}

// The three BFECC velocity modes (ZQ_CUDA_Advection3D.cu:1110-1182, 1240-1336, 1389-1474).  bf holds
// bfecc_scratch_values(d) values; scratch the Nu+Nv+Nw face scratch; vel4 the cell velocity.
template <class R>
void advect_velocity_bfecc(const Dims& d, R* u, R* v, R* w, const uint8_t* occ, int is_open, int mode,
                           const AdvParams& ap, vec4_t<R>* vel4, R* scratch, R* bf, hipStream_t s) {
    using V4 = vec4_t<R>;
    const dim3 blk(BX, BY);
    const size_t N = (size_t)d.W * d.H * d.D;
    const size_t Nu = (size_t)(d.W + 1) * d.H * d.D, Nv = (size_t)d.W * (d.H + 1) * d.D;
    const size_t Nw = (size_t)d.W * d.H * (d.D + 1), NF = Nu + Nv + Nw;
    R* ou = scratch;
    R* ov = scratch + Nu;
    R* ow = scratch + Nu + Nv;
    if (mode == 1) {  // inReg_outReg: vel4 = phi; A = adv(phi; v); B = adv(A; -v); B = phi~; adv(phi~; v) -> SoA
        V4* A = (V4*)bf;
        V4* B = (V4*)bf + N;
        launch_mac_to_cell4<R>(d, u, v, w, vel4, s, occ);
        const RegularVelCached<R> Vf{vel4, d, {}};
        const NegVel<R, RegularVelCached<R>> Vb{Vf};
        cells_in<R, false>(d, Vf, RegularVel<R>{vel4, d}, occ, A, ap, s);
        cells_in<R, false>(d, Vb, RegularVel<R>{A, d}, occ, B, ap, s);
        increment<R>((R*)vel4, (const R*)B, nullptr, (R*)B, 4 * N, s);
        R* adv = scratch;  // 3N values <= Nu+Nv+Nw
        cells_in<R, true>(d, Vf, RegularVel<R>{B, d}, occ, adv, ap, s);
        launch_cells_to_faces_apply<R>(d, adv, occ, u, v, w, is_open, s);
        return;
    }
    if (mode == 3) {  // inMAC_outMAC: F1 = adv(u; u); F2 = adv(F1; -u); keep u, u = u~; adv(u~; kept) -> F1
        R* F2 = bf;
        R* K = bf + NF;
        const MacVel<R> Vf{u, v, w, d};
        faces_in<R>(d, Vf, Vf, occ, ou, ov, ow, ap, s);
        faces_in<R>(d, NegVel<R, MacVel<R>>{Vf}, MacVel<R>{ou, ov, ow, d}, occ, F2, F2 + Nu, F2 + Nu + Nv, ap, s);
        increment<R>(u, F2, K, u, Nu, s);
        increment<R>(v, F2 + Nu, K + Nu, v, Nv, s);
        increment<R>(w, F2 + Nu + Nv, K + Nu + Nv, w, Nw, s);
        faces_in<R>(d, MacVel<R>{K, K + Nu, K + Nu + Nv, d}, MacVel<R>{u, v, w, d}, occ, ou, ov, ow, ap, s);
    } else {  // mode 5, inReg_outMAC: F = adv(phi; v); C = cells(F); F = adv(C; -v); C = cells(F); C = phi~; F = adv(C; v)
        V4* C = (V4*)bf;
        launch_mac_to_cell4<R>(d, u, v, w, vel4, s, occ);
        const RegularVelCached<R> Vf{vel4, d, {}};
        faces_in<R>(d, Vf, RegularVel<R>{vel4, d}, occ, ou, ov, ow, ap, s);
        launch_mac_to_cell4<R>(d, ou, ov, ow, C, s);
        faces_in<R>(d, NegVel<R, RegularVelCached<R>>{Vf}, RegularVel<R>{C, d}, occ, ou, ov, ow, ap, s);
        launch_mac_to_cell4<R>(d, ou, ov, ow, C, s);
        increment<R>((R*)vel4, (const R*)C, nullptr, (R*)C, 4 * N, s);
        faces_in<R>(d, Vf, RegularVel<R>{C, d}, occ, ou, ov, ow, ap, s);
    }
    hipLaunchKernelGGL((k_apply_faces<R, 0>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, ou, occ, u, d, is_open);
    hipLaunchKernelGGL((k_apply_faces<R, 1>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, ov, occ, v, d, is_open);
    hipLaunchKernelGGL((k_apply_faces<R, 2>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, ow, occ, w, d, is_open);
}

}  // namespace

template <class R>
void launch_advect_velocity(const Dims& d, R* u, R* v, R* w, const uint8_t* occ, int is_open, int mode,
                            const AdvParams& ap, vec4_t<R>* vel4, R* scratch, hipStream_t s, R* bfecc) {
    if (mode == 1 || mode == 3 || mode == 5) {
        advect_velocity_bfecc<R>(d, u, v, w, occ, is_open, mode, ap, vel4, scratch, bfecc, s);
        return;
    }
    const dim3 blk(BX, BY);
    const size_t Nu = (size_t)(d.W + 1) * d.H * d.D, Nv = (size_t)d.W * (d.H + 1) * d.D;
    R* ou = scratch;
    R* ov = scratch + Nu;
    R* ow = scratch + Nu + Nv;
    const StepScale<R> h = step_scale<R>(ap.h);
    const R dtt = (R)ap.dtt;
    if (mode == 0) {
        launch_mac_to_cell4<R>(d, u, v, w, vel4, s, occ);
        R* adv = scratch;  // 3N values <= Nu+Nv+Nw
        hipLaunchKernelGGL(k_advect_cells<R>, grid_for(d.W, d.H, d.D), blk, 0, s, vel4, occ, adv, d, h, ap.steps, dtt);
        launch_cells_to_faces_apply<R>(d, adv, occ, u, v, w, is_open, s);
        return;
    }
    if (mode == 4) {
        launch_mac_to_cell4<R>(d, u, v, w, vel4, s, occ);
        const RegularVelCached<R> V{vel4, d, {}};
        hipLaunchKernelGGL((k_advect_faces<R, 0, RegularVelCached<R>>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, V, occ,
                           ou, d, h, ap.steps, dtt);
        hipLaunchKernelGGL((k_advect_faces<R, 1, RegularVelCached<R>>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, V, occ,
                           ov, d, h, ap.steps, dtt);
        hipLaunchKernelGGL((k_advect_faces<R, 2, RegularVelCached<R>>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, V, occ,
                           ow, d, h, ap.steps, dtt);
    } else {  // mode 2: inMAC_outMAC
        const MacVel<R> V{u, v, w, d};
        hipLaunchKernelGGL((k_advect_faces<R, 0, MacVel<R>>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, V, occ, ou, d,
                           h, ap.steps, dtt);
        hipLaunchKernelGGL((k_advect_faces<R, 1, MacVel<R>>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, V, occ, ov, d,
                           h, ap.steps, dtt);
        hipLaunchKernelGGL((k_advect_faces<R, 2, MacVel<R>>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, V, occ, ow, d,
                           h, ap.steps, dtt);
    }
    hipLaunchKernelGGL((k_apply_faces<R, 0>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, ou, occ, u, d, is_open);
    hipLaunchKernelGGL((k_apply_faces<R, 1>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, ov, occ, v, d, is_open);
    hipLaunchKernelGGL((k_apply_faces<R, 2>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, ow, occ, w, d, is_open);
}

template <class R>
void launch_advect_scalar(const Dims& d, const R* u, const R* v, const R* w, const uint8_t* occ, const R* Tin,
                          const R* Rin, R* Tout, R* Rout, int mac, const AdvParams& ap, vec4_t<R>* vel4,
                          hipStream_t s) {
    const dim3 blk(BX, BY);
    const StepScale<R> h = step_scale<R>(ap.h);
    const R dtt = (R)ap.dtt;
    if (mac) {
        const MacVel<R> V{u, v, w, d};
        hipLaunchKernelGGL((k_advect_scalar<R, MacVel<R>>), grid_for(d.W, d.H, d.D), blk, 0, s, V, occ, Tin, Rin, Tout,
                           Rout, d, h, ap.steps, dtt);
    } else {
        launch_mac_to_cell4<R>(d, u, v, w, vel4, s, occ);
        const RegularVelCached<R> V{vel4, d, {}};
        hipLaunchKernelGGL((k_advect_scalar<R, RegularVelCached<R>>), grid_for(d.W, d.H, d.D), blk, 0, s, V, occ, Tin,
                           Rin, Tout, Rout, d, h, ap.steps, dtt);
    }
}

#define ZQ3D_ADVECT_INSTANTIATE(R)                                                                                     \
    template void launch_mac_to_cell4<R>(const Dims&, const R*, const R*, const R*, vec4_t<R>*, hipStream_t,           \
                                         const uint8_t*);                                                              \
    template void launch_advect_velocity<R>(const Dims&, R*, R*, R*, const uint8_t*, int, int, const AdvParams&,       \
                                            vec4_t<R>*, R*, hipStream_t, R*);                                          \
    template void launch_advect_scalar<R>(const Dims&, const R*, const R*, const R*, const uint8_t*, const R*,         \
                                          const R*, R*, R*, int, const AdvParams&, vec4_t<R>*, hipStream_t);
ZQ3D_ADVECT_INSTANTIATE(float)
ZQ3D_ADVECT_INSTANTIATE(double)
#undef ZQ3D_ADVECT_INSTANTIATE

}  // namespace zq3d
