/*
 * zq_oracle.cpp -- TEST INFRASTRUCTURE ONLY (see zq_oracle.h).
 *
 * Plain C++ restatement of the reference's 3D smoke step, written from the
 * reference's .cu/.cpp text (cited per function).  Built with
 * -ffp-contract=off so every float operation is a single IEEE op in the order
 * written; the HIP kernels follow the same order.  The parallel loops below
 * (OpenMP over z) never change results: every output element is computed by
 * one iteration, and reductions are per-plane partials summed in plane order.
 */
#include "zq_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Dims {
    int W, H, D;
    size_t N() const { return (size_t)W * H * D; }
    size_t Nu() const { return (size_t)(W + 1) * H * D; }
    size_t Nv() const { return (size_t)W * (H + 1) * D; }
    size_t Nw() const { return (size_t)W * H * (D + 1); }
    size_t c(int i, int j, int k) const { return ((size_t)k * H + j) * W + i; }
    size_t fu(int i, int j, int k) const { return ((size_t)k * H + j) * (W + 1) + i; }
    size_t fv(int i, int j, int k) const { return ((size_t)k * (H + 1) + j) * W + i; }
    size_t fw(int i, int j, int k) const { return ((size_t)k * H + j) * W + i; }
    size_t pp(int i, int j, int k) const { return ((size_t)(k + 1) * (H + 2) + (j + 1)) * (W + 2) + (i + 1); }
    size_t Npad() const { return (size_t)(W + 2) * (H + 2) * (D + 2); }
};

inline float lerpf(float a, float b, float t) { return a + t * (b - a); }
inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* Texel-space trilinear sample with clamp-to-edge addressing.  CUDA's
 * normalized-coordinate linear filter (ZQ_CUDA_Advection3D.cu:656-660) reads
 * texels floor(x-0.5) and +1 with weight frac(x-0.5); here x is already the
 * texel-space coordinate (the "-0.5" applied by the caller).  Weights are
 * exact FP32, not the hardware's 9-bit fixed point. */
struct Tex {
    const float* g; int nx, ny, nz; int stride; int comp;
    float at(int i, int j, int k) const { return g[(((size_t)k * ny + j) * nx + i) * stride + comp]; }
    float sample(float x, float y, float z) const {
        x = std::fmin(std::fmax(x, -1.0f), (float)nx);
        y = std::fmin(std::fmax(y, -1.0f), (float)ny);
        z = std::fmin(std::fmax(z, -1.0f), (float)nz);
        float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
        float ax = x - fx, ay = y - fy, az = z - fz;
        int i1 = clampi(i0 + 1, 0, nx - 1), j1 = clampi(j0 + 1, 0, ny - 1), k1 = clampi(k0 + 1, 0, nz - 1);
        i0 = clampi(i0, 0, nx - 1); j0 = clampi(j0, 0, ny - 1); k0 = clampi(k0, 0, nz - 1);
        float c00 = lerpf(at(i0, j0, k0), at(i1, j0, k0), ax);
        float c10 = lerpf(at(i0, j1, k0), at(i1, j1, k0), ax);
        float c01 = lerpf(at(i0, j0, k1), at(i1, j0, k1), ax);
        float c11 = lerpf(at(i0, j1, k1), at(i1, j1, k1), ax);
        float c0 = lerpf(c00, c10, ay);
        float c1 = lerpf(c01, c11, ay);
        return lerpf(c0, c1, az);
    }
};

/* Velocity sampler used by the backtrace.
 * regular: cell-centred float4 grid at pos-0.5 (ZQ_CUDA_Advection3D.cu:89-90)
 * MAC    : u at (x, y-0.5, z-0.5) on (W+1)xHxD etc. (ZQ_CUDA_Advection3D.cu:245-251) */
struct VelSampler {
    Dims d; bool mac;
    Tex r0, r1, r2;  /* regular: vel4 comps 0..2 ; MAC: u, v, w */
    void sample(float px, float py, float pz, float& vx, float& vy, float& vz) const {
        if (!mac) {
            float x = px - 0.5f, y = py - 0.5f, z = pz - 0.5f;
            vx = r0.sample(x, y, z); vy = r1.sample(x, y, z); vz = r2.sample(x, y, z);
        } else {
            vx = r0.sample(px, py - 0.5f, pz - 0.5f);
            vy = r1.sample(px - 0.5f, py, pz - 0.5f);
            vz = r2.sample(px - 0.5f, py - 0.5f, pz);
        }
    }
};

VelSampler make_regular_sampler(const Dims& d, const float* vel4) {
    VelSampler s; s.d = d; s.mac = false;
    s.r0 = Tex{vel4, d.W, d.H, d.D, 4, 0};
    s.r1 = Tex{vel4, d.W, d.H, d.D, 4, 1};
    s.r2 = Tex{vel4, d.W, d.H, d.D, 4, 2};
    return s;
}
VelSampler make_mac_sampler(const Dims& d, const float* u, const float* v, const float* w) {
    VelSampler s; s.d = d; s.mac = true;
    s.r0 = Tex{u, d.W + 1, d.H, d.D, 1, 0};
    s.r1 = Tex{v, d.W, d.H + 1, d.D, 1, 0};
    s.r2 = Tex{w, d.W, d.H, d.D + 1, 1, 0};
    return s;
}

/* Backtrace of ZQ_CUDA_Advection3D.cu:87-109 (identical in every kernel of that file):
 *   stop when pos leaves [0,W]x[0,H]x[0,D] or the occupancy texel at pos is solid;
 *   lastpos = pos; pos -= vel*dtt/h; vel = V(pos); at most `steps` passes.
 * The returned position is `lastpos`, so at most steps-1 displacements apply. */
inline void backtrace(const Dims& d, const VelSampler& V, const uint8_t* occ, float h, int steps, float dtt,
                      float px, float py, float pz, float& lx, float& ly, float& lz) {
    lx = px; ly = py; lz = pz;
    float vx, vy, vz;
    V.sample(px, py, pz, vx, vy, vz);
    const float fW = (float)d.W, fH = (float)d.H, fD = (float)d.D;
    unsigned istep = 0;
    do {
        if (!(px >= 0.0f && px <= fW && py >= 0.0f && py <= fH && pz >= 0.0f && pz <= fD)) break;
        /* point-filtered occupancy, clamp addressing: texel floor(pos) */
        int oi = clampi((int)std::floor(px), 0, d.W - 1);
        int oj = clampi((int)std::floor(py), 0, d.H - 1);
        int ok = clampi((int)std::floor(pz), 0, d.D - 1);
        if (occ[d.c(oi, oj, ok)]) break;
        lx = px; ly = py; lz = pz;
        px = px - (vx * dtt) / h;
        py = py - (vy * dtt) / h;
        pz = pz - (vz * dtt) / h;
        V.sample(px, py, pz, vx, vy, vz);
        istep++;
    } while (istep < (unsigned)steps);
}

/* Apply_Advect_Velocity_Result_{Open,Closed}_{u,v,w}_Kernel (ZQ_CUDA_Advection3D.cu:471-650) */
void apply_advect_result(const Dims& d, const uint8_t* occ, int is_open, const float* au, const float* av,
                         const float* aw, float* u, float* v, float* w) {
    const int W = d.W, H = d.H, D = d.D;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i <= W; i++) {
                bool ok;
                if (i == 0) ok = is_open && !occ[d.c(0, j, k)];
                else if (i == W) ok = is_open && !occ[d.c(W - 1, j, k)];
                else ok = !occ[d.c(i, j, k)] && !occ[d.c(i - 1, j, k)];
                if (ok) u[d.fu(i, j, k)] = au[d.fu(i, j, k)];
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j <= H; j++)
            for (int i = 0; i < W; i++) {
                bool ok;
                if (j == 0) ok = is_open && !occ[d.c(i, 0, k)];
                else if (j == H) ok = is_open && !occ[d.c(i, H - 1, k)];
                else ok = !occ[d.c(i, j, k)] && !occ[d.c(i, j - 1, k)];
                if (ok) v[d.fv(i, j, k)] = av[d.fv(i, j, k)];
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                bool ok;
                if (k == 0) ok = is_open && !occ[d.c(i, j, 0)];
                else if (k == D) ok = is_open && !occ[d.c(i, j, D - 1)];
                else ok = !occ[d.c(i, j, k)] && !occ[d.c(i, j, k - 1)];
                if (ok) w[d.fw(i, j, k)] = aw[d.fw(i, j, k)];
            }
}

inline uint32_t hash_u32(uint32_t x) {
    /* lowbias32 integer finalizer (public domain, C. Wellons) */
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

}  // namespace

extern "C" {

/* MAC_to_Regular4_Kernel (ZQ_CUDA_MACtoRegular.cu:9-30) */
void orc_mac_to_cell4(int W, int H, int D, const float* u, const float* v, const float* w, float* vel4) {
    Dims d{W, H, D};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                size_t o = d.c(i, j, k);
                vel4[o * 4 + 0] = 0.5f * (u[d.fu(i, j, k)] + u[d.fu(i + 1, j, k)]);
                vel4[o * 4 + 1] = 0.5f * (v[d.fv(i, j, k)] + v[d.fv(i, j + 1, k)]);
                vel4[o * 4 + 2] = 0.5f * (w[d.fw(i, j, k)] + w[d.fw(i, j, k + 1)]);
                vel4[o * 4 + 3] = 0.0f;
            }
}

/* Regular4_to_MAC_{u,v,w}_Kernel (ZQ_CUDA_MACtoRegular.cu:32-99): interior faces average the
 * two cells (cell[i] + cell[i-1]); domain faces copy the single inside cell. */
void orc_cell4_to_mac(int W, int H, int D, const float* vel4, float* u, float* v, float* w) {
    Dims d{W, H, D};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i <= W; i++) {
                float val;
                if (i == 0) val = vel4[d.c(0, j, k) * 4 + 0];
                else if (i == W) val = vel4[d.c(W - 1, j, k) * 4 + 0];
                else val = 0.5f * (vel4[d.c(i, j, k) * 4 + 0] + vel4[d.c(i - 1, j, k) * 4 + 0]);
                u[d.fu(i, j, k)] = val;
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j <= H; j++)
            for (int i = 0; i < W; i++) {
                float val;
                if (j == 0) val = vel4[d.c(i, 0, k) * 4 + 1];
                else if (j == H) val = vel4[d.c(i, H - 1, k) * 4 + 1];
                else val = 0.5f * (vel4[d.c(i, j, k) * 4 + 1] + vel4[d.c(i, j - 1, k) * 4 + 1]);
                v[d.fv(i, j, k)] = val;
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                float val;
                if (k == 0) val = vel4[d.c(i, j, 0) * 4 + 2];
                else if (k == D) val = vel4[d.c(i, j, D - 1) * 4 + 2];
                else val = 0.5f * (vel4[d.c(i, j, k) * 4 + 2] + vel4[d.c(i, j, k - 1) * 4 + 2]);
                w[d.fw(i, j, k)] = val;
            }
}

float orc_sample3(const float* g, int nx, int ny, int nz, float x, float y, float z) {
    Tex t{g, nx, ny, nz, 1, 0};
    return t.sample(x, y, z);
}

/* cu_AddForce3D (ZQ_CUDA_AddForce3D.cu:510-556):
 *  1. cell velocity = MAC face average (ZQlibCuda cu_MAC_to_Regular_vel, absent; taken to be the
 *     same 0.5f*(lo+hi) as MAC_to_Regular4_Kernel);
 *  2. vorticity on interior cells [1,N-2]^3, zero elsewhere (:11-34);
 *  3. normalized gradient of |w|, one-sided at the domain edges, unnormalized when 0 (:64-111);
 *  4. f = eps*h*(N x w), f_y += beta*(T - Tamb) (:161-184);  extension: f_y -= alpha*rho when
 *     alpha != 0, and f += extra_force when given;
 *  5. faces += dt*0.5*(f_l + f_r) when both cells are fluid; a domain face uses its inside cell
 *     (:211-385).  The z=0 iteration of AddForce_w_Kernel's loop reads cell z-1 out of bounds
 *     (:380-384); this follows the slice kernel (:400-419), which handles z=0 and z=D once. */
void orc_add_force(int W, int H, int D, float* u, float* v, float* w, const uint8_t* occ,
                   const float* T, const float* rho, float h, float dt, float beta, float alpha,
                   float eps, float Tamb, const float* extra_force,
                   float* vort4, float* gradn3, float* force3) {
    Dims d{W, H, D};
    const size_t N = d.N();
    const int HW = H * W;
    std::vector<float> cu(N), cv(N), cw(N), vv(N * 3, 0.0f), vs(N, 0.0f), gv(N * 3, 0.0f), f(N * 3, 0.0f);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                size_t o = d.c(i, j, k);
                cu[o] = 0.5f * (u[d.fu(i, j, k)] + u[d.fu(i + 1, j, k)]);
                cv[o] = 0.5f * (v[d.fv(i, j, k)] + v[d.fv(i, j + 1, k)]);
                cw[o] = 0.5f * (w[d.fw(i, j, k)] + w[d.fw(i, j, k + 1)]);
            }
    const float two_h = 2 * h;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < D - 1; k++)
        for (int j = 1; j < H - 1; j++)
            for (int i = 1; i < W - 1; i++) {
                size_t o = d.c(i, j, k);
                float wx = (cw[o + W] - cw[o - W] - cv[o + HW] + cv[o - HW]) / two_h;
                float wy = (cu[o + HW] - cu[o - HW] - cw[o + 1] + cw[o - 1]) / two_h;
                float wz = (cv[o + 1] - cv[o - 1] - cu[o + W] + cu[o - W]) / two_h;
                vv[o * 3 + 0] = wx; vv[o * 3 + 1] = wy; vv[o * 3 + 2] = wz;
                vs[o] = std::sqrt(wx * wx + wy * wy + wz * wz);
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                size_t o = d.c(i, j, k);
                float gx, gy, gz;
                if (i == 0) gx = vs[o + 1] - vs[o];
                else if (i == W - 1) gx = vs[o] - vs[o - 1];
                else gx = 0.5f * (vs[o + 1] - vs[o - 1]);
                if (j == 0) gy = vs[o + W] - vs[o];
                else if (j == H - 1) gy = vs[o] - vs[o - W];
                else gy = 0.5f * (vs[o + W] - vs[o - W]);
                if (k == 0) gz = vs[o + HW] - vs[o];
                else if (k == D - 1) gz = vs[o] - vs[o - HW];
                else gz = 0.5f * (vs[o + HW] - vs[o - HW]);
                float len = std::sqrt(gx * gx + gy * gy + gz * gz);
                if (len != 0.0f) { gx /= len; gy /= len; gz /= len; }
                gv[o * 3 + 0] = gx; gv[o * 3 + 1] = gy; gv[o * 3 + 2] = gz;
                const float ch = eps * h;
                const float wx = vv[o * 3 + 0], wy = vv[o * 3 + 1], wz = vv[o * 3 + 2];
                float fx = ch * (gy * wz - gz * wy);
                float fy = ch * (gz * wx - gx * wz);
                float fz = ch * (gx * wy - gy * wx);
                fy += beta * (T[o] - Tamb);
                if (alpha != 0.0f) fy -= alpha * rho[o];
                if (extra_force) { fx += extra_force[o * 3 + 0]; fy += extra_force[o * 3 + 1]; fz += extra_force[o * 3 + 2]; }
                f[o * 3 + 0] = fx; f[o * 3 + 1] = fy; f[o * 3 + 2] = fz;
            }
    const float dth = dt * 0.5f;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i <= W; i++) {
                size_t fi = d.fu(i, j, k);
                if (i == 0) { if (!occ[d.c(0, j, k)]) u[fi] += dt * f[d.c(0, j, k) * 3 + 0]; }
                else if (i == W) { if (!occ[d.c(W - 1, j, k)]) u[fi] += dt * f[d.c(W - 1, j, k) * 3 + 0]; }
                else if (!occ[d.c(i - 1, j, k)] && !occ[d.c(i, j, k)])
                    u[fi] += dth * (f[d.c(i, j, k) * 3 + 0] + f[d.c(i - 1, j, k) * 3 + 0]);
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j <= H; j++)
            for (int i = 0; i < W; i++) {
                size_t fi = d.fv(i, j, k);
                if (j == 0) { if (!occ[d.c(i, 0, k)]) v[fi] += dt * f[d.c(i, 0, k) * 3 + 1]; }
                else if (j == H) { if (!occ[d.c(i, H - 1, k)]) v[fi] += dt * f[d.c(i, H - 1, k) * 3 + 1]; }
                else if (!occ[d.c(i, j - 1, k)] && !occ[d.c(i, j, k)])
                    v[fi] += dth * (f[d.c(i, j, k) * 3 + 1] + f[d.c(i, j - 1, k) * 3 + 1]);
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                size_t fi = d.fw(i, j, k);
                if (k == 0) { if (!occ[d.c(i, j, 0)]) w[fi] += dt * f[d.c(i, j, 0) * 3 + 2]; }
                else if (k == D) { if (!occ[d.c(i, j, D - 1)]) w[fi] += dt * f[d.c(i, j, D - 1) * 3 + 2]; }
                else if (!occ[d.c(i, j, k - 1)] && !occ[d.c(i, j, k)])
                    w[fi] += dth * (f[d.c(i, j, k) * 3 + 2] + f[d.c(i, j, k - 1) * 3 + 2]);
            }
    if (vort4)
        for (size_t o = 0; o < N; o++) {
            vort4[o * 4 + 0] = vv[o * 3 + 0]; vort4[o * 4 + 1] = vv[o * 3 + 1];
            vort4[o * 4 + 2] = vv[o * 3 + 2]; vort4[o * 4 + 3] = vs[o];
        }
    if (gradn3) std::memcpy(gradn3, gv.data(), sizeof(float) * N * 3);
    if (force3) std::memcpy(force3, f.data(), sizeof(float) * N * 3);
}

/* Atten_{u,v,w}_Kernel, Atten_temperature_density_Kernel (ZQ_CUDA_Attenuation3D.cu:8-116) */
void orc_attenuation(int W, int H, int D, float* u, float* v, float* w, float* T, float* rho,
                     const uint8_t* occ, float aV, float aT, float aR) {
    Dims d{W, H, D};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i <= W; i++) {
                bool ok = (i == 0) ? !occ[d.c(0, j, k)]
                        : (i == W) ? !occ[d.c(W - 1, j, k)]
                        : (!occ[d.c(i - 1, j, k)] && !occ[d.c(i, j, k)]);
                if (ok) u[d.fu(i, j, k)] *= aV;
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j <= H; j++)
            for (int i = 0; i < W; i++) {
                bool ok = (j == 0) ? !occ[d.c(i, 0, k)]
                        : (j == H) ? !occ[d.c(i, H - 1, k)]
                        : (!occ[d.c(i, j - 1, k)] && !occ[d.c(i, j, k)]);
                if (ok) v[d.fv(i, j, k)] *= aV;
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                bool ok = (k == 0) ? !occ[d.c(i, j, 0)]
                        : (k == D) ? !occ[d.c(i, j, D - 1)]
                        : (!occ[d.c(i, j, k - 1)] && !occ[d.c(i, j, k)]);
                if (ok) w[d.fw(i, j, k)] *= aV;
            }
    const size_t N = d.N();
    for (size_t o = 0; o < N; o++) { T[o] *= aT; rho[o] *= aR; }
}

void orc_backtrace(int W, int H, int D, float h, int steps, float dtt, const float* u, const float* v,
                   const float* w, const uint8_t* occ, int mac_velocity, float px, float py, float pz,
                   float* out3) {
    Dims d{W, H, D};
    std::vector<float> vel4;
    VelSampler V;
    if (mac_velocity) V = make_mac_sampler(d, u, v, w);
    else { vel4.resize(d.N() * 4); orc_mac_to_cell4(W, H, D, u, v, w, vel4.data()); V = make_regular_sampler(d, vel4.data()); }
    backtrace(d, V, occ, h, steps, dtt, px, py, pz, out3[0], out3[1], out3[2]);
}

/* Advect_Velocity (ZQ_CUDA_Advection3D.cu:1686-1749), non-BFECC modes:
 *  inReg_outReg (:76-113, 1061-1108): MAC->cell float4, backtrace per cell, sample the cell grid
 *      at lastpos, cell->MAC average, masked apply;
 *  inMAC_outMAC (:232-377, 1184-1238): per face, backtrace through the staggered grids, sample
 *      the face's own component;
 *  inReg_outMAC (:115-230, 1338-1387): per face, backtrace through the cell grid, sample the
 *      cell grid's matching component.
 * The BFECC modes compose these (below). */
namespace {

/* Input_Increment_Kernel (ZQ_CUDA_Advection3D.cu:56-64): input*1.5 - input_star*0.5 with double
 * literals, so in double, rounded once to float.  The 4-channel kernel (:66-74) goes through float4
 * operators of the absent ZQlibCudaDefines.cuh; the same double expression is used for it (equal to
 * a contracted float evaluation except in double-rounding corner cases: parity unpinned there). */
inline float bfecc_inc(float a, float b) { return (float)((double)a * 1.5 - (double)b * 0.5); }

/* Cells: backtrace from the centre with V, sample the regular input grid In at the end point
 * (Advect_Velocity_inRegular_outRegular_Kernel, :76-113; input and velocity textures differ under BFECC). */
void adv_cells(const Dims& d, const VelSampler& V, const VelSampler& In, const uint8_t* occ, float h, int steps,
               float dtt, float* out4) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < d.D; k++)
        for (int j = 0; j < d.H; j++)
            for (int i = 0; i < d.W; i++) {
                float lx, ly, lz, sx, sy, sz;
                backtrace(d, V, occ, h, steps, dtt, i + 0.5f, j + 0.5f, k + 0.5f, lx, ly, lz);
                In.sample(lx, ly, lz, sx, sy, sz);
                const size_t o = d.c(i, j, k);
                out4[o * 4 + 0] = sx; out4[o * 4 + 1] = sy; out4[o * 4 + 2] = sz; out4[o * 4 + 3] = 0.0f;
            }
}

/* Faces: backtrace from each face centre with V, the face's component of In at the end point
 * (Advect_Velocity_inRegular_outMAC_*, :115-230; inMAC_outMAC_*, :232-377). */
void adv_faces(const Dims& d, const VelSampler& V, const VelSampler& In, const uint8_t* occ, float h, int steps,
               float dtt, float* ou, float* ov, float* ow) {
    const int W = d.W, H = d.H, D = d.D;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i <= W; i++) {
                float lx, ly, lz, sx, sy, sz;
                backtrace(d, V, occ, h, steps, dtt, (float)i, j + 0.5f, k + 0.5f, lx, ly, lz);
                In.sample(lx, ly, lz, sx, sy, sz);
                ou[d.fu(i, j, k)] = sx;
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j <= H; j++)
            for (int i = 0; i < W; i++) {
                float lx, ly, lz, sx, sy, sz;
                backtrace(d, V, occ, h, steps, dtt, i + 0.5f, (float)j, k + 0.5f, lx, ly, lz);
                In.sample(lx, ly, lz, sx, sy, sz);
                ov[d.fv(i, j, k)] = sy;
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                float lx, ly, lz, sx, sy, sz;
                backtrace(d, V, occ, h, steps, dtt, i + 0.5f, j + 0.5f, (float)k, lx, ly, lz);
                In.sample(lx, ly, lz, sx, sy, sz);
                ow[d.fw(i, j, k)] = sz;
            }
}

std::vector<float> negated(const float* a, size_t n) {
    std::vector<float> r(n);
    for (size_t i = 0; i < n; i++) r[i] = -a[i];
    return r;
}

}  // namespace

/* Advect_Velocity (ZQ_CUDA_Advection3D.cu:1696-1742).  The BFECC modes follow
 * cu_Advect_Velocity_inRegular_outRegular_BFECC (:1110-1182), ..._inMAC_outMAC_BFECC (:1240-1336) and
 * ..._inRegular_outMAC_BFECC (:1389-1474): advect phi with v, advect that with -v (the velocity
 * negated, :997-1019), phi~ = 1.5 phi - 0.5 phi^ (:56-74), advect phi~ with v, apply.  In the MAC
 * variant phi is the caller's faces, overwritten by phi~, so faces the final apply skips keep phi~. */
int orc_advect_velocity(int W, int H, int D, float h, int steps, float dtt, float* u, float* v, float* w,
                        const uint8_t* occ, int is_open, int type) {
    Dims d{W, H, D};
    const size_t N = d.N(), Nu = d.Nu(), Nv = d.Nv(), Nw = d.Nw();
    std::vector<float> ou(Nu, 0.0f), ov(Nv, 0.0f), ow(Nw, 0.0f);
    switch (type) {
    case ORC_ADV_VEL_INREG_OUTREG: {
        std::vector<float> vel4(N * 4), adv4(N * 4);
        orc_mac_to_cell4(W, H, D, u, v, w, vel4.data());
        const VelSampler V = make_regular_sampler(d, vel4.data());
        adv_cells(d, V, V, occ, h, steps, dtt, adv4.data());
        orc_cell4_to_mac(W, H, D, adv4.data(), ou.data(), ov.data(), ow.data());
        break;
    }
    case ORC_ADV_VEL_INREG_OUTREG_BFECC: {
        std::vector<float> vel4(N * 4), a4(N * 4), b4(N * 4);
        orc_mac_to_cell4(W, H, D, u, v, w, vel4.data());
        const VelSampler V = make_regular_sampler(d, vel4.data());
        adv_cells(d, V, V, occ, h, steps, dtt, a4.data());
        const std::vector<float> nvel = negated(vel4.data(), N * 4);
        adv_cells(d, make_regular_sampler(d, nvel.data()), make_regular_sampler(d, a4.data()), occ, h, steps, dtt,
                  b4.data());
        for (size_t i = 0; i < N * 4; i++) b4[i] = bfecc_inc(vel4[i], b4[i]);
        adv_cells(d, V, make_regular_sampler(d, b4.data()), occ, h, steps, dtt, a4.data());
        orc_cell4_to_mac(W, H, D, a4.data(), ou.data(), ov.data(), ow.data());
        break;
    }
    case ORC_ADV_VEL_INMAC_OUTMAC: {
        const VelSampler V = make_mac_sampler(d, u, v, w);
        adv_faces(d, V, V, occ, h, steps, dtt, ou.data(), ov.data(), ow.data());
        break;
    }
    case ORC_ADV_VEL_INMAC_OUTMAC_BFECC: {
        const VelSampler V = make_mac_sampler(d, u, v, w);
        std::vector<float> fu(Nu), fv(Nv), fw(Nw);
        adv_faces(d, V, V, occ, h, steps, dtt, fu.data(), fv.data(), fw.data());
        const std::vector<float> nu = negated(u, Nu), nv = negated(v, Nv), nw = negated(w, Nw);
        adv_faces(d, make_mac_sampler(d, nu.data(), nv.data(), nw.data()), make_mac_sampler(d, fu.data(), fv.data(), fw.data()),
                  occ, h, steps, dtt, ou.data(), ov.data(), ow.data());
        const std::vector<float> su(u, u + Nu), sv(v, v + Nv), sw(w, w + Nw);  // the velocity texture, before the increment
        for (size_t i = 0; i < Nu; i++) u[i] = bfecc_inc(u[i], ou[i]);
        for (size_t i = 0; i < Nv; i++) v[i] = bfecc_inc(v[i], ov[i]);
        for (size_t i = 0; i < Nw; i++) w[i] = bfecc_inc(w[i], ow[i]);
        adv_faces(d, make_mac_sampler(d, su.data(), sv.data(), sw.data()), make_mac_sampler(d, u, v, w), occ, h, steps,
                  dtt, ou.data(), ov.data(), ow.data());
        break;
    }
    case ORC_ADV_VEL_INREG_OUTMAC: {
        std::vector<float> vel4(N * 4);
        orc_mac_to_cell4(W, H, D, u, v, w, vel4.data());
        const VelSampler V = make_regular_sampler(d, vel4.data());
        adv_faces(d, V, V, occ, h, steps, dtt, ou.data(), ov.data(), ow.data());
        break;
    }
    case ORC_ADV_VEL_INREG_OUTMAC_BFECC: {
        std::vector<float> vel4(N * 4), c4(N * 4);
        orc_mac_to_cell4(W, H, D, u, v, w, vel4.data());
        const VelSampler V = make_regular_sampler(d, vel4.data());
        adv_faces(d, V, V, occ, h, steps, dtt, ou.data(), ov.data(), ow.data());
        orc_mac_to_cell4(W, H, D, ou.data(), ov.data(), ow.data(), c4.data());
        const std::vector<float> nvel = negated(vel4.data(), N * 4);
        adv_faces(d, make_regular_sampler(d, nvel.data()), make_regular_sampler(d, c4.data()), occ, h, steps, dtt,
                  ou.data(), ov.data(), ow.data());
        orc_mac_to_cell4(W, H, D, ou.data(), ov.data(), ow.data(), c4.data());
        for (size_t i = 0; i < N * 4; i++) c4[i] = bfecc_inc(vel4[i], c4[i]);
        adv_faces(d, V, make_regular_sampler(d, c4.data()), occ, h, steps, dtt, ou.data(), ov.data(), ow.data());
        break;
    }
    default:
        return -1;
    }
    apply_advect_result(d, occ, is_open, ou.data(), ov.data(), ow.data(), u, v, w);
    return 0;
}

/* Advect_Scalar (ZQ_CUDA_Advection3D.cu:1751-1821): Reg (:379-418) and MAC (:420-469) velocity
 * sampling; T and rho sampled at lastpos-0.5 with clamp.  The BFECC modes are not handled by the
 * reference's switch and return the memset-zero outputs (:1782-1794); reproduced. */
int orc_advect_scalar(int W, int H, int D, float h, int steps, float dtt, const float* u, const float* v,
                      const float* w, const uint8_t* occ, const float* Tin, const float* Rin,
                      float* Tout, float* Rout, int type) {
    Dims d{W, H, D};
    const size_t N = d.N();
    if (type == ORC_ADV_SCA_MAC_BFECC || type == ORC_ADV_SCA_REG_BFECC) {
        std::memset(Tout, 0, sizeof(float) * N);
        std::memset(Rout, 0, sizeof(float) * N);
        return 0;
    }
    if (type != ORC_ADV_SCA_MAC && type != ORC_ADV_SCA_REG) return -1;
    std::vector<float> vel4, tT(Tin, Tin + N), tR(Rin, Rin + N);
    VelSampler V;
    if (type == ORC_ADV_SCA_MAC) V = make_mac_sampler(d, u, v, w);
    else { vel4.resize(N * 4); orc_mac_to_cell4(W, H, D, u, v, w, vel4.data()); V = make_regular_sampler(d, vel4.data()); }
    Tex tt{tT.data(), W, H, D, 1, 0}, tr{tR.data(), W, H, D, 1, 0};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                float lx, ly, lz;
                backtrace(d, V, occ, h, steps, dtt, i + 0.5f, j + 0.5f, k + 0.5f, lx, ly, lz);
                size_t o = d.c(i, j, k);
                Tout[o] = tt.sample(lx - 0.5f, ly - 0.5f, lz - 0.5f);
                Rout[o] = tr.sample(lx - 0.5f, ly - 0.5f, lz - 0.5f);
            }
    return 0;
}

/* Unscaled divergence, the reference's own form (ZQ_SmokeSimulationUtils3D.cpp:2403-2405):
 * u[i+1] - u[i] + v[j+1] - v[j] + w[k+1] - w[k], evaluated left to right; 0 on solid cells. */
void orc_divergence(int W, int H, int D, const float* u, const float* v, const float* w,
                    const uint8_t* occ, float* div) {
    Dims d{W, H, D};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                size_t o = d.c(i, j, k);
                div[o] = occ[o] ? 0.0f
                       : u[d.fu(i + 1, j, k)] - u[d.fu(i, j, k)] + v[d.fv(i, j + 1, k)] - v[d.fv(i, j, k)]
                         + w[d.fw(i, j, k + 1)] - w[d.fw(i, j, k)];
            }
}

/* Neighbour count of the 7-point operator: a neighbour counts when it is an in-domain fluid
 * cell, or lies outside an open domain face (p = 0 there, as in the open analog
 * ZQ_SmokeSimulationUtils3D.cpp:2824-2827).  Solid cells and walls are Neumann (not counted). */
void orc_sor_count(int W, int H, int D, const uint8_t* occ, int open_mask, uint8_t* ncount) {
    Dims d{W, H, D};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                size_t o = d.c(i, j, k);
                if (occ[o]) { ncount[o] = 0; continue; }
                int n = 0;
                n += (i > 0) ? !occ[o - 1] : ((open_mask & ORC_OPEN_XLO) != 0);
                n += (i < W - 1) ? !occ[o + 1] : ((open_mask & ORC_OPEN_XHI) != 0);
                n += (j > 0) ? !occ[o - W] : ((open_mask & ORC_OPEN_YLO) != 0);
                n += (j < H - 1) ? !occ[o + W] : ((open_mask & ORC_OPEN_YHI) != 0);
                n += (k > 0) ? !occ[o - (size_t)W * H] : ((open_mask & ORC_OPEN_ZLO) != 0);
                n += (k < D - 1) ? !occ[o + (size_t)W * H] : ((open_mask & ORC_OPEN_ZHI) != 0);
                ncount[o] = (uint8_t)n;
            }
}

/* Red-black SOR (colour (i+j+k)&1, colour 0 first) on the padded pressure.  For a cell with
 * count n > 0:  s = p[x-1]+p[x+1]+p[y-1]+p[y+1]+p[z-1]+p[z+1]  (left to right; ghosts and solid
 * cells hold 0 so they add nothing), pg = (s - b) * (1/n), p = (1-omega)*p + omega*pg.
 * With omega = 1 this is the Gauss-Seidel update p = (sum_nb p - div)/#nb of the in-tree analog
 * CPU_PoissonSolver3D_OpenOctreePoisson.cpp:416-438. */
}  // extern "C"

namespace {

uint32_t f32_bits(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return u;
}

/* Red-black SOR iterations (the one loop both entry points run).  gated: before iteration it >= 1,
 * stop when the largest |new - old| of iteration it-1 (FP32 bits) is <= thresh_bits, or, past
 * iteration 2K, when it is no smaller than that of iteration it-1-K (K = 16: the updates have
 * stopped shrinking -- FP32 rounding of the iterate bounds them from below). */
constexpr int GATE_K = 16;
int sor_iterations(int W, int H, int D, const uint8_t* ncount, const float* b, float* p, float omega, int iters,
                   bool gated, uint32_t thresh_bits) {
    Dims d{W, H, D};
    const size_t sx = 1, sy = (size_t)(W + 2), sz = (size_t)(W + 2) * (H + 2);
    float rcp[7];
    rcp[0] = 0.0f;
    for (int n = 1; n <= 6; n++) rcp[n] = 1.0f / (float)n;
    const float omw = 1.0f - omega;
    std::vector<uint32_t> hist;  // largest update of each iteration run (gated)
    int it = 0;
    for (; it < iters; it++) {
        if (gated && it > 0) {
            const uint32_t g1 = hist[it - 1];
            if (g1 <= thresh_bits || (it > 2 * GATE_K && g1 >= hist[it - 1 - GATE_K])) break;
        }
        uint32_t dm = 0;
        for (int color = 0; color < 2; color++) {
#pragma omp parallel for schedule(static) reduction(max : dm)
            for (int k = 0; k < D; k++)
                for (int j = 0; j < H; j++)
                    for (int i = (color + j + k) & 1; i < W; i += 2) {
                        size_t o = d.c(i, j, k);
                        int n = ncount[o];
                        if (!n) continue;
                        size_t q = d.pp(i, j, k);
                        const float old = p[q];
                        float s = p[q - sx] + p[q + sx] + p[q - sy] + p[q + sy] + p[q - sz] + p[q + sz];
                        float pg = (s - b[o]) * rcp[n];
                        p[q] = omw * p[q] + omega * pg;
                        if (gated) {
                            const uint32_t db = f32_bits(std::fabs(p[q] - old));
                            dm = db > dm ? db : dm;
                        }
                    }
        }
        if (gated) hist.push_back(dm);
    }
    return it;
}

}  // namespace

extern "C" {

void orc_sor(int W, int H, int D, const uint8_t* ncount, const float* b, float* p, float omega, int iters) {
    sor_iterations(W, H, D, ncount, b, p, omega, iters, false, 0u);
}

int orc_sor_gated(int W, int H, int D, const uint8_t* ncount, const float* b, float* p, float omega, int iters,
                  float thresh) {
    return sor_iterations(W, H, D, ncount, b, p, omega, iters, true, f32_bits(thresh));
}

float orc_gate_threshold(double tol, float omega) { return (float)(0.75 * tol * (double)omega / 6.0); }

}  // extern "C"

namespace {

/* Gradient correction.  Interior faces with two fluid cells: u -= p_r - p_l; a domain face of a
 * fluid cell is corrected only when that face is open, against p = 0 outside
 * (ZQ_SmokeSimulationUtils3D.cpp:2815-2861 open, :2888-2922 closed).  P is float or double. */
template <class P>
void correct_faces(const Dims& d, const uint8_t* occ, int open_mask, const P* p, float* u, float* v, float* w) {
    const int W = d.W, H = d.H, D = d.D;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i <= W; i++) {
                bool ok = (i == 0) ? ((open_mask & ORC_OPEN_XLO) && !occ[d.c(0, j, k)])
                        : (i == W) ? ((open_mask & ORC_OPEN_XHI) && !occ[d.c(W - 1, j, k)])
                        : (!occ[d.c(i - 1, j, k)] && !occ[d.c(i, j, k)]);
                if (ok) {
                    size_t fi = d.fu(i, j, k);
                    u[fi] = (float)((P)u[fi] - (p[d.pp(i, j, k)] - p[d.pp(i - 1, j, k)]));
                }
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j <= H; j++)
            for (int i = 0; i < W; i++) {
                bool ok = (j == 0) ? ((open_mask & ORC_OPEN_YLO) && !occ[d.c(i, 0, k)])
                        : (j == H) ? ((open_mask & ORC_OPEN_YHI) && !occ[d.c(i, H - 1, k)])
                        : (!occ[d.c(i, j - 1, k)] && !occ[d.c(i, j, k)]);
                if (ok) {
                    size_t fi = d.fv(i, j, k);
                    v[fi] = (float)((P)v[fi] - (p[d.pp(i, j, k)] - p[d.pp(i, j - 1, k)]));
                }
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                bool ok = (k == 0) ? ((open_mask & ORC_OPEN_ZLO) && !occ[d.c(i, j, 0)])
                        : (k == D) ? ((open_mask & ORC_OPEN_ZHI) && !occ[d.c(i, j, D - 1)])
                        : (!occ[d.c(i, j, k - 1)] && !occ[d.c(i, j, k)]);
                if (ok) {
                    size_t fi = d.fw(i, j, k);
                    w[fi] = (float)((P)w[fi] - (p[d.pp(i, j, k)] - p[d.pp(i, j, k - 1)]));
                }
            }
}

/* FP64 divergence of FP32 faces: ((u1-u0) + (v1-v0)) + (w1-w0) in double */
inline double div64_at(const Dims& d, const float* u, const float* v, const float* w, int i, int j, int k) {
    double du = (double)u[d.fu(i + 1, j, k)] - (double)u[d.fu(i, j, k)];
    double dv = (double)v[d.fv(i, j + 1, k)] - (double)v[d.fv(i, j, k)];
    double dw = (double)w[d.fw(i, j, k + 1)] - (double)w[d.fw(i, j, k)];
    return (du + dv) + dw;
}

/* mean of div64 over fluid cells; per-plane partials summed in plane order */
double mean_div64(const Dims& d, const float* u, const float* v, const float* w, const uint8_t* occ) {
    std::vector<double> ps(d.D, 0.0);
    std::vector<long long> pc(d.D, 0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < d.D; k++) {
        double s = 0.0; long long c = 0;
        for (int j = 0; j < d.H; j++)
            for (int i = 0; i < d.W; i++)
                if (!occ[d.c(i, j, k)]) { s += div64_at(d, u, v, w, i, j, k); c++; }
        ps[k] = s; pc[k] = c;
    }
    double s = 0.0; long long c = 0;
    for (int k = 0; k < d.D; k++) { s += ps[k]; c += pc[k]; }
    return c ? s / (double)c : 0.0;
}

/* r = b - (sum_nb p - n p), FP64; returns max|r| over counted cells */
double residual64(const Dims& d, const uint8_t* nc, const double* b, const double* p, float* r32) {
    const size_t sx = 1, sy = (size_t)(d.W + 2), sz = (size_t)(d.W + 2) * (d.H + 2);
    std::vector<double> pm(d.D, 0.0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < d.D; k++) {
        double m = 0.0;
        for (int j = 0; j < d.H; j++)
            for (int i = 0; i < d.W; i++) {
                size_t o = d.c(i, j, k);
                int n = nc[o];
                if (!n) { r32[o] = 0.0f; continue; }
                size_t q = d.pp(i, j, k);
                double s = p[q - sx] + p[q + sx] + p[q - sy] + p[q + sy] + p[q - sz] + p[q + sz];
                double r = b[o] - (s - (double)n * p[q]);
                r32[o] = (float)r;
                m = std::max(m, std::fabs(r));
            }
        pm[k] = m;
    }
    double m = 0.0;
    for (int k = 0; k < d.D; k++) m = std::max(m, pm[k]);
    return m;
}

}  // namespace

extern "C" {

void orc_project(int W, int H, int D, float* u, float* v, float* w, const uint8_t* occ, int open_mask,
                 float mean, float omega, int iters, float* p_pad_out) {
    Dims d{W, H, D};
    const size_t N = d.N();
    std::vector<float> div(N), b(N), p(d.Npad(), 0.0f);
    std::vector<uint8_t> nc(N);
    orc_divergence(W, H, D, u, v, w, occ, div.data());
    orc_sor_count(W, H, D, occ, open_mask, nc.data());
    for (size_t o = 0; o < N; o++) b[o] = nc[o] ? div[o] - mean : 0.0f;
    orc_sor(W, H, D, nc.data(), b.data(), p.data(), omega, iters);
    correct_faces<float>(d, occ, open_mask, p.data(), u, v, w);
    if (p_pad_out) std::memcpy(p_pad_out, p.data(), sizeof(float) * d.Npad());
}

/* The pass-length rule of the mixed-precision projection with fitted passes.  Over a pass of nprev
 * sweeps max|r| fell from rprev to r, a per-sweep factor rho = (r/rprev)^(1/nprev); reaching tol at
 * that rate takes log(tol/r)/log(rho) more sweeps.  The count is taken 1.25 times because the decay
 * slows as the smooth error modes take over (a stationary iteration's early sweeps remove the rough
 * modes fastest).  Evaluated as 1.25 * nprev * (log(tol/r) / log(r/rprev)). */
int orc_fitted_pass_length(int iters, int fit, double r, double rprev, int nprev, double tol) {
    if (!fit || !(rprev > 0.0) || !(r < rprev) || !(r > tol) || nprev < 1) return iters;
    const double need = 1.25 * (double)nprev * (std::log(tol / r) / std::log(r / rprev));
    if (!(need < (double)iters)) return iters;
    const int n = (int)std::ceil(need);
    return n < 1 ? 1 : n;
}

int orc_project_mp(int W, int H, int D, float* u, float* v, float* w, const uint8_t* occ, int open_mask,
                   int subtract_mean, float omega, int inner_iters, int max_outer, double tol,
                   double* p64, double* stats, int pass_rule) {
    Dims d{W, H, D};
    const size_t N = d.N();
    std::vector<uint8_t> nc(N);
    orc_sor_count(W, H, D, occ, open_mask, nc.data());
    double mean = subtract_mean ? mean_div64(d, u, v, w, occ) : 0.0;
    std::vector<double> b(N);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                size_t o = d.c(i, j, k);
                b[o] = nc[o] ? div64_at(d, u, v, w, i, j, k) - mean : 0.0;
                if (!nc[o]) p64[d.pp(i, j, k)] = 0.0;   /* keeps the zero-at-solid invariant */
            }
    std::vector<float> r32(N), e(d.Npad());
    long long sweeps = 0;
    int outer = 0;
    double rmax = 0.0;
    /* The residual is the divergence of the faces corrected in FP64; the stored faces are rounded to
     * FP32 once more.  So a residual within tol ends the passes only when the faces the correction
     * will store also have max|div - mean| <= tol (evaluated on a corrected copy). */
    std::vector<float> cu(u, u + d.Nu()), cv(v, v + d.Nv()), cw(w, w + d.Nw());
    double dprev = -1.0, rprev = -1.0;
    int nprev = 0;
    for (;;) {
        rmax = residual64(d, nc.data(), b.data(), p64, r32.data());
        if (!(rmax > tol)) {
            std::copy(u, u + d.Nu(), cu.begin());
            std::copy(v, v + d.Nv(), cv.begin());
            std::copy(w, w + d.Nw(), cw.begin());
            correct_faces<double>(d, occ, open_mask, p64, cu.data(), cv.data(), cw.data());
            const double dmax = orc_max_abs_div(W, H, D, cu.data(), cv.data(), cw.data(), occ, open_mask, mean);
            if (!(dmax > tol)) break;
            /* a pass that no longer shrinks it below 0.9 of the previous one meets the faces' FP32
             * rounding floor: the bound is out of reach, stop */
            if (dprev >= 0.0 && !(dmax < 0.9 * dprev)) break;
            dprev = dmax;
        }
        if (outer >= max_outer) break;
        std::fill(e.begin(), e.end(), 0.0f);
        int n;
        if (pass_rule == 2) {
            n = orc_sor_gated(W, H, D, nc.data(), r32.data(), e.data(), omega, inner_iters,
                              orc_gate_threshold(tol, omega));
        } else {
            n = orc_fitted_pass_length(inner_iters, pass_rule == 1, rmax, rprev, nprev, tol);
            orc_sor(W, H, D, nc.data(), r32.data(), e.data(), omega, n);
        }
        sweeps += n;
        for (int k = 0; k < D; k++)
            for (int j = 0; j < H; j++)
                for (int i = 0; i < W; i++) {
                    size_t q = d.pp(i, j, k);
                    p64[q] += (double)e[q];
                }
        rprev = rmax;
        nprev = n;
        outer++;
    }
    correct_faces<double>(d, occ, open_mask, p64, u, v, w);
    if (stats) {
        stats[0] = (double)sweeps;
        stats[1] = rmax;
        stats[2] = orc_max_abs_div(W, H, D, u, v, w, occ, open_mask, mean);
        stats[3] = mean;
    }
    return outer;
}

void orc_extrapolate_warm_start(size_t n, double* p, double* prev, int valid, double beta) {
    for (size_t i = 0; i < n; i++) {
        const double t = p[i];
        if (valid) p[i] = t + beta * (t - prev[i]);
        prev[i] = t;
    }
}

int orc_project_cg64(int W, int H, int D, float* u, float* v, float* w, const uint8_t* occ, int open_mask,
                     int subtract_mean, double tol, int max_iter, double* p64) {
    Dims d{W, H, D};
    const size_t N = d.N(), NP = d.Npad();
    std::vector<uint8_t> nc(N);
    orc_sor_count(W, H, D, occ, open_mask, nc.data());
    double mean = subtract_mean ? mean_div64(d, u, v, w, occ) : 0.0;
    const size_t sx = 1, sy = (size_t)(W + 2), sz = (size_t)(W + 2) * (H + 2);
    /* A = -L (SPD on counted cells): (A x)_c = n x_c - sum_nb x_nb ; solve A p = -b */
    auto applyA = [&](const std::vector<double>& x, std::vector<double>& y) {
#pragma omp parallel for schedule(static)
        for (int k = 0; k < D; k++)
            for (int j = 0; j < H; j++)
                for (int i = 0; i < W; i++) {
                    size_t o = d.c(i, j, k), q = d.pp(i, j, k);
                    if (!nc[o]) { y[q] = 0.0; continue; }
                    double s = x[q - sx] + x[q + sx] + x[q - sy] + x[q + sy] + x[q - sz] + x[q + sz];
                    y[q] = (double)nc[o] * x[q] - s;
                }
    };
    std::vector<double> x(NP, 0.0), r(NP, 0.0), pdir(NP, 0.0), Ap(NP, 0.0);
    double bnorm = 0.0;
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                size_t o = d.c(i, j, k), q = d.pp(i, j, k);
                double bb = nc[o] ? -(div64_at(d, u, v, w, i, j, k) - mean) : 0.0;
                r[q] = bb; bnorm = std::max(bnorm, std::fabs(bb));
            }
    pdir = r;
    const long long NPl = (long long)NP;
    double rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr)
    for (long long q = 0; q < NPl; q++) rr += r[q] * r[q];
    int it = 0;
    for (; it < max_iter; it++) {
        double rmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : rmax)
        for (long long q = 0; q < NPl; q++) rmax = std::max(rmax, std::fabs(r[q]));
        if (rmax <= tol) break;
        applyA(pdir, Ap);
        double pAp = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : pAp)
        for (long long q = 0; q < NPl; q++) pAp += pdir[q] * Ap[q];
        if (pAp <= 0.0) break;
        double alpha = rr / pAp;
        double rr2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr2)
        for (long long q = 0; q < NPl; q++) {
            x[q] += alpha * pdir[q];
            r[q] -= alpha * Ap[q];
            rr2 += r[q] * r[q];
        }
        double beta = rr2 / rr;
        rr = rr2;
#pragma omp parallel for schedule(static)
        for (long long q = 0; q < NPl; q++) pdir[q] = r[q] + beta * pdir[q];
    }
    correct_faces<double>(d, occ, open_mask, x.data(), u, v, w);
    if (p64) std::memcpy(p64, x.data(), sizeof(double) * NP);
    return it;
}

double orc_max_abs_div(int W, int H, int D, const float* u, const float* v, const float* w,
                       const uint8_t* occ, int open_mask, double mean) {
    Dims d{W, H, D};
    std::vector<uint8_t> nc(d.N());
    orc_sor_count(W, H, D, occ, open_mask, nc.data());
    double m = 0.0;
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++)
                if (nc[d.c(i, j, k)]) m = std::max(m, std::fabs(div64_at(d, u, v, w, i, j, k) - mean));
    return m;
}

float orc_closed_mean_divergence(int W, int H, int D, const float* u, const float* v, const float* w,
                                 const uint8_t* occ, int* count_out) {
    Dims d{W, H, D};
    float s = 0.0f;
    int count = 0;
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++)
                if (!occ[d.c(i, j, k)]) {
                    count++;
                    s += u[d.fu(i + 1, j, k)] - u[d.fu(i, j, k)] + v[d.fv(i, j + 1, k)] - v[d.fv(i, j, k)]
                       + w[d.fw(i, j, k + 1)] - w[d.fw(i, j, k)];
                }
    if (count_out) *count_out = count;
    return count ? s / count : 0.0f;
}

/* Mean divergence of the device-resident path: the FP32 cell divergences summed in FP64 over
 * fluid cells, divided by the count, rounded to float.  (The reference's host loop accumulates
 * in FP32, ZQ_SmokeSimulationUtils3D.cpp:2387-2421; see orc_closed_mean_divergence.) */
float orc_mean_div32_f64(int W, int H, int D, const float* u, const float* v, const float* w, const uint8_t* occ) {
    Dims d{W, H, D};
    std::vector<float> div(d.N());
    orc_divergence(W, H, D, u, v, w, occ, div.data());
    double s = 0.0, c = 0.0;
    for (size_t o = 0; o < d.N(); o++)
        if (!occ[o]) { s += (double)div[o]; c += 1.0; }
    return c > 0.0 ? (float)(s / c) : 0.0f;
}

/* ZQ_InitFineGrid3D (ZQ_SmokeSimulationUtils3D.cpp:1523-1664).  shell_mask selects which domain
 * faces get a one-cell solid layer (the reference's closed solvers set all six). */
void orc_init_fine_grid(int W, int H, int D, int shell_mask, int intobj, uint8_t* occ, float* u, float* v, float* w) {
    Dims d{W, H, D};
    std::memset(occ, 0, d.N());
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                bool s = ((shell_mask & ORC_OPEN_XLO) && i == 0) || ((shell_mask & ORC_OPEN_XHI) && i == W - 1)
                      || ((shell_mask & ORC_OPEN_YLO) && j == 0) || ((shell_mask & ORC_OPEN_YHI) && j == H - 1)
                      || ((shell_mask & ORC_OPEN_ZLO) && k == 0) || ((shell_mask & ORC_OPEN_ZHI) && k == D - 1);
                if (s) occ[d.c(i, j, k)] = 1;
            }
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                bool s = false;
                switch (intobj) {
                case 1: /* sphere (:1588-1601), integer halves as in the reference */
                    s = (k + 0.5 - D / 2) * (k + 0.5 - D / 2) + (j + 0.5 - H / 2) * (j + 0.5 - H / 2)
                        + (i + 0.5 - W / 2) * (i + 0.5 - W / 2) < W * W / 400.0;
                    break;
                case 2: /* box (:1603-1617) */
                    s = std::fabs(k + 0.5 - D / 2) < D / 20 && std::fabs(j + 0.5 - H / 2) < H / 20
                        && std::fabs(i + 0.5 - W / 2) < W / 10;
                    break;
                case 3: { /* oval (:1618-1636): the reference's loop variables swap i/j roles
                             (i runs over height, j over width) and index occupy[k][j][i] */
                    break;
                }
                case 4: /* special1 (:1637-1650) */
                    s = (j == H / 2 + 13 || j == H / 2 + 14) && (i % 4 == 0 || i % 4 == 3) && i >= 4 && i < W - 4;
                    break;
                default: break;
                }
                if (s) occ[d.c(i, j, k)] = 1;
            }
    if (intobj == 3) {
        /* i in [0,H), j in [0,W): distance tests use (i, j) as (x, y) and write occupy[k*H*W + j*W + i] */
        for (int k = 0; k < D; k++)
            for (int i = 0; i < H; i++)
                for (int j = 0; j < W; j++) {
                    double a = std::sqrt((i + 0.5 - 3.0 * W / 7) * (i + 0.5 - 3.0 * W / 7) + (j + 0.5 - H / 2) * (j + 0.5 - H / 2)
                                         + (k + 0.5 - D / 2) * (k + 0.5 - D / 2));
                    double c = std::sqrt((i + 0.5 - 4.0 * W / 7) * (i + 0.5 - 4.0 * W / 7) + (j + 0.5 - H / 2) * (j + 0.5 - H / 2)
                                         + (k + 0.5 - D / 2) * (k + 0.5 - D / 2));
                    if (a + c < W / 5.0) {
                        size_t idx = (size_t)k * H * W + (size_t)j * W + i;
                        if (idx < d.N()) occ[idx] = 1;
                    }
                }
    }
    std::memset(u, 0, sizeof(float) * d.Nu());
    std::memset(v, 0, sizeof(float) * d.Nv());
    std::memset(w, 0, sizeof(float) * d.Nw());
}

/* ZQ_ReinjectTemperatureDensity3D (ZQ_SmokeSimulationUtils3D.cpp:3842-3882), randDensity off */
void orc_reinject_box(int W, int H, int D, const float* box, float* T, float* rho) {
    Dims d{W, H, D};
    const float cx = box[0] * W, cy = box[1] * H, cz = box[2] * D;
    const float hx = box[3] * W, hy = box[4] * H, hz = box[5] * D;
    const float rd = box[6], rt = box[7];
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++)
                if (std::fabs(i + 0.5f - cx) <= hx && std::fabs(j + 0.5f - cy) <= hy && std::fabs(k + 0.5f - cz) <= hz) {
                    rho[d.c(i, j, k)] = rd * 1.0f;
                    T[d.c(i, j, k)] = rt;
                }
}

void orc_reinject_sphere(int W, int H, int D, float cx, float cy, float cz, float r, float rho_src, float T_src,
                         float* T, float* rho) {
    Dims d{W, H, D};
    const float x0 = cx * W, y0 = cy * H, z0 = cz * D, rr = r * W;
    const float r2 = rr * rr;
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                float dx = i + 0.5f - x0, dy = j + 0.5f - y0, dz = k + 0.5f - z0;
                if (dx * dx + dy * dy + dz * dz <= r2) { rho[d.c(i, j, k)] = rho_src; T[d.c(i, j, k)] = T_src; }
            }
}

/* Extension in the semantics of ZQ_ApplyMovingObject3D (ZQ_SmokeSimulationUtils3D.cpp:1933-2046):
 * occupancy cleared, solid voxels stamped, the six faces of each solid voxel set to the solid's
 * velocity, the shell re-added.  Here the solid is an analytic sphere (normalized centre, radius
 * in units of W) and the velocity is a float (the reference truncates an integer). */
void orc_stamp_sphere(int W, int H, int D, float cx, float cy, float cz, float r, float su, float sv, float sw,
                      int shell_mask, uint8_t* occ, float* u, float* v, float* w) {
    Dims d{W, H, D};
    std::memset(occ, 0, d.N());
    const float x0 = cx * W, y0 = cy * H, z0 = cz * D, rr = r * W;
    const float r2 = rr * rr;
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                float dx = i + 0.5f - x0, dy = j + 0.5f - y0, dz = k + 0.5f - z0;
                if (dx * dx + dy * dy + dz * dz <= r2) {
                    occ[d.c(i, j, k)] = 1;
                    u[d.fu(i, j, k)] = su; u[d.fu(i + 1, j, k)] = su;
                    v[d.fv(i, j, k)] = sv; v[d.fv(i, j + 1, k)] = sv;
                    w[d.fw(i, j, k)] = sw; w[d.fw(i, j, k + 1)] = sw;
                }
            }
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                bool s = ((shell_mask & ORC_OPEN_XLO) && i == 0) || ((shell_mask & ORC_OPEN_XHI) && i == W - 1)
                      || ((shell_mask & ORC_OPEN_YLO) && j == 0) || ((shell_mask & ORC_OPEN_YHI) && j == H - 1)
                      || ((shell_mask & ORC_OPEN_ZLO) && k == 0) || ((shell_mask & ORC_OPEN_ZHI) && k == D - 1);
                if (s) occ[d.c(i, j, k)] = 1;
            }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads of the following calls (n <= 0: the OpenMP default) */
void orc_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
#else
    (void)n;
#endif
}

void orc_noise(int W, int H, int D, uint32_t seed, float amp, float* out) {
    const size_t N = (size_t)W * H * D;
    for (size_t o = 0; o < N; o++) {
        uint32_t x = hash_u32((uint32_t)o * 0x9E3779B9u ^ hash_u32(seed + 0x632BE5ABu));
        float u01 = (float)(x >> 8) * (1.0f / 16777216.0f);
        out[o] = amp * (2.0f * u01 - 1.0f);
    }
}

/* UpdateOneFrame order (ZQ_SmokeSimulation3D.cpp:156-243): force, attenuation, velocity
 * advection, projection, scalar advection (with the projected velocity), reinjection. */
int orc_step(const orc_step_params* P, float* u, float* v, float* w, float* T, float* rho, const uint8_t* occ,
             const float* extra_force, double* p64_pad, double* stats) {
    const int W = P->W, H = P->H, D = P->D;
    Dims d{W, H, D};
    orc_add_force(W, H, D, u, v, w, occ, T, rho, P->h, P->dt, P->beta, P->alpha, P->eps, P->Tamb, extra_force,
                  nullptr, nullptr, nullptr);
    orc_attenuation(W, H, D, u, v, w, T, rho, occ, P->aV, P->aT, P->aR);
    const float dtt = P->dt / P->steps;
    if (orc_advect_velocity(W, H, D, P->h, P->steps, dtt, u, v, w, occ, 1, P->adv_vel_type)) return -1;
    if (P->mixed) {
        orc_project_mp(W, H, D, u, v, w, occ, P->open_mask, P->closed_mean, P->omega, P->sor_iters, P->max_outer,
                       P->tol, p64_pad, stats, P->pass_rule);
    } else {
        float mean = 0.0f;
        if (P->closed_mean) mean = orc_mean_div32_f64(W, H, D, u, v, w, occ);
        orc_project(W, H, D, u, v, w, occ, P->open_mask, mean, P->omega, P->sor_iters, nullptr);
    }
    std::vector<float> T2(d.N()), R2(d.N());
    if (orc_advect_scalar(W, H, D, P->h, P->steps, dtt, u, v, w, occ, T, rho, T2.data(), R2.data(), P->adv_sca_type))
        return -1;
    std::memcpy(T, T2.data(), sizeof(float) * d.N());
    std::memcpy(rho, R2.data(), sizeof(float) * d.N());
    for (int s = 0; s < P->n_box; s++) orc_reinject_box(W, H, D, P->boxes + 8 * s, T, rho);
    for (int s = 0; s < P->n_sphere; s++) {
        const float* sp = P->spheres + 6 * s;
        orc_reinject_sphere(W, H, D, sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], T, rho);
    }
    return 0;
}

}  // extern "C"
