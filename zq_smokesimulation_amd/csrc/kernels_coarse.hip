// kernels_coarse.hip -- the coarse + sub-grid projection (COARSE_OPEN_POISSON_SOR,
// ZQ_SmokeSimulationUtils3D.cpp:3688-3795) on gfx950.
//
// Set-up, averaging, the coarse red-black sweeps and the coarse -> fine correction are integer /
// gather work with the reference's operation order, so they are bit-exact with the CPU
// restatement.  The sub-projection of a regular block (no solid cell) is the reference's dense
// operator M = I - D [0; pinv(P) V] restricted to the interior faces (3 s^2 (s-1) x 3 s^2 (s+1),
// 1344 x 1728 at s = 8): here it is applied matrix-free.  M x subtracts the gradient of the
// block's Neumann pressure p with L p = div x - mean(div x); the block Laplacian L is diagonalised
// by the orthonormal DCT-II basis along each axis, so p = Q3 diag(1/lambda) Q3^T div with the
// constant mode dropped (which is the mean removal), 12 s^4 flops per block instead of
// 2 * 1344 * 1728.  Blocks with solid cells solve the block's closed Poisson problem
// (ZQ_SolveSubGrid3D, :1393-1521) by FP64 conjugate gradients inside one workgroup.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "zq3d_coarse.h"
#include "zq3d_kernels.h"

namespace zq3d {
namespace {

struct CG {
    int W, H, D, s, Wc, Hc, Dc;
};

CG make_cg(const CoarseDims& g) { return CG{g.f.W, g.f.H, g.f.D, g.s, g.Wc, g.Hc, g.Dc}; }

__device__ __forceinline__ int f_c(const CG& g, int i, int j, int k) { return (k * g.H + j) * g.W + i; }
__device__ __forceinline__ int f_u(const CG& g, int i, int j, int k) { return (k * g.H + j) * (g.W + 1) + i; }
__device__ __forceinline__ int f_v(const CG& g, int i, int j, int k) { return (k * (g.H + 1) + j) * g.W + i; }
__device__ __forceinline__ int f_w(const CG& g, int i, int j, int k) { return (k * g.H + j) * g.W + i; }
__device__ __forceinline__ int c_c(const CG& g, int i, int j, int k) { return (k * g.Hc + j) * g.Wc + i; }
__device__ __forceinline__ int c_u(const CG& g, int i, int j, int k) { return (k * g.Hc + j) * (g.Wc + 1) + i; }
__device__ __forceinline__ int c_v(const CG& g, int i, int j, int k) { return (k * (g.Hc + 1) + j) * g.Wc + i; }
__device__ __forceinline__ int c_w(const CG& g, int i, int j, int k) { return (k * g.Hc + j) * g.Wc + i; }
__device__ __forceinline__ int c_p(const CG& g, int i, int j, int k) {
    return ((k + 1) * (g.Hc + 2) + (j + 1)) * (g.Wc + 2) + (i + 1);
}

inline unsigned blocks_for(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// dst[e] = src[e] for e < n, by the workgroup's nt threads: U loads in flight per thread before
// the LDS stores (a load-store loop would wait out the memory latency once per element)
template <int U>
__device__ __forceinline__ void stage(float* dst, const float* __restrict__ src, int n, int tid, int nt) {
    for (int base = tid; base < n; base += U * nt) {
        float t[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int e = base + q * nt;
            t[q] = e < n ? src[e] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int e = base + q * nt;
            if (e < n) dst[e] = t[q];
        }
    }
}

// ------------------------------------------------------------------------------ set-up
// One thread per coarse face: 1 - (solid sub-faces)/s^2; a sub-face of an interior coarse face is
// solid when either fine cell beside it is, of a domain face when the inside cell is
// (ZQ_InitCoarseGrid3D, ZQ_SmokeSimulationUtils3D.cpp:1690-1835).
__global__ __launch_bounds__(256) void k_coarse_fractions(CG g, const uint8_t* __restrict__ occ, float* rU, float* rV,
                                                          float* rW) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nu = (size_t)(g.Wc + 1) * g.Hc * g.Dc, nv = (size_t)g.Wc * (g.Hc + 1) * g.Dc;
    const size_t nw = (size_t)g.Wc * g.Hc * (g.Dc + 1);
    const int s = g.s;
    const float sf = (float)(s * s);
    int n = 0;
    if (t < nu) {
        const int i = (int)(t % (g.Wc + 1)), j = (int)((t / (g.Wc + 1)) % g.Hc), k = (int)(t / ((size_t)(g.Wc + 1) * g.Hc));
        for (int sk = 0; sk < s; sk++)
            for (int sj = 0; sj < s; sj++) {
                const int y = j * s + sj, z = k * s + sk;
                bool solid;
                if (i == 0) solid = occ[f_c(g, 0, y, z)];
                else if (i == g.Wc) solid = occ[f_c(g, g.W - 1, y, z)];
                else solid = occ[f_c(g, i * s, y, z)] || occ[f_c(g, i * s - 1, y, z)];
                n += solid;
            }
        rU[t] = 1.0f - (float)n / sf;
        return;
    }
    t -= nu;
    if (t < nv) {
        const int i = (int)(t % g.Wc), j = (int)((t / g.Wc) % (g.Hc + 1)), k = (int)(t / ((size_t)g.Wc * (g.Hc + 1)));
        for (int sk = 0; sk < s; sk++)
            for (int si = 0; si < s; si++) {
                const int x = i * s + si, z = k * s + sk;
                bool solid;
                if (j == 0) solid = occ[f_c(g, x, 0, z)];
                else if (j == g.Hc) solid = occ[f_c(g, x, g.H - 1, z)];
                else solid = occ[f_c(g, x, j * s, z)] || occ[f_c(g, x, j * s - 1, z)];
                n += solid;
            }
        rV[t] = 1.0f - (float)n / sf;
        return;
    }
    t -= nv;
    if (t < nw) {
        const int i = (int)(t % g.Wc), j = (int)((t / g.Wc) % g.Hc), k = (int)(t / ((size_t)g.Wc * g.Hc));
        for (int sj = 0; sj < s; sj++)
            for (int si = 0; si < s; si++) {
                const int x = i * s + si, y = j * s + sj;
                bool solid;
                if (k == 0) solid = occ[f_c(g, x, y, 0)];
                else if (k == g.Dc) solid = occ[f_c(g, x, y, g.D - 1)];
                else solid = occ[f_c(g, x, y, k * s)] || occ[f_c(g, x, y, k * s - 1)];
                n += solid;
            }
        rW[t] = 1.0f - (float)n / sf;
    }
}

// One thread per coarse cell: solid when all six fractions are 0 (:1837-1858); 1/c of the
// face-fraction operator with c = (open out-of-domain sides) + r(x-) + r(x+) + r(y-) + r(y+) +
// r(z-) + r(z+) over the in-domain sides (the system of ZQ_InitMatOpenPoisson3D, :570-589), or with
// flux != 0 the sum of all six fractions, out-of-domain sides only when open (the flux KKT systems,
// :749-1150); 0 for solid cells and c == 0; regular = no solid fine cell in the block (:3318-3340).  derive == 0 takes
// the coarse occupancy as given in cocc instead (the C ABI solve's caller passes it).
__global__ __launch_bounds__(256) void k_coarse_cells(CG g, const uint8_t* __restrict__ occ, int open_mask,
                                                      const float* __restrict__ rU, const float* __restrict__ rV,
                                                      const float* __restrict__ rW, uint8_t* cocc, float* inv,
                                                      uint8_t* regular, float* vol, int derive, int flux) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)g.Wc * g.Hc * g.Dc) return;
    const int i = (int)(t % g.Wc), j = (int)((t / g.Wc) % g.Hc), k = (int)(t / ((size_t)g.Wc * g.Hc));
    const float r0 = rU[c_u(g, i, j, k)], r1 = rU[c_u(g, i + 1, j, k)];
    const float r2 = rV[c_v(g, i, j, k)], r3 = rV[c_v(g, i, j + 1, k)];
    const float r4 = rW[c_w(g, i, j, k)], r5 = rW[c_w(g, i, j, k + 1)];
    const bool solid = derive ? (r0 == 0 && r1 == 0 && r2 == 0 && r3 == 0 && r4 == 0 && r5 == 0) : cocc[t] != 0;
    float iv = 0.0f;
    if (!solid && !flux) {
        int nopen = 0;
        nopen += (i == 0) && (open_mask & 1);
        nopen += (i == g.Wc - 1) && (open_mask & 2);
        nopen += (j == 0) && (open_mask & 4);
        nopen += (j == g.Hc - 1) && (open_mask & 8);
        nopen += (k == 0) && (open_mask & 16);
        nopen += (k == g.Dc - 1) && (open_mask & 32);
        float c = (float)nopen;
        if (i > 0) c += r0;
        if (i < g.Wc - 1) c += r1;
        if (j > 0) c += r2;
        if (j < g.Hc - 1) c += r3;
        if (k > 0) c += r4;
        if (k < g.Dc - 1) c += r5;
        iv = c > 0.0f ? 1.0f / c : 0.0f;
    } else if (!solid) {
        // flux form: every side by its own fraction, an out-of-domain side only when open
        float c = 0.0f;
        if (i > 0 || (open_mask & 1)) c += r0;
        if (i < g.Wc - 1 || (open_mask & 2)) c += r1;
        if (j > 0 || (open_mask & 4)) c += r2;
        if (j < g.Hc - 1 || (open_mask & 8)) c += r3;
        if (k > 0 || (open_mask & 16)) c += r4;
        if (k < g.Dc - 1 || (open_mask & 32)) c += r5;
        iv = c > 0.0f ? 1.0f / c : 0.0f;
    }
    if (derive) cocc[t] = solid;
    inv[t] = iv;
    if (regular) {
        const int s = g.s;
        int n = 0;
        for (int sk = 0; sk < s; sk++)
            for (int sj = 0; sj < s; sj++)
                for (int si = 0; si < s; si++) n += occ[f_c(g, i * s + si, j * s + sj, k * s + sk)] != 0;
        regular[t] = n == 0;
        // BuildVolumeUnOccupyRatio (:1860-1885)
        vol[t] = 1.0f - (float)n / (float)(s * s * s);
    }
}

// Closed domains: mean = sum(div) / sum(vol) over the non-solid coarse cells, FP64, one workgroup
// of 1024 threads (thread t sums cells t, t+1024, ... in order, then a fixed tree), then
// div = (float)(div - mean * vol) on those cells.
__global__ __launch_bounds__(1024) void k_coarse_mean(int n, const float* __restrict__ div, const float* __restrict__ vol,
                                                      const uint8_t* __restrict__ cocc, double* mean) {
    __shared__ double sd[1024], sv[1024];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int c = t; c < n; c += 1024)
        if (!cocc[c]) {
            a += (double)div[c];
            b += (double)vol[c];
        }
    sd[t] = a;
    sv[t] = b;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h) {
            sd[t] += sd[t + h];
            sv[t] += sv[t + h];
        }
        __syncthreads();
    }
    if (t == 0) *mean = sv[0] > 0.0 ? sd[0] / sv[0] : 0.0;
}

__global__ __launch_bounds__(256) void k_coarse_sub_mean(int n, float* __restrict__ div, const float* __restrict__ vol,
                                                         const uint8_t* __restrict__ cocc, const double* mean) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || cocc[c]) return;
    div[c] = (float)((double)div[c] - *mean * (double)vol[c]);
}

// ------------------------------------------------------------------------------ averaging
// coarse face = (sum of its s x s fine sub-faces, slower axis outer) / s^2 (ZQ_MapFine2Coarse3D,
// :2190-2254)
__global__ __launch_bounds__(256) void k_coarse_average(CG g, const float* __restrict__ u, const float* __restrict__ v,
                                                        const float* __restrict__ w, float* cu, float* cv, float* cw) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nu = (size_t)(g.Wc + 1) * g.Hc * g.Dc, nv = (size_t)g.Wc * (g.Hc + 1) * g.Dc;
    const size_t nw = (size_t)g.Wc * g.Hc * (g.Dc + 1);
    const int s = g.s;
    const int sf = s * s;
    float a = 0.0f;
    if (t < nu) {
        const int i = (int)(t % (g.Wc + 1)), j = (int)((t / (g.Wc + 1)) % g.Hc), k = (int)(t / ((size_t)(g.Wc + 1) * g.Hc));
        for (int sk = 0; sk < s; sk++)
            for (int sj = 0; sj < s; sj++) a += u[f_u(g, i * s, j * s + sj, k * s + sk)];
        cu[t] = a / sf;
        return;
    }
    t -= nu;
    if (t < nv) {
        const int i = (int)(t % g.Wc), j = (int)((t / g.Wc) % (g.Hc + 1)), k = (int)(t / ((size_t)g.Wc * (g.Hc + 1)));
        for (int sk = 0; sk < s; sk++)
            for (int si = 0; si < s; si++) a += v[f_v(g, i * s + si, j * s, k * s + sk)];
        cv[t] = a / sf;
        return;
    }
    t -= nv;
    if (t < nw) {
        const int i = (int)(t % g.Wc), j = (int)((t / g.Wc) % g.Hc), k = (int)(t / ((size_t)g.Wc * g.Hc));
        for (int sj = 0; sj < s; sj++)
            for (int si = 0; si < s; si++) a += w[f_w(g, i * s + si, j * s + sj, k * s)];
        cw[t] = a / sf;
    }
}

// coarse divergence u[i+1]-u[i]+v[j+1]-v[j]+w[k+1]-w[k], left to right; 0 on solid coarse cells
__global__ __launch_bounds__(256) void k_coarse_div(CG g, const float* __restrict__ cu, const float* __restrict__ cv,
                                                    const float* __restrict__ cw, const uint8_t* __restrict__ cocc,
                                                    float* div) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)g.Wc * g.Hc * g.Dc) return;
    const int i = (int)(t % g.Wc), j = (int)((t / g.Wc) % g.Hc), k = (int)(t / ((size_t)g.Wc * g.Hc));
    div[t] = cocc[t] ? 0.0f
                     : cu[c_u(g, i + 1, j, k)] - cu[c_u(g, i, j, k)] + cv[c_v(g, i, j + 1, k)] - cv[c_v(g, i, j, k)] +
                           cw[c_w(g, i, j, k + 1)] - cw[c_w(g, i, j, k)];
}

// ------------------------------------------------------------------------------ coarse sweeps
// One colour of a red-black iteration of the face-fraction operator: cells (i+j+k)&1 == colour,
// s = r(x-) p(x-) + r(x+) p(x+) + r(y-) p(y-) + r(y+) p(y+) + r(z-) p(z-) + r(z+) p(z+) left to right,
// p = (1-omega) p + omega (s - div) * inv.  Ghosts and solid cells hold 0.
__global__ __launch_bounds__(256) void k_coarse_sweep(CG g, const float* __restrict__ rU, const float* __restrict__ rV,
                                                      const float* __restrict__ rW, const float* __restrict__ inv,
                                                      const float* __restrict__ div, float* p, int colour, float omega,
                                                      float omw) {
    const int Wh = (g.Wc + 1) >> 1;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)Wh * g.Hc * g.Dc) return;
    const int xh = (int)(t % Wh), j = (int)((t / Wh) % g.Hc), k = (int)(t / ((size_t)Wh * g.Hc));
    const int i = 2 * xh + ((colour + j + k) & 1);
    if (i >= g.Wc) return;
    const int o = c_c(g, i, j, k);
    const int sy = g.Wc + 2, sz = (g.Wc + 2) * (g.Hc + 2);
    const int q = c_p(g, i, j, k);
    // every operand is loaded before the test, so the loads share one memory latency
    const float iv = inv[o];
    const float s = rU[c_u(g, i, j, k)] * p[q - 1] + rU[c_u(g, i + 1, j, k)] * p[q + 1] + rV[c_v(g, i, j, k)] * p[q - sy] +
                    rV[c_v(g, i, j + 1, k)] * p[q + sy] + rW[c_w(g, i, j, k)] * p[q - sz] +
                    rW[c_w(g, i, j, k + 1)] * p[q + sz];
    const float pg = (s - div[o]) * iv;
    const float pn = omw * p[q] + omega * pg;
    if (iv != 0.0f) p[q] = pn;
}

// k red-black iterations in one launch: a workgroup owns the z-planes [z0, z0 + zc) and keeps
// pressure planes [z0 - 2k, z0 + zc + 2k) (whole padded x-y planes, ghosts included) in LDS, with
// the fractions, divergence and 1/diagonal of the planes it updates.  Half-sweep h updates the
// planes within 2k-1-h of its own range, so every value it reads is one a whole-grid sweep would
// have produced (halo planes are recomputed redundantly, with the same operations), and the owned
// planes end bit-identical to 2k single-colour launches.  src is read, dst written.
struct SlabLayout {
    int zlo, zhi, clo, chi;  // pressure planes [zlo, zhi), updated cell planes [clo, chi)
    int NP, NC;
};
__host__ __device__ inline SlabLayout slab_layout(int Dc, int z0, int k, int zc) {
    SlabLayout L;
    L.zlo = z0 - 2 * k < -1 ? -1 : z0 - 2 * k;
    L.zhi = z0 + zc + 2 * k > Dc + 1 ? Dc + 1 : z0 + zc + 2 * k;
    L.clo = z0 - 2 * k + 1 < 0 ? 0 : z0 - 2 * k + 1;
    L.chi = z0 + zc + 2 * k - 1 > Dc ? Dc : z0 + zc + 2 * k - 1;
    L.NP = L.zhi - L.zlo;
    L.NC = L.chi - L.clo;
    return L;
}
// floats of LDS for planes NP (pressure) and NC (cells): padded pressure, u/v fractions of NC planes,
// w fractions of NC+1 face planes, divergence and 1/diagonal of NC planes
__host__ __device__ inline size_t slab_floats(int Wc, int Hc, int NP, int NC) {
    const size_t PS = (size_t)(Wc + 2) * (Hc + 2), CS = (size_t)Wc * Hc;
    return NP * PS + NC * ((size_t)(Wc + 1) * Hc + (size_t)Wc * (Hc + 1) + 2 * CS) + (NC + 1) * CS;
}

__global__ __launch_bounds__(1024) void k_coarse_slab(CG g, const float* __restrict__ rU, const float* __restrict__ rV,
                                                     const float* __restrict__ rW, const float* __restrict__ inv,
                                                     const float* __restrict__ div, const float* __restrict__ src,
                                                     float* __restrict__ dst, float omega, float omw, int k, int zc) {
    extern __shared__ float lds[];
    const int z0 = blockIdx.x * zc;
    const SlabLayout L = slab_layout(g.Dc, z0, k, zc);
    const int PW = g.Wc + 2, PS = PW * (g.Hc + 2);
    const int CS = g.Wc * g.Hc, US = (g.Wc + 1) * g.Hc, VS = g.Wc * (g.Hc + 1);
    float* P = lds;
    float* sU = P + L.NP * PS;
    float* sV = sU + L.NC * US;
    float* sW = sV + L.NC * VS;
    float* sD = sW + (L.NC + 1) * CS;
    float* sI = sD + L.NC * CS;
    const int tid = threadIdx.x, nt = blockDim.x;
    stage<16>(P, src + (size_t)(L.zlo + 1) * PS, L.NP * PS, tid, nt);
    stage<16>(sU, rU + (size_t)L.clo * US, L.NC * US, tid, nt);
    stage<16>(sV, rV + (size_t)L.clo * VS, L.NC * VS, tid, nt);
    stage<16>(sW, rW + (size_t)L.clo * CS, (L.NC + 1) * CS, tid, nt);
    stage<16>(sD, div + (size_t)L.clo * CS, L.NC * CS, tid, nt);
    stage<16>(sI, inv + (size_t)L.clo * CS, L.NC * CS, tid, nt);
    __syncthreads();
    const int Wh = (g.Wc + 1) >> 1;
    for (int hs = 0; hs < 2 * k; hs++) {
        const int c = hs & 1, r = 2 * k - 1 - hs;
        const int a = max(z0 - r, 0), bnd = min(z0 + zc + r, g.Dc);
        const int n = (bnd - a) * g.Hc * Wh;
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int xh = e % Wh, j = (e / Wh) % g.Hc, kk = a + e / (Wh * g.Hc);
            const int i = 2 * xh + ((c + j + kk) & 1);
            if (i >= g.Wc) continue;
            const int lk = kk - L.clo;
            const int o = lk * CS + j * g.Wc + i;
            const float iv = sI[o];
            if (iv == 0.0f) continue;
            const int q = (kk - L.zlo) * PS + (j + 1) * PW + (i + 1);
            const int fu = lk * US + j * (g.Wc + 1) + i, fv = lk * VS + j * g.Wc + i;
            const float s = sU[fu] * P[q - 1] + sU[fu + 1] * P[q + 1] + sV[fv] * P[q - PW] + sV[fv + g.Wc] * P[q + PW] +
                            sW[o] * P[q - PS] + sW[o + CS] * P[q + PS];
            const float pg = (s - sD[o]) * iv;
            P[q] = omw * P[q] + omega * pg;
        }
        __syncthreads();
    }
    const int zend = min(z0 + zc, g.Dc);
    for (int e = threadIdx.x; e < (zend - z0) * PS; e += blockDim.x)
        dst[(size_t)(z0 + 1) * PS + e] = P[(z0 - L.zlo) * PS + e];
}

// ------------------------------------------------------------------------------ coarse delta
// delta of a coarse face: (u - g) - u with g = p_r - p_l on faces with fraction > 0 (domain faces:
// open ones, p = 0 outside), else 0 (ZQ_SolveCoarseOpenPoisson3DSOR, :2611-2642, with the correction
// of ZQ_AdjustOpenVelocity3D, :2794-2865)
__device__ __forceinline__ float delta_of(float r, bool ok, float uf, float pr, float pl) {
    if (!(r > 0.0f) || !ok) return 0.0f;
    const float gr = pr - pl;
    return (uf - gr) - uf;
}
__device__ __forceinline__ float delta_u(const CG& g, const float* rU, const float* cu, const float* p, int om, int i,
                                         int j, int k) {
    const int f = c_u(g, i, j, k);
    const bool ok = (i > 0 && i < g.Wc) || (i == 0 && (om & 1)) || (i == g.Wc && (om & 2));
    return delta_of(rU[f], ok, cu[f], p[c_p(g, i, j, k)], p[c_p(g, i - 1, j, k)]);
}
__device__ __forceinline__ float delta_v(const CG& g, const float* rV, const float* cv, const float* p, int om, int i,
                                         int j, int k) {
    const int f = c_v(g, i, j, k);
    const bool ok = (j > 0 && j < g.Hc) || (j == 0 && (om & 4)) || (j == g.Hc && (om & 8));
    return delta_of(rV[f], ok, cv[f], p[c_p(g, i, j, k)], p[c_p(g, i, j - 1, k)]);
}
__device__ __forceinline__ float delta_w(const CG& g, const float* rW, const float* cw, const float* p, int om, int i,
                                         int j, int k) {
    const int f = c_w(g, i, j, k);
    const bool ok = (k > 0 && k < g.Dc) || (k == 0 && (om & 16)) || (k == g.Dc && (om & 32));
    return delta_of(rW[f], ok, cw[f], p[c_p(g, i, j, k)], p[c_p(g, i, j, k - 1)]);
}

// Fine faces on coarse planes += the delta of their coarse face when the fine cells beside them are
// fluid (ZQ_MapCoarse2Fine3D, :3172-3296).  The reference's x = W faces test the cell at x = W, one
// past the row (:3209); this reads the inside cell x = W-1, as its y and z faces do.  Its y = H
// faces take the delta of coarse v row Hc-1 (:3251), which is kept.
// With guide faces (gU non-null), each delta first gets + gcoef * (guide - coarse face)
// (ZQ_ApplyGuidingVelocity3D, :3162-3167).
__global__ __launch_bounds__(256) void k_coarse_prolong(CG g, const float* __restrict__ rU,
                                                        const float* __restrict__ rV, const float* __restrict__ rW,
                                                        const float* __restrict__ cu, const float* __restrict__ cv,
                                                        const float* __restrict__ cw, const float* __restrict__ p,
                                                        int om, const uint8_t* __restrict__ occ, float* u, float* v,
                                                        float* w, const float* __restrict__ gU,
                                                        const float* __restrict__ gV, const float* __restrict__ gW,
                                                        float gcoef) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = g.s;
    const size_t nu = (size_t)(g.Wc + 1) * g.H * g.D, nv = (size_t)g.W * (g.Hc + 1) * g.D;
    const size_t nw = (size_t)g.W * g.H * (g.Dc + 1);
    if (t < nu) {
        const int ic = (int)(t % (g.Wc + 1)), y = (int)((t / (g.Wc + 1)) % g.H), z = (int)(t / ((size_t)(g.Wc + 1) * g.H));
        const int x = ic * s;
        bool ok;
        if (ic == 0) ok = !occ[f_c(g, 0, y, z)];
        else if (ic == g.Wc) ok = !occ[f_c(g, g.W - 1, y, z)];
        else ok = !occ[f_c(g, x, y, z)] && !occ[f_c(g, x - 1, y, z)];
        if (ok) {
            float dl = delta_u(g, rU, cu, p, om, ic, y / s, z / s);
            if (gU) {
                const int c = c_u(g, ic, y / s, z / s);
                dl = dl + gcoef * (gU[c] - cu[c]);
            }
            const int f = f_u(g, x, y, z);
            u[f] = u[f] + dl;
        }
        return;
    }
    t -= nu;
    if (t < nv) {
        const int x = (int)(t % g.W), jc = (int)((t / g.W) % (g.Hc + 1)), z = (int)(t / ((size_t)g.W * (g.Hc + 1)));
        const int y = jc * s;
        bool ok;
        int row = jc;
        if (jc == 0) ok = !occ[f_c(g, x, 0, z)];
        else if (jc == g.Hc) {
            ok = !occ[f_c(g, x, g.H - 1, z)];
            row = g.Hc - 1;
        } else ok = !occ[f_c(g, x, y, z)] && !occ[f_c(g, x, y - 1, z)];
        if (ok) {
            float dl = delta_v(g, rV, cv, p, om, x / s, row, z / s);
            if (gV) {
                const int c = c_v(g, x / s, row, z / s);
                dl = dl + gcoef * (gV[c] - cv[c]);
            }
            const int f = f_v(g, x, y, z);
            v[f] = v[f] + dl;
        }
        return;
    }
    t -= nv;
    if (t < nw) {
        const int x = (int)(t % g.W), y = (int)((t / g.W) % g.H), kc = (int)(t / ((size_t)g.W * g.H));
        const int z = kc * s;
        bool ok;
        if (kc == 0) ok = !occ[f_c(g, x, y, 0)];
        else if (kc == g.Dc) ok = !occ[f_c(g, x, y, g.D - 1)];
        else ok = !occ[f_c(g, x, y, z)] && !occ[f_c(g, x, y, z - 1)];
        if (ok) {
            float dl = delta_w(g, rW, cw, p, om, x / s, y / s, kc);
            if (gW) {
                const int c = c_w(g, x / s, y / s, kc);
                dl = dl + gcoef * (gW[c] - cw[c]);
            }
            const int f = f_w(g, x, y, z);
            w[f] = w[f] + dl;
        }
    }
}

// u -= p_r - p_l in place on the coarse faces a delta applies to (the projection of the C ABI call)
__global__ __launch_bounds__(256) void k_coarse_correct(CG g, const float* __restrict__ rU,
                                                        const float* __restrict__ rV, const float* __restrict__ rW,
                                                        const float* __restrict__ p, int om, float* u, float* v,
                                                        float* w) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nu = (size_t)(g.Wc + 1) * g.Hc * g.Dc, nv = (size_t)g.Wc * (g.Hc + 1) * g.Dc;
    const size_t nw = (size_t)g.Wc * g.Hc * (g.Dc + 1);
    if (t < nu) {
        const int i = (int)(t % (g.Wc + 1)), j = (int)((t / (g.Wc + 1)) % g.Hc), k = (int)(t / ((size_t)(g.Wc + 1) * g.Hc));
        const bool ok = (i > 0 && i < g.Wc) || (i == 0 && (om & 1)) || (i == g.Wc && (om & 2));
        if (rU[t] > 0.0f && ok) u[t] = u[t] - (p[c_p(g, i, j, k)] - p[c_p(g, i - 1, j, k)]);
        return;
    }
    t -= nu;
    if (t < nv) {
        const int i = (int)(t % g.Wc), j = (int)((t / g.Wc) % (g.Hc + 1)), k = (int)(t / ((size_t)g.Wc * (g.Hc + 1)));
        const bool ok = (j > 0 && j < g.Hc) || (j == 0 && (om & 4)) || (j == g.Hc && (om & 8));
        if (rV[t] > 0.0f && ok) v[t] = v[t] - (p[c_p(g, i, j, k)] - p[c_p(g, i, j - 1, k)]);
        return;
    }
    t -= nv;
    if (t < nw) {
        const int i = (int)(t % g.Wc), j = (int)((t / g.Wc) % g.Hc), k = (int)(t / ((size_t)g.Wc * g.Hc));
        const bool ok = (k > 0 && k < g.Dc) || (k == 0 && (om & 16)) || (k == g.Dc && (om & 32));
        if (rW[t] > 0.0f && ok) w[t] = w[t] - (p[c_p(g, i, j, k)] - p[c_p(g, i, j, k - 1)]);
    }
}

// ------------------------------------------------------------------------------ regular blocks
// DCT-II basis and Neumann eigenvalues of one axis of an s-cell block: q[m*s+n] = c_m cos(pi m (2n+1)/(2s)),
// c_0 = sqrt(1/s), c_m = sqrt(2/s); lam[m] = 4 sin^2(pi m/(2s)).  The block Laplacian (1 per fluid
// neighbour, -count on the diagonal) has eigenvalues -(lam[mx] + lam[my] + lam[mz]).
struct SubTables {
    float q[COARSE_MAX_S * COARSE_MAX_S];
    float lam[COARSE_MAX_S];
};

// A workgroup takes NB consecutive blocks of one block row (NB = 256 / s^2, so one thread per grid
// line of each block: s^2 lines per block along each axis).  Faces are staged through LDS with
// coalesced row loads; each thread runs the 1D transforms of its line in registers; the cell
// buffers pad every x row to s+1 floats so lines along y and z hit distinct banks.
template <int S>
struct RegTile {
    static constexpr int NB = 256 / (S * S);
    static constexpr int TW = NB * S;
    static constexpr int RP = S + 1;                   // padded row
    static constexpr int NCELL = NB * S * S * RP;      // padded cells of the tile
};

template <int S>
__device__ __forceinline__ int rt_cell(int b, int z, int y, int x) {
    return ((b * S + z) * S + y) * (S + 1) + x;
}

// w_first != 0: every interior w face of a block takes the block's first one (z = 1, y = 0, x = 0),
// as the octree's level-1 write-back does
template <int S>
__global__ __launch_bounds__(256) void k_sub_regular(CG g, SubTables tab, const uint8_t* __restrict__ regular, float* u,
                                                     float* v, float* w, int w_first) {
    constexpr int NB = RegTile<S>::NB, TW = RegTile<S>::TW, NC = RegTile<S>::NCELL;
    __shared__ float Us[S][S][TW + 1];
    __shared__ float Vs[S][S + 1][TW];
    __shared__ float Ws[S + 1][S][TW];
    __shared__ float A[NC], B[NC];
    __shared__ uint8_t reg[NB];
    const int tid = threadIdx.x;
    const int bi0 = blockIdx.x * NB, bj = blockIdx.y, bk = blockIdx.z;
    const int nb = min(NB, g.Wc - bi0);
    int mine = 0;
    if (tid < NB) {
        mine = tid < nb ? regular[c_c(g, bi0 + tid, bj, bk)] : 0;
        reg[tid] = (uint8_t)mine;
    }
    if (!__syncthreads_or(mine)) return;
    const int x0 = bi0 * S, y0 = bj * S, z0 = bk * S;
    const int tw = nb * S;
    {
        // all of a thread's face loads in flight before the LDS stores
        constexpr int NU = S * S * (TW + 1), NV = S * (S + 1) * TW, NW = (S + 1) * S * TW;
        constexpr int KU = (NU + 255) / 256, KV = (NV + 255) / 256, KW = (NW + 255) / 256;
        float tu[KU], tv[KV], tw_[KW];
#pragma unroll
        for (int q = 0; q < KU; q++) {
            const int e = tid + 256 * q, x = e % (TW + 1), r = e / (TW + 1);
            tu[q] = (e < NU && x <= tw) ? u[f_u(g, x0 + x, y0 + r % S, z0 + r / S)] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < KV; q++) {
            const int e = tid + 256 * q, x = e % TW, r = e / TW;
            tv[q] = (e < NV && x < tw) ? v[f_v(g, x0 + x, y0 + r % (S + 1), z0 + r / (S + 1))] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < KW; q++) {
            const int e = tid + 256 * q, x = e % TW, r = e / TW;
            tw_[q] = (e < NW && x < tw) ? w[f_w(g, x0 + x, y0 + r % S, z0 + r / S)] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < KU; q++) {
            const int e = tid + 256 * q;
            if (e < NU) Us[(e / (TW + 1)) / S][(e / (TW + 1)) % S][e % (TW + 1)] = tu[q];
        }
#pragma unroll
        for (int q = 0; q < KV; q++) {
            const int e = tid + 256 * q;
            if (e < NV) Vs[(e / TW) / (S + 1)][(e / TW) % (S + 1)][e % TW] = tv[q];
        }
#pragma unroll
        for (int q = 0; q < KW; q++) {
            const int e = tid + 256 * q;
            if (e < NW) Ws[(e / TW) / S][(e / TW) % S][e % TW] = tw_[q];
        }
    }
    __syncthreads();
    const int b = tid / (S * S), l = tid % (S * S);
    const bool act = b < nb;
    const int la = l % S, lb = l / S;
    float q[S][S];
#pragma unroll
    for (int m = 0; m < S; m++)
#pragma unroll
        for (int n = 0; n < S; n++) q[m][n] = tab.q[m * S + n];
    float in[S], out[S];
    // x lines (y = la, z = lb): divergence, forward transform -> A
    if (act) {
        const int xb = b * S;
#pragma unroll
        for (int x = 0; x < S; x++)
            in[x] = Us[lb][la][xb + x + 1] - Us[lb][la][xb + x] + Vs[lb][la + 1][xb + x] - Vs[lb][la][xb + x] +
                    Ws[lb + 1][la][xb + x] - Ws[lb][la][xb + x];
#pragma unroll
        for (int m = 0; m < S; m++) {
            float a = 0.0f;
#pragma unroll
            for (int n = 0; n < S; n++) a += q[m][n] * in[n];
            A[rt_cell<S>(b, lb, la, m)] = a;
        }
    }
    __syncthreads();
    // y lines (mx = la, z = lb): forward -> B
    if (act) {
#pragma unroll
        for (int n = 0; n < S; n++) in[n] = A[rt_cell<S>(b, lb, n, la)];
#pragma unroll
        for (int m = 0; m < S; m++) {
            float a = 0.0f;
#pragma unroll
            for (int n = 0; n < S; n++) a += q[m][n] * in[n];
            B[rt_cell<S>(b, lb, m, la)] = a;
        }
    }
    __syncthreads();
    // z lines (mx = la, my = lb): forward, divide by the eigenvalue, inverse -> A
    if (act) {
#pragma unroll
        for (int n = 0; n < S; n++) in[n] = B[rt_cell<S>(b, n, lb, la)];
        const float lxy = tab.lam[la] + tab.lam[lb];
#pragma unroll
        for (int m = 0; m < S; m++) {
            float a = 0.0f;
#pragma unroll
            for (int n = 0; n < S; n++) a += q[m][n] * in[n];
            const float lam = lxy + tab.lam[m];
            out[m] = (la | lb | m) ? a / (-lam) : 0.0f;
        }
#pragma unroll
        for (int n = 0; n < S; n++) {
            float a = 0.0f;
#pragma unroll
            for (int m = 0; m < S; m++) a += q[m][n] * out[m];
            A[rt_cell<S>(b, n, lb, la)] = a;
        }
    }
    __syncthreads();
    // y lines inverse -> B
    if (act) {
#pragma unroll
        for (int m = 0; m < S; m++) in[m] = A[rt_cell<S>(b, lb, m, la)];
#pragma unroll
        for (int n = 0; n < S; n++) {
            float a = 0.0f;
#pragma unroll
            for (int m = 0; m < S; m++) a += q[m][n] * in[m];
            B[rt_cell<S>(b, lb, n, la)] = a;
        }
    }
    __syncthreads();
    // x lines inverse: the pressure -> A
    if (act) {
#pragma unroll
        for (int m = 0; m < S; m++) in[m] = B[rt_cell<S>(b, lb, la, m)];
#pragma unroll
        for (int n = 0; n < S; n++) {
            float a = 0.0f;
#pragma unroll
            for (int m = 0; m < S; m++) a += q[m][n] * in[m];
            A[rt_cell<S>(b, lb, la, n)] = a;
        }
    }
    __syncthreads();
    // interior faces of the regular blocks: face -= p_r - p_l, written in row order
    for (int e = tid; e < S * S * (TW + 1); e += 256) {
        const int x = e % (TW + 1), r = e / (TW + 1), y = r % S, z = r / S;
        const int bx = x / S, xi = x % S;
        if (xi == 0 || x >= tw || !reg[bx]) continue;
        u[f_u(g, x0 + x, y0 + y, z0 + z)] =
            Us[z][y][x] - (A[rt_cell<S>(bx, z, y, xi)] - A[rt_cell<S>(bx, z, y, xi - 1)]);
    }
    for (int e = tid; e < S * (S + 1) * TW; e += 256) {
        const int x = e % TW, r = e / TW, y = r % (S + 1), z = r / (S + 1);
        const int bx = x / S, xi = x % S;
        if (y == 0 || y == S || x >= tw || !reg[bx]) continue;
        v[f_v(g, x0 + x, y0 + y, z0 + z)] =
            Vs[z][y][x] - (A[rt_cell<S>(bx, z, y, xi)] - A[rt_cell<S>(bx, z, y - 1, xi)]);
    }
    for (int e = tid; e < (S + 1) * S * TW; e += 256) {
        const int x = e % TW, r = e / TW, y = r % S, z = r / S;
        const int bx = x / S, xi = x % S;
        if (z == 0 || z == S || x >= tw || !reg[bx]) continue;
        const float val = w_first ? Ws[1][0][bx * S] - (A[rt_cell<S>(bx, 1, 0, 0)] - A[rt_cell<S>(bx, 0, 0, 0)])
                                  : Ws[z][y][x] - (A[rt_cell<S>(bx, z, y, xi)] - A[rt_cell<S>(bx, z - 1, y, xi)]);
        w[f_w(g, x0 + x, y0 + y, z0 + z)] = val;
    }
}

// ------------------------------------------------------------------------------ blocks with solids
template <int NT>
__device__ __forceinline__ double block_sum(double x, double* red) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wv] = x;
    __syncthreads();
    double a = 0.0;
    for (int i = 0; i < NT / 64; i++) a += red[i];
    __syncthreads();
    return a;
}
template <int NT>
__device__ __forceinline__ int block_min(int x, int* redi) {
    for (int off = 32; off > 0; off >>= 1) x = min(x, __shfl_xor(x, off, 64));
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) redi[wv] = x;
    __syncthreads();
    int a = INT_MAX;
    for (int i = 0; i < NT / 64; i++) a = min(a, redi[i]);
    __syncthreads();
    return a;
}

// ZQ_SolveSubGrid3D (:1393-1521) on one block with solid cells, one workgroup (one thread per cell):
// the closed Poisson problem over the block's fluid cells with unit face weights between fluid cells
// and Neumann walls everywhere else, solved in least squares with the first fluid cell pinned
// (ZQ_InitMatClosedPoisson3D, :603-747; ZQ_SolveClosedPressure3D, :2522-2609).  Its solution is
// the pressure of L p = div - (mean of div over p's fluid component) with p = 0 at the first fluid
// cell on that cell's component and mean 0 on the others (the minimum-norm least-squares choice),
// found here by FP64 CG from 0 to |r| <= 1e-12 |b| (at most the reference's 2(s^3 + 3 s^2 (s+1) + 1)
// iterations).  The pressure is rounded to float, the interior faces get input + delta with delta =
// -(p_r - p_l) between two fluid cells and 0 elsewhere (ZQ_AdjustClosedVelocity3D, :2867-2927).
// Blocks with fewer than two fluid cells are left unchanged (:1492-1497).
template <int S>
__global__ __launch_bounds__(((S * S * S + 63) / 64) * 64) void k_sub_solids(CG g, const uint8_t* __restrict__ regular,
                                                                             const uint8_t* __restrict__ occ, float* u,
                                                                             float* v, float* w) {
    constexpr int N = S * S * S;
    constexpr int NT = ((N + 63) / 64) * 64;
    const int bi = blockIdx.x, bj = blockIdx.y, bk = blockIdx.z;
    if (regular[c_c(g, bi, bj, bk)]) return;
    __shared__ float Us[S][S][S + 1];
    __shared__ float Vs[S][S + 1][S];
    __shared__ float Ws[S + 1][S][S];
    __shared__ uint8_t fl[N];
    __shared__ int lab[N];
    __shared__ double dsh[N];
    __shared__ float p32[N];
    __shared__ double red[NT / 64];
    __shared__ int redi[NT / 64];
    const int t = threadIdx.x;
    const int x0 = bi * S, y0 = bj * S, z0 = bk * S;
    {
        constexpr int NF = S * S * (S + 1), KF = (NF + NT - 1) / NT;
        float tu[KF], tv[KF], tw_[KF];
#pragma unroll
        for (int q = 0; q < KF; q++) {
            const int e = t + NT * q;
            const int xu = e % (S + 1), ru = e / (S + 1), xv = e % S, rv = e / S;
            tu[q] = e < NF ? u[f_u(g, x0 + xu, y0 + ru % S, z0 + ru / S)] : 0.0f;
            tv[q] = e < NF ? v[f_v(g, x0 + xv, y0 + rv % (S + 1), z0 + rv / (S + 1))] : 0.0f;
            tw_[q] = e < NF ? w[f_w(g, x0 + xv, y0 + rv % S, z0 + rv / S)] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < KF; q++) {
            const int e = t + NT * q;
            if (e < NF) {
                Us[(e / (S + 1)) / S][(e / (S + 1)) % S][e % (S + 1)] = tu[q];
                Vs[(e / S) / (S + 1)][(e / S) % (S + 1)][e % S] = tv[q];
                Ws[(e / S) / S][(e / S) % S][e % S] = tw_[q];
            }
        }
    }
    const int ci = t % S, cj = (t / S) % S, ck = t / (S * S);
    const bool me = t < N && !occ[f_c(g, x0 + ci, y0 + cj, z0 + ck)];
    if (t < N) fl[t] = me;
    const int nfl = __syncthreads_count(me);
    if (nfl < 2) return;
    // fluid neighbours inside the block
    int nbr[6];
    int nn = 0;
    if (me) {
        if (ci > 0 && fl[t - 1]) nbr[nn++] = t - 1;
        if (ci < S - 1 && fl[t + 1]) nbr[nn++] = t + 1;
        if (cj > 0 && fl[t - S]) nbr[nn++] = t - S;
        if (cj < S - 1 && fl[t + S]) nbr[nn++] = t + S;
        if (ck > 0 && fl[t - S * S]) nbr[nn++] = t - S * S;
        if (ck < S - 1 && fl[t + S * S]) nbr[nn++] = t + S * S;
    }
    // divergence (left to right, as ZQ_SolveClosedPressure3D, :2569-2571)
    double b = 0.0;
    if (me)
        b = (double)(Us[ck][cj][ci + 1] - Us[ck][cj][ci] + Vs[ck][cj + 1][ci] - Vs[ck][cj][ci] + Ws[ck + 1][cj][ci] -
                     Ws[ck][cj][ci]);
    // fluid components: labels = smallest cell index of the component
    int mylab = me ? t : INT_MAX;
    if (t < N) lab[t] = mylab;
    __syncthreads();
    for (;;) {
        int m = mylab;
        for (int n = 0; n < nn; n++) m = min(m, lab[nbr[n]]);
        __syncthreads();
        const int ch = me && m < mylab;
        if (ch) {
            mylab = m;
            lab[t] = m;
        }
        if (!__syncthreads_or(ch)) break;
    }
    // remove each component's mean divergence, one component at a time in label order
    const int first = block_min<NT>(me ? t : INT_MAX, redi);
    {
        bool pending = me;
        for (;;) {
            const int root = block_min<NT>(pending ? mylab : INT_MAX, redi);
            if (root == INT_MAX) break;
            const bool in = pending && mylab == root;
            const double sum = block_sum<NT>(in ? b : 0.0, red);
            const double cnt = block_sum<NT>(in ? 1.0 : 0.0, red);
            if (in) {
                b -= sum / cnt;
                pending = false;
            }
        }
    }
    // CG on A = -L (cnt p - sum of fluid neighbours), A x = -b, from x = 0
    double x = 0.0, r = -b, d = r;
    double rr = block_sum<NT>(r * r, red);
    const double stop = rr * 1e-24;
    const int maxit = 2 * (N + 3 * S * S * (S + 1) + 1);
    for (int it = 0; it < maxit && rr > stop; it++) {
        if (t < N) dsh[t] = d;
        __syncthreads();
        double q = 0.0;
        if (me) {
            double sn = 0.0;
            for (int n = 0; n < nn; n++) sn += dsh[nbr[n]];
            q = (double)nn * d - sn;
        }
        const double dq = block_sum<NT>(d * q, red);
        if (!(dq > 0.0)) break;
        const double alpha = rr / dq;
        x += alpha * d;
        r -= alpha * q;
        const double rn = block_sum<NT>(r * r, red);
        d = r + (rn / rr) * d;
        rr = rn;
    }
    // pin the first fluid cell's component at that cell
    if (t == first) dsh[0] = x;
    __syncthreads();
    const double x1 = dsh[0];
    if (me && mylab == first) x -= x1;
    if (t < N) p32[t] = me ? (float)x : 0.0f;
    __syncthreads();
    // interior faces: input + delta
    for (int e = t; e < N; e += NT) {
        const int i = e % S, j = (e / S) % S, k = e / (S * S);
        const float pc = p32[e];
        const bool fc = fl[e];
        if (i) {
            const float dl = (fc && fl[e - 1]) ? -(pc - p32[e - 1]) : 0.0f;
            u[f_u(g, x0 + i, y0 + j, z0 + k)] = Us[k][j][i] + dl;
        }
        if (j) {
            const float dl = (fc && fl[e - S]) ? -(pc - p32[e - S]) : 0.0f;
            v[f_v(g, x0 + i, y0 + j, z0 + k)] = Vs[k][j][i] + dl;
        }
        if (k) {
            const float dl = (fc && fl[e - S * S]) ? -(pc - p32[e - S * S]) : 0.0f;
            w[f_w(g, x0 + i, y0 + j, z0 + k)] = Ws[k][j][i] + dl;
        }
    }
}

// ------------------------------------------------------------------------------ batched product
// out[r][c] = sum_k M[r][k] in[k][c], k ascending, one rounding per operation
__global__ __launch_bounds__(256) void k_batched_matvec(int rows, int cols, int n, const float* __restrict__ M,
                                                        const float* __restrict__ in, float* out) {
    __shared__ float Ms[16][17];
    __shared__ float Is[16][17];
    const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
    const int r = blockIdx.y * 16 + tr, c = blockIdx.x * 16 + tc;
    float acc = 0.0f;
    for (int k0 = 0; k0 < cols; k0 += 16) {
        const int mr = blockIdx.y * 16 + tr;
        Ms[tr][tc] = (mr < rows && k0 + tc < cols) ? M[(size_t)mr * cols + k0 + tc] : 0.0f;
        Is[tr][tc] = (k0 + tr < cols && c < n) ? in[(size_t)(k0 + tr) * n + c] : 0.0f;
        __syncthreads();
        const int kn = min(16, cols - k0);
        for (int kk = 0; kk < kn; kk++) acc = acc + Ms[tr][kk] * Is[kk][tc];
        __syncthreads();
    }
    if (r < rows && c < n) out[(size_t)r * n + c] = acc;
}

SubTables make_tables(int s) {
    SubTables t{};
    const double pi = 3.14159265358979323846;
    for (int m = 0; m < s; m++) {
        const double cm = m == 0 ? std::sqrt(1.0 / s) : std::sqrt(2.0 / s);
        for (int n = 0; n < s; n++) t.q[m * s + n] = (float)(cm * std::cos(pi * m * (2 * n + 1) / (2.0 * s)));
        const double sn = std::sin(pi * m / (2.0 * s));
        t.lam[m] = (float)(4.0 * sn * sn);
    }
    return t;
}

template <int S>
void sub_launch(const CG& g, const CoarseBufs& B, const uint8_t* occ, float* u, float* v, float* w, hipStream_t s) {
    constexpr int NB = RegTile<S>::NB;
    const SubTables tab = make_tables(S);
    const dim3 gr((unsigned)((g.Wc + NB - 1) / NB), (unsigned)g.Hc, (unsigned)g.Dc);
    hipLaunchKernelGGL(k_sub_regular<S>, gr, dim3(256), 0, s, g, tab, B.regular, u, v, w, 0);
    constexpr int NT = ((S * S * S + 63) / 64) * 64;
    hipLaunchKernelGGL(k_sub_solids<S>, dim3((unsigned)g.Wc, (unsigned)g.Hc, (unsigned)g.Dc), dim3(NT), 0, s, g,
                       B.regular, occ, u, v, w);
}

}  // namespace

void launch_coarse_setup(const CoarseDims& cd, const uint8_t* occ, int open_mask, const CoarseBufs& B, hipStream_t s,
                         int flux) {
    const CG g = make_cg(cd);
    const size_t nf = cd.Nu() + cd.Nv() + cd.Nw();
    hipLaunchKernelGGL(k_coarse_fractions, dim3(blocks_for(nf, 256)), dim3(256), 0, s, g, occ, B.rU, B.rV, B.rW);
    hipLaunchKernelGGL(k_coarse_cells, dim3(blocks_for(cd.Nc(), 256)), dim3(256), 0, s, g, occ, open_mask,
                       (const float*)B.rU, (const float*)B.rV, (const float*)B.rW, B.cocc, B.inv, B.regular, B.vol, 1,
                       flux);
}

void launch_coarse_restrict(const CoarseDims& cd, const float* u, const float* v, const float* w, const CoarseBufs& B,
                            hipStream_t s) {
    const CG g = make_cg(cd);
    const size_t nf = cd.Nu() + cd.Nv() + cd.Nw();
    hipLaunchKernelGGL(k_coarse_average, dim3(blocks_for(nf, 256)), dim3(256), 0, s, g, u, v, w, B.cu, B.cv, B.cw);
    hipLaunchKernelGGL(k_coarse_div, dim3(blocks_for(cd.Nc(), 256)), dim3(256), 0, s, g, (const float*)B.cu,
                       (const float*)B.cv, (const float*)B.cw, (const uint8_t*)B.cocc, B.div);
    launch_fill_f32(B.p, cd.Np(), 0.0f, s);
    launch_fill_f32(B.p2, cd.Np(), 0.0f, s);
}

void launch_coarse_prepare(const CoarseDims& cd, const float* u, const float* v, const float* w, int open_mask,
                           const CoarseBufs& B, hipStream_t s) {
    const CG g = make_cg(cd);
    // 1/diagonal from the caller's fractions and coarse occupancy (B.rU/rV/rW, B.cocc)
    hipLaunchKernelGGL(k_coarse_cells, dim3(blocks_for(cd.Nc(), 256)), dim3(256), 0, s, g, (const uint8_t*)nullptr,
                       open_mask, (const float*)B.rU, (const float*)B.rV, (const float*)B.rW, B.cocc, B.inv,
                       (uint8_t*)nullptr, (float*)nullptr, 0, 0);
    hipLaunchKernelGGL(k_coarse_div, dim3(blocks_for(cd.Nc(), 256)), dim3(256), 0, s, g, u, v, w,
                       (const uint8_t*)B.cocc, B.div);
    launch_fill_f32(B.p, cd.Np(), 0.0f, s);
    launch_fill_f32(B.p2, cd.Np(), 0.0f, s);
}

constexpr size_t kSlabMaxLds = 160 * 1024;

// LDS of the largest slab a schedule launches (the first workgroup's halo may be clipped, an interior
// one's is not)
void launch_coarse_remove_mean(const CoarseDims& cd, const CoarseBufs& B, hipStream_t s) {
    const int n = (int)cd.Nc();
    hipLaunchKernelGGL(k_coarse_mean, dim3(1), dim3(1024), 0, s, n, (const float*)B.div, (const float*)B.vol,
                       (const uint8_t*)B.cocc, B.mean);
    hipLaunchKernelGGL(k_coarse_sub_mean, dim3(blocks_for((size_t)n, 256)), dim3(256), 0, s, n, B.div,
                       (const float*)B.vol, (const uint8_t*)B.cocc, (const double*)B.mean);
}

size_t coarse_slab_lds(const CoarseDims& cd, CoarseSweep sw) {
    if (sw.k <= 0) return 0;
    size_t most = 0;
    for (int z0 = 0; z0 < cd.Dc; z0 += sw.zc) {
        const SlabLayout L = slab_layout(cd.Dc, z0, sw.k, sw.zc);
        most = std::max(most, slab_floats(cd.Wc, cd.Hc, L.NP, L.NC) * sizeof(float));
    }
    return most;
}

bool slab_lds_enabled() {
    static const bool ok =
        hipFuncSetAttribute((const void*)k_coarse_slab, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSlabMaxLds) ==
        hipSuccess;
    return ok;
}

bool coarse_sweep_valid(const CoarseDims& cd, CoarseSweep sw) {
    if (sw.k == 0) return true;
    if (sw.k < 0 || sw.zc <= 0) return false;
    const size_t lds = coarse_slab_lds(cd, sw);
    return lds <= kSlabMaxLds && (lds <= 64 * 1024 || slab_lds_enabled());
}

const float* launch_coarse_sor(const CoarseDims& cd, const CoarseBufs& B, float omega, int iters, CoarseSweep sw,
                               hipStream_t s) {
    const CG g = make_cg(cd);
    const float omw = 1.0f - omega;
    if (sw.k <= 0) {
        const size_t n = (size_t)((cd.Wc + 1) / 2) * cd.Hc * cd.Dc;
        for (int it = 0; it < iters; it++)
            for (int c = 0; c < 2; c++)
                hipLaunchKernelGGL(k_coarse_sweep, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, (const float*)B.rU,
                                   (const float*)B.rV, (const float*)B.rW, (const float*)B.inv, (const float*)B.div,
                                   B.p, c, omega, omw);
        return B.p;
    }
    float* src = B.p;
    float* dst = B.p2;
    const unsigned nwg = (unsigned)((cd.Dc + sw.zc - 1) / sw.zc);
    for (int done = 0; done < iters;) {
        const int k = std::min(sw.k, iters - done);
        const CoarseSweep cur{k, sw.zc};
        hipLaunchKernelGGL(k_coarse_slab, dim3(nwg), dim3(1024), coarse_slab_lds(cd, cur), s, g, (const float*)B.rU,
                           (const float*)B.rV, (const float*)B.rW, (const float*)B.inv, (const float*)B.div,
                           (const float*)src, dst, omega, omw, k, sw.zc);
        std::swap(src, dst);
        done += k;
    }
    return src;
}

int coarse_sweep_candidates(const CoarseDims& cd, CoarseSweep* out, int max) {
    static const CoarseSweep all[] = {{0, 0}, {1, 1}, {1, 2}, {2, 1}, {2, 2}, {2, 4}, {3, 1}, {3, 2}, {4, 1}, {4, 2},
                                      {4, 4}, {6, 1}, {8, 1}};
    int n = 0;
    for (const CoarseSweep& c : all)
        if (n < max && coarse_sweep_valid(cd, c)) out[n++] = c;
    return n;
}

void launch_coarse_prolong(const CoarseDims& cd, const CoarseBufs& B, const float* p, int open_mask, const uint8_t* occ,
                           float* u, float* v, float* w, hipStream_t s, int guide, float guide_coeff) {
    const CG g = make_cg(cd);
    const size_t n = (size_t)(cd.Wc + 1) * cd.f.H * cd.f.D + (size_t)cd.f.W * (cd.Hc + 1) * cd.f.D +
                     (size_t)cd.f.W * cd.f.H * (cd.Dc + 1);
    hipLaunchKernelGGL(k_coarse_prolong, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, (const float*)B.rU,
                       (const float*)B.rV, (const float*)B.rW, (const float*)B.cu, (const float*)B.cv,
                       (const float*)B.cw, p, open_mask, occ, u, v, w, guide ? (const float*)B.gU : nullptr,
                       guide ? (const float*)B.gV : nullptr, guide ? (const float*)B.gW : nullptr, guide_coeff);
}

void launch_coarse_correct(const CoarseDims& cd, const CoarseBufs& B, const float* p, int open_mask, float* u, float* v,
                           float* w, hipStream_t s) {
    const CG g = make_cg(cd);
    const size_t nf = cd.Nu() + cd.Nv() + cd.Nw();
    hipLaunchKernelGGL(k_coarse_correct, dim3(blocks_for(nf, 256)), dim3(256), 0, s, g, (const float*)B.rU,
                       (const float*)B.rV, (const float*)B.rW, p, open_mask, u, v, w);
}

void launch_sub_project(const CoarseDims& cd, const CoarseBufs& B, const uint8_t* occ, float* u, float* v, float* w,
                        hipStream_t s) {
    const CG g = make_cg(cd);
    switch (cd.s) {
    case 2: sub_launch<2>(g, B, occ, u, v, w, s); break;
    case 3: sub_launch<3>(g, B, occ, u, v, w, s); break;
    case 4: sub_launch<4>(g, B, occ, u, v, w, s); break;
    case 5: sub_launch<5>(g, B, occ, u, v, w, s); break;
    case 6: sub_launch<6>(g, B, occ, u, v, w, s); break;
    case 7: sub_launch<7>(g, B, occ, u, v, w, s); break;
    case 8: sub_launch<8>(g, B, occ, u, v, w, s); break;
    default: break;
    }
}

void launch_sub_regular_mask(const Dims& f, int bs, const uint8_t* mask, float* u, float* v, float* w, int w_first,
                             hipStream_t s) {
    const CG g = make_cg(coarse_dims(f, bs));
    auto go = [&](auto tag) {
        constexpr int S = decltype(tag)::value;
        constexpr int NB = RegTile<S>::NB;
        const dim3 gr((unsigned)((g.Wc + NB - 1) / NB), (unsigned)g.Hc, (unsigned)g.Dc);
        hipLaunchKernelGGL(k_sub_regular<S>, gr, dim3(256), 0, s, g, make_tables(S), mask, u, v, w, w_first);
    };
    switch (bs) {
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    default: break;
    }
}

void launch_batched_matvec(int rows, int cols, int n, const float* M, const float* in, float* out, hipStream_t s) {
    const dim3 gr(blocks_for((size_t)n, 16), blocks_for((size_t)rows, 16));
    hipLaunchKernelGGL(k_batched_matvec, gr, dim3(256), 0, s, rows, cols, n, M, in, out);
}

}  // namespace zq3d
