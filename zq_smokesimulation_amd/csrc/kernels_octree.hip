// kernels_octree.hip -- gfx950 kernels of the octree projection with the red-black solver
// (OPEN_OCTREE_POISSON_SOR / CLOSED_OCTREE_POISSON_SOR, ZQ_SmokeOctreeGrid3D.cpp).
//
// Build.  The reference fills a distance field breadth-first from the solid cells
// (_seedFilling, :5034-5167); that is the taxicab distance, computed here exactly by three
// separable line passes (each a forward and a backward min-plus scan).  Block minima per level and
// one pass over the fine cells then give the leaves (_buildOctreeInfo, :5377-5534) as one byte per
// fine cell: the level of the leaf holding it.
//
// Solve.  The reference precomputes per-leaf neighbour lists on the host
// (_selectInfos_for_CUDAOpenPoissonRedBlack3, :6330-9305) and hands them to the red-black kernels.
// Here each leaf walks its six faces over the level map instead: the fine cells behind a face,
// visited in the order of their interleaved (slower axis high) coordinates, meet the leaves behind
// the face in exactly the order the reference's lists hold them (same level or coarser first,
// else the finer leaves square by square), and a level-m leaf's cells form one aligned run of 4^m
// positions, skipped at once.  The sums coeff += w, sigma += w * p run in that order, so the
// result is the lists' bit for bit.  One launch per level and colour, the levels in order 0..3.
#include "zq3d_coarse.h"
#include "zq3d_kernels.h"
#include "zq3d_octree.h"

namespace zq3d {
namespace {

constexpr int OCT_INF = 1 << 29;

struct OG {
    int W, H, D;
};

__device__ __forceinline__ int ow(const OG& g, int L) { return g.W >> L; }
__device__ __forceinline__ int oh(const OG& g, int L) { return g.H >> L; }
__device__ __forceinline__ int od(const OG& g, int L) { return g.D >> L; }
__device__ __forceinline__ size_t oat(const OG& g, int L, int x, int y, int z) {
    return ((size_t)z * (g.H >> L) + y) * (size_t)(g.W >> L) + x;
}

// weight[L][M] = 4^min(L,M) / (len_L + len_M), len = 0.5 * 2^L (the reference's weight tables,
// CPU_PoissonSolver3D_OpenOctreePoisson.cpp:408,452,496,540)
struct OctW {
    float w[4][4];
};

OctW make_weights() {
    OctW t{};
    for (int L = 0; L < 4; L++)
        for (int M = 0; M < 4; M++) {
            const int m = L < M ? L : M;
            const float area = (float)(1 << (2 * m));
            const float a = 0.5f * (float)(1 << L), b = 0.5f * (float)(1 << M);
            t.w[L][M] = area / (a + b);
        }
    return t;
}

struct LevelPtrs {
    const float* p[4];
};

// ---------------------------------------------------------------------------- distance field
// x lines: one thread per (y, z) row
__global__ __launch_bounds__(256) void k_oct_dist_x(OG g, const uint8_t* __restrict__ occ, int32_t* __restrict__ dist) {
    const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= (size_t)g.H * g.D) return;
    const size_t base = line * g.W;
    int d = OCT_INF;
    for (int i = 0; i < g.W; i++) {
        d = occ[base + i] ? 0 : min(d + 1, OCT_INF);
        dist[base + i] = d;
    }
    d = OCT_INF;
    for (int i = g.W - 1; i >= 0; i--) {
        d = min(dist[base + i], min(d + 1, OCT_INF));
        dist[base + i] = d;
    }
}

// lines along y (axis 1) or z (axis 2): one thread per (x, other) pair, x fastest across lanes
__global__ __launch_bounds__(256) void k_oct_dist_yz(OG g, int axis, int32_t* __restrict__ dist, int finish) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = axis == 1 ? g.H : g.D;
    const int other = axis == 1 ? g.D : g.H;
    if (t >= (size_t)g.W * other) return;
    const int x = (int)(t % g.W), o = (int)(t / g.W);
    const size_t stride = axis == 1 ? (size_t)g.W : (size_t)g.W * g.H;
    const size_t base = axis == 1 ? (size_t)o * g.H * g.W + x : (size_t)o * g.W + x;
    int d = OCT_INF;
    for (int i = 0; i < n; i++) {
        d = min(dist[base + i * stride], min(d + 1, OCT_INF));
        dist[base + i * stride] = d;
    }
    d = OCT_INF;
    for (int i = n - 1; i >= 0; i--) {
        d = min(dist[base + i * stride], min(d + 1, OCT_INF));
        // no solid cell anywhere: every distance is W + H + D (_buildOctreeInfo, :5365-5373)
        dist[base + i * stride] = (finish && d >= OCT_INF) ? g.W + g.H + g.D : d;
    }
}

// smallest distance in each level-L block from the level below
__global__ __launch_bounds__(256) void k_oct_min(OG g, int L, const int32_t* __restrict__ src, int32_t* __restrict__ dst) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int wl = ow(g, L), hl = oh(g, L);
    if (t >= (size_t)wl * hl * od(g, L)) return;
    const int x = (int)(t % wl), y = (int)((t / wl) % hl), z = (int)(t / ((size_t)wl * hl));
    int m = OCT_INF;
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) m = min(m, src[oat(g, L - 1, 2 * x + i, 2 * y + j, 2 * z + k)]);
    dst[t] = m;
}

// level map and leaf flags: a level-3 block is a leaf when its smallest distance exceeds t3, a
// level-2 block outside level-3 leaves when its smallest distance exceeds t2, a level-1 block
// outside coarser leaves likewise with t1; a fine cell in no coarser leaf is a level-0 leaf unless
// solid
__global__ __launch_bounds__(256) void k_oct_levels(OG g, int t1, int t2, int t3, const uint8_t* __restrict__ occ,
                                                    const int32_t* __restrict__ m1, const int32_t* __restrict__ m2,
                                                    const int32_t* __restrict__ m3, uint8_t* __restrict__ lev,
                                                    uint8_t* __restrict__ l1, uint8_t* __restrict__ l2,
                                                    uint8_t* __restrict__ l3) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= g.W || y >= g.H) return;
    const bool a3 = m3[oat(g, 3, x >> 3, y >> 3, z >> 3)] > t3;
    const bool a2 = !a3 && m2[oat(g, 2, x >> 2, y >> 2, z >> 2)] > t2;
    const bool a1 = !a3 && !a2 && m1[oat(g, 1, x >> 1, y >> 1, z >> 1)] > t1;
    const size_t c = oat(g, 0, x, y, z);
    lev[c] = a3 ? 3 : a2 ? 2 : a1 ? 1 : (occ[c] ? OCT_SOLID : 0);
    if (((x | y | z) & 1) == 0) l1[oat(g, 1, x >> 1, y >> 1, z >> 1)] = a1;
    if (((x | y | z) & 3) == 0) l2[oat(g, 2, x >> 2, y >> 2, z >> 2)] = a2;
    if (((x | y | z) & 7) == 0) l3[oat(g, 3, x >> 3, y >> 3, z >> 3)] = a3;
}

// level map from caller-given leaf flags (the first level whose flag covers the cell; none: solid)
__global__ __launch_bounds__(256) void k_oct_levels_given(OG g, const uint8_t* __restrict__ a0,
                                                          const uint8_t* __restrict__ a1,
                                                          const uint8_t* __restrict__ a2,
                                                          const uint8_t* __restrict__ a3, uint8_t* __restrict__ lev,
                                                          uint8_t* __restrict__ l1, uint8_t* __restrict__ l2,
                                                          uint8_t* __restrict__ l3) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= g.W || y >= g.H) return;
    const size_t c = oat(g, 0, x, y, z);
    const bool b1 = a1[oat(g, 1, x >> 1, y >> 1, z >> 1)] != 0, b2 = a2[oat(g, 2, x >> 2, y >> 2, z >> 2)] != 0;
    const bool b3 = a3[oat(g, 3, x >> 3, y >> 3, z >> 3)] != 0;
    lev[c] = a0[c] ? 0 : b1 ? 1 : b2 ? 2 : b3 ? 3 : OCT_SOLID;
    if (((x | y | z) & 1) == 0) l1[oat(g, 1, x >> 1, y >> 1, z >> 1)] = b1;
    if (((x | y | z) & 3) == 0) l2[oat(g, 2, x >> 2, y >> 2, z >> 2)] = b2;
    if (((x | y | z) & 7) == 0) l3[oat(g, 3, x >> 3, y >> 3, z >> 3)] = b3;
}

// ---------------------------------------------------------------------------- divergence
// fine cells: u(i+1) - u(i) + v(j+1) - v(j) + w(k+1) - w(k), left to right; p0 = 0
__global__ __launch_bounds__(256) void k_oct_div0(OG g, const float* __restrict__ u, const float* __restrict__ v,
                                                  const float* __restrict__ w, float* __restrict__ d0,
                                                  float* __restrict__ p0) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= g.W || y >= g.H) return;
    const size_t cu = ((size_t)z * g.H + y) * (g.W + 1) + x, cv = ((size_t)z * (g.H + 1) + y) * g.W + x;
    const size_t cw = ((size_t)z * g.H + y) * g.W + x;
    d0[cw] = u[cu + 1] - u[cu] + v[cv + g.W] - v[cv] + w[cw + (size_t)g.W * g.H] - w[cw];
    p0[cw] = 0.0f;
}

// level L: the sum of the eight children in x-fastest order, left to right; p_L = 0
__global__ __launch_bounds__(256) void k_oct_restrict(OG g, int L, const float* __restrict__ src,
                                                      float* __restrict__ dst, float* __restrict__ pl) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int wl = ow(g, L), hl = oh(g, L);
    if (t >= (size_t)wl * hl * od(g, L)) return;
    const int x = (int)(t % wl), y = (int)((t / wl) % hl), z = (int)(t / ((size_t)wl * hl));
    const int M = L - 1, X = 2 * x, Y = 2 * y, Z = 2 * z;
    dst[t] = src[oat(g, M, X, Y, Z)] + src[oat(g, M, X + 1, Y, Z)] + src[oat(g, M, X, Y + 1, Z)] +
             src[oat(g, M, X + 1, Y + 1, Z)] + src[oat(g, M, X, Y, Z + 1)] + src[oat(g, M, X + 1, Y, Z + 1)] +
             src[oat(g, M, X, Y + 1, Z + 1)] + src[oat(g, M, X + 1, Y + 1, Z + 1)];
    pl[t] = 0.0f;
}

// ---------------------------------------------------------------------------- Gauss-Seidel
// one neighbour of level m at fine cell (x, y, z): coeff += w, sigma += w * p_m
__device__ __forceinline__ void oct_take(const OG& g, const OctW& wt, const LevelPtrs& P, int L, int m, int x, int y,
                                         int z, float& coeff, float& sigma) {
    const float wg = wt.w[L][m];
    coeff += wg;
    sigma += wg * P.p[m][oat(g, m, x >> m, y >> m, z >> m)];
}

// level-0 leaves of one colour: the six face neighbours read straight off the level map
__global__ __launch_bounds__(256) void k_oct_gs0(OG g, OctW wt, LevelPtrs P, const uint8_t* __restrict__ lev,
                                                 const float* __restrict__ d0, float* __restrict__ p0, int colour,
                                                 const float* __restrict__ mean) {
    const int hx = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    const int x = 2 * hx + ((y + z + colour) & 1);
    if (x >= g.W || y >= g.H) return;
    const size_t c = oat(g, 0, x, y, z);
    if (lev[c] != 0) return;
    float coeff = 0.0f, sigma = 0.0f;
    const int nx[6] = {x - 1, x + 1, x, x, x, x}, ny[6] = {y, y, y - 1, y + 1, y, y}, nz[6] = {z, z, z, z, z - 1, z + 1};
#pragma unroll
    for (int f = 0; f < 6; f++) {
        if (nx[f] < 0 || ny[f] < 0 || nz[f] < 0 || nx[f] >= g.W || ny[f] >= g.H || nz[f] >= g.D) {
            coeff += wt.w[0][3];  // open domain face
            continue;
        }
        const int m = lev[oat(g, 0, nx[f], ny[f], nz[f])];
        if (m != OCT_SOLID) oct_take(g, wt, P, 0, m, nx[f], ny[f], nz[f], coeff, sigma);
    }
    const float dd = d0[c];
    sigma -= mean ? dd - *mean : dd;
    p0[c] = sigma / coeff;
}

// level L >= 1 leaves of one colour: each face's fine cells walked in interleaved order
__global__ __launch_bounds__(64) void k_oct_gsL(OG g, OctW wt, LevelPtrs P, const uint8_t* __restrict__ lev, int L,
                                                const float* __restrict__ dl, float* __restrict__ pl, int colour,
                                                const float* __restrict__ mean) {
    const int wl = ow(g, L), hl = oh(g, L), dlv = od(g, L);
    const int hx = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    const int x = 2 * hx + ((y + z + colour) & 1);
    if (x >= wl) return;
    const int S = 1 << L;
    const int X0 = x * S, Y0 = y * S, Z0 = z * S;
    if (lev[oat(g, 0, X0, Y0, Z0)] != L) return;
    // level 3: the reference's lists take a block when the block at z = 0 of its column is a leaf
    // (ZQ_SmokeOctreeGrid3D.cpp:8416); blocks it takes that are no leaf are never read
    if (L == 3 && lev[oat(g, 0, X0, Y0, 0)] != 3) return;
    float coeff = 0.0f, sigma = 0.0f;
    const int cl[3] = {x, y, z}, dims[3] = {wl, hl, dlv}, base[3] = {X0, Y0, Z0};
    for (int ax = 0; ax < 3; ax++) {
        const int slow = ax == 2 ? 1 : 2, fast = ax == 0 ? 1 : 0;
        for (int sd = -1; sd <= 1; sd += 2) {
            if ((sd < 0 && cl[ax] == 0) || (sd > 0 && cl[ax] == dims[ax] - 1)) {
                coeff += wt.w[L][3];  // open domain face
                continue;
            }
            const int f = sd < 0 ? base[ax] - 1 : base[ax] + S;
            for (int t = 0; t < S * S;) {
                int a = 0, b = 0;
                for (int k = 0; k < L; k++) {
                    b |= ((t >> (2 * k)) & 1) << k;
                    a |= ((t >> (2 * k + 1)) & 1) << k;
                }
                int q[3];
                q[ax] = f;
                q[slow] = base[slow] + a;
                q[fast] = base[fast] + b;
                const int m = lev[oat(g, 0, q[0], q[1], q[2])];
                if (m == OCT_SOLID) {
                    t += 1;
                    continue;
                }
                oct_take(g, wt, P, L, m, q[0], q[1], q[2], coeff, sigma);
                if (m >= L) break;  // a leaf as large as this one covers the whole face
                t += 1 << (2 * m);
            }
        }
    }
    const size_t c = oat(g, L, x, y, z);
    const float dd = dl[c];
    sigma -= mean ? dd - *mean * (float)(1 << (3 * L)) : dd;
    pl[c] = sigma / coeff;
}

// ---------------------------------------------------------------------------- face correction
// the leaf beside a face at coordinate f holding fine cell c: the cell itself (level 0) or a coarser
// leaf whose block the face bounds
__device__ __forceinline__ bool oct_side(const OG& g, const uint8_t* lev, const LevelPtrs& P, int f, int x, int y,
                                         int z, float& len, float& pv) {
    const int m = lev[oat(g, 0, x, y, z)];
    if (m == OCT_SOLID) return false;
    if (m > 0 && (f & ((1 << m) - 1)) != 0) return false;
    len = 0.5f * (float)(1 << m);
    pv = P.p[m][oat(g, m, x >> m, y >> m, z >> m)];
    return true;
}

// face -= (p2 - p1) / (len1 + len2) between two leaves; domain faces against a ghost 4 cells out
// (Adjust_MAC_{u,v,w}_OpenOctreePoisson_Kernel, CPU_PoissonSolver3D_OpenOctreePoisson.cpp:6-373)
__global__ __launch_bounds__(256) void k_oct_adjust(OG g, LevelPtrs P, const uint8_t* __restrict__ lev, float* u,
                                                    float* v, float* w) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nu = (size_t)(g.W + 1) * g.H * g.D, nv = (size_t)g.W * (g.H + 1) * g.D;
    const size_t nw = (size_t)g.W * g.H * (g.D + 1);
    int ax, i, j, k, ex[3];
    float* face;
    if (t < nu) {
        ax = 0;
        ex[0] = g.W + 1, ex[1] = g.H, ex[2] = g.D;
        face = u + t;
    } else if ((t -= nu) < nv) {
        ax = 1;
        ex[0] = g.W, ex[1] = g.H + 1, ex[2] = g.D;
        face = v + t;
    } else if ((t -= nv) < nw) {
        ax = 2;
        ex[0] = g.W, ex[1] = g.H, ex[2] = g.D + 1;
        face = w + t;
    } else {
        return;
    }
    i = (int)(t % ex[0]);
    j = (int)((t / ex[0]) % ex[1]);
    k = (int)(t / ((size_t)ex[0] * ex[1]));
    const int fc[3] = {i, j, k};
    const int f = fc[ax], n = ax == 0 ? g.W : ax == 1 ? g.H : g.D;
    int lo[3] = {i, j, k};
    lo[ax] -= 1;
    float len1, len2, pa, pb;
    if (f == 0) {
        if (!oct_side(g, lev, P, 0, i, j, k, len2, pb)) return;
        *face = *face - (pb - 0.0f) / (4.0f + len2);
    } else if (f == n) {
        if (!oct_side(g, lev, P, f, lo[0], lo[1], lo[2], len1, pa)) return;
        *face = *face - (0.0f - pa) / (4.0f + len1);
    } else {
        if (!oct_side(g, lev, P, f, lo[0], lo[1], lo[2], len1, pa)) return;
        if (!oct_side(g, lev, P, f, i, j, k, len2, pb)) return;
        *face = *face - (pb - pa) / (len1 + len2);
    }
}

unsigned blocks_of(size_t n, int b) { return (unsigned)((n + b - 1) / b); }
dim3 fine_grid(const Dims& g) { return dim3((unsigned)((g.W + 63) / 64), (unsigned)((g.H + 3) / 4), (unsigned)g.D); }
size_t level_cells(const Dims& g, int L) { return (size_t)(g.W >> L) * (g.H >> L) * (g.D >> L); }
size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

size_t octree_bytes(const Dims& g) {
    size_t b = up256(level_cells(g, 0)) + up256(4 * level_cells(g, 0));  // lev, dist
    for (int L = 1; L <= 3; L++) b += up256(level_cells(g, L)) + up256(4 * level_cells(g, L));  // leaf, mn
    for (int L = 0; L <= 3; L++) b += 2 * up256(4 * level_cells(g, L));  // p, d
    return b;
}

void octree_layout(const Dims& g, char* mem, OctreeBufs* B) {
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char* p = mem + o;
        o += up256(bytes);
        return p;
    };
    B->lev = (uint8_t*)take(level_cells(g, 0));
    B->dist = (int32_t*)take(4 * level_cells(g, 0));
    B->leaf[0] = nullptr;
    B->mn[0] = B->dist;
    for (int L = 1; L <= 3; L++) {
        B->leaf[L] = (uint8_t*)take(level_cells(g, L));
        B->mn[L] = (int32_t*)take(4 * level_cells(g, L));
    }
    for (int L = 0; L <= 3; L++) {
        B->p[L] = (float*)take(4 * level_cells(g, L));
        B->d[L] = (float*)take(4 * level_cells(g, L));
    }
}

void launch_octree_build(const Dims& d, const uint8_t* occ, int t1, int t2, int t3, const OctreeBufs& B,
                         hipStream_t s) {
    const OG g{d.W, d.H, d.D};
    hipLaunchKernelGGL(k_oct_dist_x, dim3(blocks_of((size_t)d.H * d.D, 256)), dim3(256), 0, s, g, occ, B.dist);
    hipLaunchKernelGGL(k_oct_dist_yz, dim3(blocks_of((size_t)d.W * d.D, 256)), dim3(256), 0, s, g, 1, B.dist, 0);
    hipLaunchKernelGGL(k_oct_dist_yz, dim3(blocks_of((size_t)d.W * d.H, 256)), dim3(256), 0, s, g, 2, B.dist, 1);
    for (int L = 1; L <= 3; L++)
        hipLaunchKernelGGL(k_oct_min, dim3(blocks_of(level_cells(d, L), 256)), dim3(256), 0, s, g, L,
                           (const int32_t*)B.mn[L - 1], B.mn[L]);
    hipLaunchKernelGGL(k_oct_levels, fine_grid(d), dim3(64, 4), 0, s, g, t1, t2, t3, occ, (const int32_t*)B.mn[1],
                       (const int32_t*)B.mn[2], (const int32_t*)B.mn[3], B.lev, B.leaf[1], B.leaf[2], B.leaf[3]);
}

void launch_octree_from_leaves(const Dims& d, const uint8_t* l0, const uint8_t* l1, const uint8_t* l2,
                               const uint8_t* l3, const OctreeBufs& B, hipStream_t s) {
    const OG g{d.W, d.H, d.D};
    hipLaunchKernelGGL(k_oct_levels_given, fine_grid(d), dim3(64, 4), 0, s, g, l0, l1, l2, l3, B.lev, B.leaf[1],
                       B.leaf[2], B.leaf[3]);
}

void launch_octree_prepare(const Dims& d, const float* u, const float* v, const float* w, const OctreeBufs& B,
                           hipStream_t s) {
    const OG g{d.W, d.H, d.D};
    hipLaunchKernelGGL(k_oct_div0, fine_grid(d), dim3(64, 4), 0, s, g, u, v, w, B.d[0], B.p[0]);
    for (int L = 1; L <= 3; L++)
        hipLaunchKernelGGL(k_oct_restrict, dim3(blocks_of(level_cells(d, L), 256)), dim3(256), 0, s, g, L,
                           (const float*)B.d[L - 1], B.d[L], B.p[L]);
}

void launch_octree_sweeps(const Dims& d, const OctreeBufs& B, int iters, const float* mean, hipStream_t s) {
    const OG g{d.W, d.H, d.D};
    const OctW wt = make_weights();
    const LevelPtrs P{{B.p[0], B.p[1], B.p[2], B.p[3]}};
    const dim3 g0((unsigned)((d.W / 2 + 63) / 64), (unsigned)((d.H + 3) / 4), (unsigned)d.D);
    for (int it = 0; it < iters; it++) {
        for (int c = 0; c < 2; c++)
            hipLaunchKernelGGL(k_oct_gs0, g0, dim3(64, 4), 0, s, g, wt, P, (const uint8_t*)B.lev,
                               (const float*)B.d[0], B.p[0], c, mean);
        for (int L = 1; L <= 3; L++) {
            const dim3 gl((unsigned)(((d.W >> L) / 2 + 63) / 64), (unsigned)(d.H >> L), (unsigned)(d.D >> L));
            for (int c = 0; c < 2; c++)
                hipLaunchKernelGGL(k_oct_gsL, gl, dim3(64), 0, s, g, wt, P, (const uint8_t*)B.lev, L,
                                   (const float*)B.d[L], B.p[L], c, mean);
        }
    }
}

void launch_octree_adjust(const Dims& d, const OctreeBufs& B, float* u, float* v, float* w, hipStream_t s) {
    const OG g{d.W, d.H, d.D};
    const LevelPtrs P{{B.p[0], B.p[1], B.p[2], B.p[3]}};
    const size_t nf = (size_t)(d.W + 1) * d.H * d.D + (size_t)d.W * (d.H + 1) * d.D + (size_t)d.W * d.H * (d.D + 1);
    hipLaunchKernelGGL(k_oct_adjust, dim3(blocks_of(nf, 256)), dim3(256), 0, s, g, P, (const uint8_t*)B.lev, u, v, w);
}

void launch_octree_sub_project(const Dims& d, const OctreeBufs& B, float* u, float* v, float* w, hipStream_t s) {
    // level 1 keeps the reference's w write-back (ZQ_SmokeOctreeGrid3D.cpp:1125-1135)
    launch_sub_regular_mask(d, 2, B.leaf[1], u, v, w, 1, s);
    launch_sub_regular_mask(d, 4, B.leaf[2], u, v, w, 0, s);
    launch_sub_regular_mask(d, 8, B.leaf[3], u, v, w, 0, s);
}

}  // namespace zq3d
