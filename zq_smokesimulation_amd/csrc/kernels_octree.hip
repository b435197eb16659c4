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
// result is the lists' bit for bit.  The walks run once per projection: they fill compact
// per-level, per-colour cell lists and, for the 2^3 to 8^3 leaves, their neighbour entries (counted,
// scanned, written), so that each of the iters x 8 sweep launches is a gather over a short list
// (a scene far from its obstacles is mostly 8^3 leaves, and the launches touch only those).
#include <algorithm>

#include <hipcub/hipcub.hpp>

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
// Wave-aggregated append of `idx` to `list` where `pred` holds: one atomic per wave and list.
__device__ __forceinline__ void wave_append(bool pred, uint32_t idx, uint32_t* list, uint32_t* count) {
    const unsigned long long mask = __ballot(pred);
    if (!mask) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (pred) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = idx;
}

// The cells each sweep updates, per level and colour (red: even x + y + z at the level): every
// level-0..2 leaf, and every level-3 leaf whose column's block at z = 0 is a leaf (the reference's
// level-3 list test, ZQ_SmokeOctreeGrid3D.cpp:8416; the blocks that test takes that are no leaf are
// never read, so they are left out).  Within a list the order is free: cells of one colour and level
// never neighbour each other.
__global__ __launch_bounds__(256) void k_oct_lists(OG g, const uint8_t* __restrict__ lev, uint32_t* l00, uint32_t* l01,
                                                   uint32_t* l10, uint32_t* l11, uint32_t* l20, uint32_t* l21,
                                                   uint32_t* l30, uint32_t* l31, uint32_t* count) {
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t N = (size_t)g.W * g.H * g.D;
    const bool valid = c < N;
    int x = 0, y = 0, z = 0, m = -1;
    if (valid) {
        x = (int)(c % g.W);
        y = (int)((c / g.W) % g.H);
        z = (int)(c / ((size_t)g.W * g.H));
        m = lev[c];
    }
    uint32_t* lists[4][2] = {{l00, l01}, {l10, l11}, {l20, l21}, {l30, l31}};
#pragma unroll
    for (int L = 0; L < 4; L++) {
        const int mask = (1 << L) - 1;
        bool take = valid && m == L && ((x | y | z) & mask) == 0;
        if (L == 3 && take) take = lev[oat(g, 0, x, y, 0)] == 3;
        const int col = ((x >> L) + (y >> L) + (z >> L)) & 1;
        const uint32_t idx = (uint32_t)oat(g, L, x >> L, y >> L, z >> L);
        wave_append(take && col == 0, idx, lists[L][0], count + 2 * L);
        wave_append(take && col == 1, idx, lists[L][1], count + 2 * L + 1);
    }
}

__global__ void k_oct_zero(uint32_t* count) {
    if (threadIdx.x < 8) count[threadIdx.x] = 0u;
}

// A neighbour entry: level in bits 30-31, the linear index at that level below; OCT_OPEN marks an
// open domain face (its weight enters the diagonal only).
constexpr uint32_t OCT_OPEN = 0xFFFFFFFFu;
constexpr int OCT_HEAD = 6;  // entries per cell kept in the transposed head (one per face)

__device__ __forceinline__ uint32_t oct_entry(const OG& g, int m, int x, int y, int z) {
    return ((uint32_t)m << 30) | (uint32_t)oat(g, m, x >> m, y >> m, z >> m);
}

// The neighbour entries of the level-L cell (x, y, z), in the reference's order, handed to put(k, e);
// returns their number.  Per face: an open domain face; else the leaf behind the face's first fine
// cell when it is as large as this one (at level 0: the fine neighbour unless solid); else the face's
// fine cells in interleaved order (slower axis high), one entry per leaf met (its run of 4^m
// positions skipped), solid cells skipped.
template <class Put>
__device__ int oct_walk(const OG& g, const uint8_t* __restrict__ lev, int L, int x, int y, int z, Put put) {
    const int S = 1 << L;
    const int cl[3] = {x, y, z}, dims[3] = {ow(g, L), oh(g, L), od(g, L)}, base[3] = {x * S, y * S, z * S};
    int k = 0;
    for (int f = 0; f < 6; f++) {
        const int ax = f >> 1, sd = (f & 1) ? 1 : -1;
        if ((sd < 0 && cl[ax] == 0) || (sd > 0 && cl[ax] == dims[ax] - 1)) {
            put(k++, OCT_OPEN);
            continue;
        }
        const int slow = ax == 2 ? 1 : 2, fast = ax == 0 ? 1 : 0;
        int q0[3] = {base[0], base[1], base[2]};
        q0[ax] = sd < 0 ? base[ax] - 1 : base[ax] + S;
        const int m0 = lev[oat(g, 0, q0[0], q0[1], q0[2])];
        if (m0 != OCT_SOLID && m0 >= L) {  // one leaf as large as this one covers the face
            put(k++, oct_entry(g, m0, q0[0], q0[1], q0[2]));
            continue;
        }
        for (int t = 0; t < S * S;) {
            int a = 0, b = 0;
            for (int j = 0; j < L; j++) {
                b |= ((t >> (2 * j)) & 1) << j;
                a |= ((t >> (2 * j + 1)) & 1) << j;
            }
            int q[3] = {q0[0], q0[1], q0[2]};
            q[slow] += a;
            q[fast] += b;
            const int m = lev[oat(g, 0, q[0], q[1], q[2])];
            if (m == OCT_SOLID) {
                t += 1;
                continue;
            }
            put(k++, oct_entry(g, m, q[0], q[1], q[2]));
            t += 1 << (2 * m);  // m < L here: the rest of that leaf's run on this face
        }
    }
    return k;
}

__device__ __forceinline__ void oct_cell(const OG& g, int L, uint32_t c, int& x, int& y, int& z) {
    const uint32_t wl = (uint32_t)ow(g, L), hl = (uint32_t)oh(g, L);
    x = (int)(c % wl);
    y = (int)((c / wl) % hl);
    z = (int)(c / (wl * hl));
}

// entries per listed cell (tot) and beyond the head (ovc, the scan input; ovc[n] = 0 closes it)
__global__ __launch_bounds__(256) void k_oct_count(OG g, const uint8_t* __restrict__ lev, int L,
                                                   const uint32_t* __restrict__ list, uint32_t n,
                                                   uint32_t* __restrict__ tot, uint32_t* __restrict__ ovc) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        ovc[n] = 0;
        return;
    }
    int x, y, z;
    oct_cell(g, L, list[i], x, y, z);
    const int k = oct_walk(g, lev, L, x, y, z, [](int, uint32_t) {});
    tot[i] = (uint32_t)k;
    ovc[i] = k > OCT_HEAD ? (uint32_t)(k - OCT_HEAD) : 0u;
}

// entries: the first OCT_HEAD transposed (entry j of cell i at j * n + i, so a wave's loads of one
// j are contiguous), the rest at the cell's scanned offset in ovf
__global__ __launch_bounds__(256) void k_oct_fill(OG g, const uint8_t* __restrict__ lev, int L,
                                                  const uint32_t* __restrict__ list, uint32_t n,
                                                  uint32_t* __restrict__ head, const uint32_t* __restrict__ off,
                                                  uint32_t* __restrict__ ovf) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int x, y, z;
    oct_cell(g, L, list[i], x, y, z);
    uint32_t* ov = ovf + off[i];
    oct_walk(g, lev, L, x, y, z, [&](int k, uint32_t e) {
        if (k < OCT_HEAD) head[(size_t)k * n + i] = e;
        else ov[k - OCT_HEAD] = e;
    });
}

struct TotPtrs {
    const uint32_t* off[8];
    uint32_t n[8];
};

__global__ void k_oct_totals(TotPtrs t, uint32_t* out) {
    const int k = threadIdx.x;
    if (k < 8) out[k] = t.n[k] ? t.off[k][t.n[k]] : 0u;
}

// right-hand side per listed cell in list order: the level's divergence, with `mean` the closed form
// div - mean * 8^L
__global__ __launch_bounds__(256) void k_oct_rhs(OG g, int L, const uint32_t* __restrict__ list, uint32_t n,
                                                 const float* __restrict__ dl, const float* __restrict__ mean,
                                                 float* __restrict__ rhs) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float dd = dl[list[i]];
    rhs[i] = mean ? dd - *mean * (float)(1 << (3 * L)) : dd;
}

// One listed cell of one colour of one level: coeff += w, sigma += w * p over its entries in order,
// then p = (sigma - rhs) / coeff (CPU_PoissonSolver3D_OpenOctreePoisson.cpp:404-425).  The head
// entries and the right-hand side load without any dependence, then all the pressures they name.
__device__ __forceinline__ void oct_gs_cell(const OctW& wt, const LevelPtrs& P, int L, float* pl,
                                            const uint32_t* __restrict__ list, uint32_t n,
                                            const uint32_t* __restrict__ head, const uint32_t* __restrict__ tot,
                                            const uint32_t* __restrict__ off, const uint32_t* __restrict__ ovf,
                                            const float* __restrict__ rhs, uint32_t i) {
    const uint32_t c = list[i], nt = tot[i];
    const float r = rhs[i];
    uint32_t ev[OCT_HEAD];
    float pv[OCT_HEAD];
#pragma unroll
    for (int j = 0; j < OCT_HEAD; j++) ev[j] = (uint32_t)j < nt ? head[(size_t)j * n + i] : OCT_OPEN;
#pragma unroll
    for (int j = 0; j < OCT_HEAD; j++)
        pv[j] = ((uint32_t)j < nt && ev[j] != OCT_OPEN) ? P.p[ev[j] >> 30][ev[j] & 0x3FFFFFFFu] : 0.0f;
    float coeff = 0.0f, sigma = 0.0f;
#pragma unroll
    for (int j = 0; j < OCT_HEAD; j++) {
        if ((uint32_t)j >= nt) break;
        if (ev[j] == OCT_OPEN) {
            coeff += wt.w[L][3];  // open domain face
        } else {
            const float wg = wt.w[L][ev[j] >> 30];
            coeff += wg;
            sigma += wg * pv[j];
        }
    }
    if (nt > OCT_HEAD) {
        // a face with finer leaves behind it: the rest, eight at a time, in order
        const uint32_t* ov = ovf + off[i];
        const uint32_t no = nt - OCT_HEAD;
        for (uint32_t k0 = 0; k0 < no; k0 += 8) {
            uint32_t e8[8];
            float p8[8];
#pragma unroll
            for (int j = 0; j < 8; j++) e8[j] = k0 + j < no ? ov[k0 + j] : OCT_OPEN;
#pragma unroll
            for (int j = 0; j < 8; j++)
                p8[j] = (k0 + j < no && e8[j] != OCT_OPEN) ? P.p[e8[j] >> 30][e8[j] & 0x3FFFFFFFu] : 0.0f;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (k0 + j >= no) break;
                if (e8[j] == OCT_OPEN) {
                    coeff += wt.w[L][3];
                } else {
                    const float wg = wt.w[L][e8[j] >> 30];
                    coeff += wg;
                    sigma += wg * p8[j];
                }
            }
        }
    }
    sigma -= r;
    pl[c] = sigma / coeff;
}

// One colour of one level, one launch.
__global__ __launch_bounds__(256) void k_oct_gs(OctW wt, LevelPtrs P, int L, float* __restrict__ pl,
                                                const uint32_t* __restrict__ list, uint32_t n,
                                                const uint32_t* __restrict__ head, const uint32_t* __restrict__ tot,
                                                const uint32_t* __restrict__ off, const uint32_t* __restrict__ ovf,
                                                const float* __restrict__ rhs) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) oct_gs_cell(wt, P, L, pl, list, n, head, tot, off, ovf, rhs, i);
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

hipError_t grow(void** p, size_t* cap, size_t need) {
    if (need <= *cap && *p) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t c = need + need / 2 + 4096;
    hipError_t e = hipMalloc(p, c);
    if (e == hipSuccess) *cap = c;
    return e;
}

}  // namespace

size_t octree_bytes(const Dims& g) {
    size_t b = up256(level_cells(g, 0)) + up256(4 * level_cells(g, 0));  // lev, dist
    for (int L = 1; L <= 3; L++) b += up256(level_cells(g, L)) + up256(4 * level_cells(g, L));  // leaf, mn
    for (int L = 0; L <= 3; L++) b += 2 * up256(4 * level_cells(g, L));  // p, d
    for (int L = 0; L <= 3; L++) b += 2 * up256(4 * (level_cells(g, L) / 2 + 1));  // lists
    return b + 256;  // counts
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
    for (int L = 0; L <= 3; L++)
        for (int c = 0; c < 2; c++) B->list[L][c] = (uint32_t*)take(4 * (level_cells(g, L) / 2 + 1));
    B->count = (uint32_t*)take(64);
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

hipError_t octree_build_lists(const Dims& d, const OctreeBufs& B, OctreeLists* Ls, hipStream_t s) {
    const OG g{d.W, d.H, d.D};
    hipLaunchKernelGGL(k_oct_zero, dim3(1), dim3(64), 0, s, B.count);
    hipLaunchKernelGGL(k_oct_lists, dim3(blocks_of(level_cells(d, 0), 256)), dim3(256), 0, s, g,
                       (const uint8_t*)B.lev, B.list[0][0], B.list[0][1], B.list[1][0], B.list[1][1], B.list[2][0],
                       B.list[2][1], B.list[3][0], B.list[3][1], B.count);
    hipError_t e = hipMemcpyAsync(Ls->counts, B.count, 32, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    // per list: head (OCT_HEAD n), tot (n), ovc (n + 1), off (n + 1), rhs (n) in one buffer
    size_t bytes = 0, nmax = 1;
    for (int k = 0; k < 8; k++) {
        const size_t n = Ls->counts[k];
        bytes += up256(4 * OCT_HEAD * n) + up256(4 * n) + 2 * up256(4 * (n + 1)) + up256(4 * n);
        nmax = std::max(nmax, n + 1);
    }
    if ((e = grow(&Ls->dyn, &Ls->dyn_cap, bytes)) != hipSuccess) return e;
    size_t scan = 0;
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)nmax,
                                              s)) != hipSuccess)
        return e;
    if ((e = grow(&Ls->scan, &Ls->scan_cap, scan)) != hipSuccess) return e;
    char* p = (char*)Ls->dyn;
    TotPtrs tp{};
    for (int k = 0; k < 8; k++) {
        const uint32_t n = Ls->counts[k];
        Ls->head[k] = (uint32_t*)p;
        p += up256(4 * OCT_HEAD * (size_t)n);
        Ls->tot[k] = (uint32_t*)p;
        p += up256(4 * (size_t)n);
        Ls->ovc[k] = (uint32_t*)p;
        p += up256(4 * ((size_t)n + 1));
        Ls->off[k] = (uint32_t*)p;
        p += up256(4 * ((size_t)n + 1));
        Ls->rhs[k] = (float*)p;
        p += up256(4 * (size_t)n);
        tp.off[k] = Ls->off[k];
        tp.n[k] = n;
        if (!n) continue;
        hipLaunchKernelGGL(k_oct_count, dim3(blocks_of((size_t)n + 1, 256)), dim3(256), 0, s, g,
                           (const uint8_t*)B.lev, k / 2, (const uint32_t*)B.list[k / 2][k % 2], n, Ls->tot[k],
                           Ls->ovc[k]);
        size_t sb = Ls->scan_cap;
        if ((e = hipcub::DeviceScan::ExclusiveSum(Ls->scan, sb, (const uint32_t*)Ls->ovc[k], Ls->off[k], (int)(n + 1),
                                                  s)) != hipSuccess)
            return e;
    }
    hipLaunchKernelGGL(k_oct_totals, dim3(1), dim3(64), 0, s, tp, B.count + 8);
    uint32_t tot[8];
    e = hipMemcpyAsync(tot, B.count + 8, 32, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    size_t need = 0;
    for (int k = 0; k < 8; k++) {
        Ls->ovf_base[k] = need;
        need += tot[k];
    }
    if ((e = grow((void**)&Ls->ovf, &Ls->ovf_cap, 4 * need + 4)) != hipSuccess) return e;
    for (int k = 0; k < 8; k++) {
        const uint32_t n = Ls->counts[k];
        if (!n) continue;
        hipLaunchKernelGGL(k_oct_fill, dim3(blocks_of(n, 256)), dim3(256), 0, s, g, (const uint8_t*)B.lev, k / 2,
                           (const uint32_t*)B.list[k / 2][k % 2], n, Ls->head[k], (const uint32_t*)Ls->off[k],
                           Ls->ovf + Ls->ovf_base[k]);
    }
    return hipGetLastError();
}

void octree_free_lists(OctreeLists* Ls) {
    if (!Ls) return;
    for (void** p : {&Ls->dyn, (void**)&Ls->ovf, &Ls->scan})
        if (*p) (void)hipFree(*p);
    Ls->dyn = Ls->scan = nullptr;
    Ls->ovf = nullptr;
    Ls->dyn_cap = Ls->ovf_cap = Ls->scan_cap = 0;
}

int launch_octree_sweeps(const Dims& d, const OctreeBufs& B, const OctreeLists& Ls, int iters, const float* mean,
                         hipStream_t s) {
    const OG g{d.W, d.H, d.D};
    const OctW wt = make_weights();
    uint32_t nmax = 0;
    for (int k = 0; k < 8; k++) {
        const uint32_t n = Ls.counts[k];
        nmax = std::max(nmax, n);
        if (n)
            hipLaunchKernelGGL(k_oct_rhs, dim3(blocks_of(n, 256)), dim3(256), 0, s, g, k / 2,
                               (const uint32_t*)B.list[k / 2][k % 2], n, (const float*)B.d[k / 2], mean, Ls.rhs[k]);
    }
    if (iters <= 0 || nmax == 0) return 0;
    const LevelPtrs P{{B.p[0], B.p[1], B.p[2], B.p[3]}};
    int launches = 0;
    for (int it = 0; it < iters; it++)
        for (int k = 0; k < 8; k++) {
            const uint32_t n = Ls.counts[k];
            if (!n) continue;
            hipLaunchKernelGGL(k_oct_gs, dim3(blocks_of(n, 256)), dim3(256), 0, s, wt, P, k / 2, B.p[k / 2],
                               (const uint32_t*)B.list[k / 2][k % 2], n, (const uint32_t*)Ls.head[k],
                               (const uint32_t*)Ls.tot[k], (const uint32_t*)Ls.off[k],
                               (const uint32_t*)(Ls.ovf + Ls.ovf_base[k]), (const float*)Ls.rhs[k]);
            launches++;
        }
    return launches;
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
