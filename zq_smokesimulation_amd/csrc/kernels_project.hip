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
//
// Reductions (closed-domain mean, residual and divergence maxima) are two-level and deterministic:
// one partial per block (wave butterfly + LDS), then one block folds the partials in index order.
#include "zq3d_device.h"
#include "zq3d_kernels.h"
#include "zq3d_sor_split.h"
#include "zq3d_stencils.h"

namespace zq3d {

namespace {

constexpr int BX = 64, BY = 4;
// planes per thread of the z-marching kernels (each thread owns one (i, j) column of a chunk)
constexpr int RES_RZ = 16;

inline dim3 grid_for(int nx, int ny, int nz) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz); }
inline dim3 grid_z(const Dims& d) { return grid_for(d.W, d.H, (d.D + RES_RZ - 1) / RES_RZ); }

__device__ __forceinline__ int block_linear() { return (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x; }

// Neighbour counts, divergence (the reference's unscaled sum, ZQ_SmokeSimulationUtils3D.cpp:2403-2405,
// evaluated left to right), zero pressure where the count is 0, and per-block FP64 sum/count of the
// divergence over fluid cells (closed-domain mean).
__global__ __launch_bounds__(BX* BY) void k_proj_setup(const float* __restrict__ u, const float* __restrict__ v,
                                                       const float* __restrict__ w, const uint8_t* __restrict__ occ,
                                                       int open_mask, uint8_t* __restrict__ nc,
                                                       float* __restrict__ b, float* __restrict__ p,
                                                       double* __restrict__ part, Dims d) {
    __shared__ double red[4];
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double dsum = 0.0, dcnt = 0.0;
    st_proj_setup_plain(u, v, w, occ, open_mask, nc, b, p, d, i, j, k, dsum, dcnt);
    if (part) {
        const double s = block_reduce<false>(dsum, red);
        __syncthreads();
        const double c = block_reduce<false>(dcnt, red);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            part[2 * block_linear()] = s;
            part[2 * block_linear() + 1] = c;
        }
    }
}

// The mixed-precision set-up: neighbour counts, FP64 divergence b64 (the same sum as the plain
// set-up, in double), p64 = 0 where the count is 0, per-block FP64 sum/count of the divergence over
// fluid cells.  One (i, j) column per thread up a chunk of RES_RZ planes: the occupancy of planes
// k-1, k, k+1 and the w face between planes stay in registers, so every input is read once.
__global__ __launch_bounds__(BX* BY) void k_proj_setup_z(const float* __restrict__ u, const float* __restrict__ v,
                                                         const float* __restrict__ w, const uint8_t* __restrict__ occ,
                                                         int open_mask, uint8_t* __restrict__ nc,
                                                         double* __restrict__ b64, double* __restrict__ p64,
                                                         double* __restrict__ part, Dims d) {
    __shared__ double red[4];
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    double dsum = 0.0, dcnt = 0.0;
    if (i < d.W && j < d.H) {
        const int HW = d.W * d.H;
        const int k0 = blockIdx.z * RES_RZ, k1 = min(k0 + RES_RZ, d.D);
        int o = cidx(d, i, j, k0);
        int ob = k0 > 0 ? occ[o - HW] : 1;  // plane below (not read at k = 0)
        int oc = occ[o];
        float w0 = w[widx(d, i, j, k0)];
        for (int k = k0; k < k1; k++, o += HW) {
            const int oa = k < d.D - 1 ? occ[o + HW] : 1;  // plane above (not read at k = D-1)
            const bool fluid = !oc;
            // the in-plane neighbours load together, whatever this cell is (an index off the grid is
            // replaced by the cell's own and its value not used)
            const int oxl = occ[i > 0 ? o - 1 : o], oxh = occ[i < d.W - 1 ? o + 1 : o];
            const int oyl = occ[j > 0 ? o - d.W : o], oyh = occ[j < d.H - 1 ? o + d.W : o];
            int n = 0;
            if (fluid) {
                n += (i > 0) ? !oxl : ((open_mask & 1) != 0);
                n += (i < d.W - 1) ? !oxh : ((open_mask & 2) != 0);
                n += (j > 0) ? !oyl : ((open_mask & 4) != 0);
                n += (j < d.H - 1) ? !oyh : ((open_mask & 8) != 0);
                n += (k > 0) ? !ob : ((open_mask & 16) != 0);
                n += (k < d.D - 1) ? !oa : ((open_mask & 32) != 0);
            }
            nc[o] = (uint8_t)n;
            const float u0 = u[uidx(d, i, j, k)], u1 = u[uidx(d, i + 1, j, k)];
            const float v0 = v[vidx(d, i, j, k)], v1 = v[vidx(d, i, j + 1, k)];
            const float w1 = w[widx(d, i, j, k + 1)];
            const double dd = (((double)u1 - (double)u0) + ((double)v1 - (double)v0)) + ((double)w1 - (double)w0);
            b64[o] = n ? dd : 0.0;
            if (!n) p64[pidx(d, i, j, k)] = 0.0;
            if (fluid) {
                dsum += dd;
                dcnt += 1.0;
            }
            ob = oc;
            oc = oa;
            w0 = w1;
        }
    }
    if (part) {
        const double s = block_reduce<false>(dsum, red);
        __syncthreads();
        const double c = block_reduce<false>(dcnt, red);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            part[2 * block_linear()] = s;
            part[2 * block_linear() + 1] = c;
        }
    }
}

// Fold the per-block (sum, count) pairs in index order; mean = sum/count when want != 0.
__global__ __launch_bounds__(1024) void k_finish_mean(const double* __restrict__ part, int nblocks,
                                                      float* __restrict__ mean32, double* __restrict__ mean64,
                                                      int want) {
    __shared__ double ss[1024], sc[1024];
    st_fold_mean(part, nblocks, mean32, mean64, want, ss, sc);
}

// Fold per-block maxima; the result is stored as the bit pattern of a non-negative double.
__global__ __launch_bounds__(1024) void k_finish_max(const double* __restrict__ part, int nblocks,
                                                     unsigned long long* __restrict__ out) {
    __shared__ double sm[1024];
    const int t = threadIdx.x;
    double m = 0.0;
    for (int i = t; i < nblocks; i += 1024) m = fmax(m, part[i]);
    sm[t] = m;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h) sm[t] = fmax(sm[t], sm[t + h]);
        __syncthreads();
    }
    if (t == 0) *out = (unsigned long long)__double_as_longlong(sm[0]);
}

// One colour of a red-black SOR iteration, in place.  Thread (t, j, k) owns cell
// x = 2t + ((colour+j+k)&1).  Used when the fused kernel's plane ring does not fit in LDS.
__global__ __launch_bounds__(BX* BY) void k_sor_half(const uint8_t* __restrict__ nc, const float* __restrict__ b,
                                                     const float* __restrict__ mean32, float* __restrict__ p, Dims d,
                                                     float omega, float omw, int colour) {
    st_sor_half(nc, b, mean32, p, d, omega, omw, colour, blockIdx.x * BX + threadIdx.x, blockIdx.y * BY + threadIdx.y,
                blockIdx.z);
}

// One full red-black iteration p_in -> p_out in a single pass over HBM.
//
// A block owns TY whole rows (all of x, so no x halo) of a chunk of zc planes and walks z with a
// ring of four planes in LDS, each plane holding rows y0-2 .. y0+TY+1 of the old pressure.  At
// step zz it updates colour 0 of plane zz (tile rows plus one halo row on each side, recomputed
// here so no other block's result is needed), then colour 1 of plane zz-1 (tile rows; every
// neighbour it reads is a colour-0 value already updated in LDS), and stores plane zz-1.
// Colour 0 reads only old colour-1 values and colour 1 only new colour-0 values, exactly as in two
// single-colour sweeps, so results are bit-identical to the in-place half sweeps.  The chunk
// recomputes colour 0 on the plane before and after it for the same reason.
// Plane zz+2 is read into registers while zz is computed (one float per thread per row for each
// 256 columns, CPT = columns per thread), hiding the HBM latency behind the updates.  The ghost
// columns x = -1 and x = W are zero in every plane and are set once.
// Blocks are numbered so each XCD (hardware dispatches block b to XCD b % 8) receives a contiguous
// run of tiles, which share halo rows through that XCD's L2.
template <int TY, int CPT>
__global__ __launch_bounds__(256) void k_sor_fused(const uint8_t* __restrict__ nc, const float* __restrict__ b,
                                                   const float* __restrict__ mean32, const float* __restrict__ pin,
                                                   float* __restrict__ pout, Dims d, float omega, float omw, int zc,
                                                   int ntiles_y) {
    extern __shared__ float lds[];
    constexpr int ROWS = TY + 4;
    constexpr int NT = 256;
    const int tid = threadIdx.x;
    const int per = gridDim.x >> 3;  // launched with a multiple of 8 blocks
    const int bid = blockIdx.x;
    const int tile = (bid & 7) * per + (bid >> 3);
    const int ty = tile % ntiles_y, tz = tile / ntiles_y;
    const int y0 = ty * TY, z0 = tz * zc;
    const int z1 = min(z0 + zc, d.D);
    if (z0 >= d.D) return;
    const int WP = d.W + 2;
    const int PLANE = ROWS * WP;
    const int sy = d.W + 2, sz = (d.W + 2) * (d.H + 2);
    const float m = *mean32;
    const int hb = NT / 2;  // threads per row when two rows are processed together
    const int sub = tid >= hb ? 1 : 0;
    const int lt = tid - sub * hb;

    auto buf = [&](int z) { return lds + (z & 3) * PLANE; };  // z >= -2: (z & 3) is z mod 4

    // ghost columns of every row of the ring
    for (int i = tid; i < 4 * ROWS; i += NT) {
        lds[i * WP] = 0.0f;
        lds[i * WP + WP - 1] = 0.0f;
    }

    float pf[ROWS][CPT];
    auto fetch = [&](int z) {  // plane z, rows y0-2.., interior columns, into registers
        const bool zok = z >= -1 && z <= d.D;
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            const int y = y0 - 2 + r;
            const bool ok = zok && y >= -1 && y <= d.H;
            const float* src = pin + (size_t)(z + 1) * sz + (size_t)(y + 1) * sy + 1;
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = tid + c * NT;
                pf[r][c] = (ok && x < d.W) ? src[x] : 0.0f;
            }
        }
    };
    auto commit = [&](int z) {  // registers -> LDS plane z
        float* B = buf(z);
#pragma unroll
        for (int r = 0; r < ROWS; r++)
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = tid + c * NT;
                if (x < d.W) B[r * WP + x + 1] = pf[r][c];
            }
    };

    // Each thread updates rows first+sub, first+sub+2, ... and, in each, the cells of the colour at
    // x = 2*(lt + c*hb) + parity.  Their counts and right-hand sides are read one step ahead into
    // registers (cur/next pairs) so no update waits on HBM.
    constexpr int RR = (TY + 2) / 2;  // colour-0 rows per thread (tile rows plus halo rows)
    constexpr int RB = TY / 2;        // colour-1 rows per thread (tile rows)
    uint8_t n0[RR][CPT], n0x[RR][CPT], n1[RB][CPT], n1x[RB][CPT];
    float b0[RR][CPT], b0x[RR][CPT], b1[RB][CPT], b1x[RB][CPT];

    auto fetch_rhs = [&](int z, int col, int first, int nrows, auto& nn, auto& bv) {
        for (int k = 0; k < nrows; k++) {
            const int r = first + sub + 2 * k;
            const int y = y0 - 2 + r;
            const int par = (col + y + z) & 1;
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = 2 * (lt + c * hb) + par;
                const bool ok = z >= 0 && z < d.D && y >= 0 && y < d.H && x < d.W;
                const int o = (z * d.H + y) * d.W + x;
                nn[k][c] = ok ? nc[o] : (uint8_t)0;
                bv[k][c] = ok ? b[o] : 0.0f;
            }
        }
    };

    auto update = [&](int z, int col, int first, int nrows, auto& nn, auto& bv) {
        if (z < 0 || z >= d.D) return;
        float* Bc = buf(z);
        const float* Bm = buf(z - 1);
        const float* Bp = buf(z + 1);
        for (int k = 0; k < nrows; k++) {
            const int r = first + sub + 2 * k;
            const int y = y0 - 2 + r;
            if (y < 0 || y >= d.H) continue;
            const int par = (col + y + z) & 1;
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int x = 2 * (lt + c * hb) + par;
                const int n = nn[k][c];
                if (x >= d.W || !n) continue;
                const int q = r * WP + x + 1;
                const float s = Bc[q - 1] + Bc[q + 1] + Bc[q - WP] + Bc[q + WP] + Bm[q] + Bp[q];
                const float bb = bv[k][c] - m;
                const float pg = (s - bb) * rcp_count(n);
                Bc[q] = omw * Bc[q] + omega * pg;
            }
        }
    };

    auto store_plane = [&](int z) {
        const float* B = buf(z);
        for (int r = 2; r < TY + 2; r++) {
            const int y = y0 - 2 + r;
            if (y >= d.H) break;
            float* dst = pout + (size_t)(z + 1) * sz + (size_t)(y + 1) * sy + 1;
            for (int x = tid; x < d.W; x += NT) dst[x] = B[r * WP + x + 1];
        }
    };

    fetch(z0 - 2);
    commit(z0 - 2);
    fetch(z0 - 1);
    commit(z0 - 1);
    fetch(z0);
    commit(z0);
    fetch(z0 + 1);
    fetch_rhs(z0 - 1, 0, 1, RR, n0, b0);
    __syncthreads();
    for (int zz = z0 - 1; zz <= z1; zz++) {
        // LDS: planes zz-1, zz, zz+1; registers: plane zz+2, colour-0 rhs of zz, colour-1 rhs of zz-1
        fetch_rhs(zz + 1, 0, 1, RR, n0x, b0x);
        fetch_rhs(zz, 1, 2, RB, n1x, b1x);
        update(zz, 0, 1, RR, n0, b0);  // colour 0, rows y0-1 .. y0+TY
        __syncthreads();
        if (zz - 1 >= z0) {
            update(zz - 1, 1, 2, RB, n1, b1);  // colour 1, rows y0 .. y0+TY-1
            __syncthreads();
            store_plane(zz - 1);
        }
        commit(zz + 2);  // into the slot of plane zz-2, no longer read
        if (zz + 1 <= z1) fetch(zz + 3);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RR; k++)
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                n0[k][c] = n0x[k][c];
                b0[k][c] = b0x[k][c];
            }
#pragma unroll
        for (int k = 0; k < RB; k++)
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                n1[k][c] = n1x[k][c];
                b1[k][c] = b1x[k][c];
            }
    }
}

// interleaved (nc, b, padded p) -> split (n_c, b_c, p_c), both colours; blockIdx.z = 2*z + colour
__global__ __launch_bounds__(BX* BY) void k_split_in(const uint8_t* __restrict__ nc, const float* __restrict__ b,
                                                     const float* __restrict__ p, uint8_t* __restrict__ n0,
                                                     uint8_t* __restrict__ n1, float* __restrict__ b0,
                                                     float* __restrict__ b1, float* __restrict__ p0,
                                                     float* __restrict__ p1, Dims d, int Wh) {
    const int xh = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    const int z = blockIdx.z >> 1, c = blockIdx.z & 1;
    if (xh >= Wh || y >= d.H) return;
    const int x = 2 * xh + ((c + y + z) & 1);
    const int oh = (z * d.H + y) * Wh + xh;
    const int qh = split_q(Wh, d.H, xh, y, z);
    const bool ok = x < d.W;
    const uint8_t n = ok ? nc[cidx(d, x, y, z)] : (uint8_t)0;
    const float bv = ok ? b[cidx(d, x, y, z)] : 0.0f;
    const float pv = ok ? p[pidx(d, x, y, z)] : 0.0f;
    (c ? n1 : n0)[oh] = n;
    (c ? b1 : b0)[oh] = bv;
    (c ? p1 : p0)[qh] = pv;
}

// split padded p_c -> interleaved padded p (interior cells)
__global__ __launch_bounds__(BX* BY) void k_split_out(const float* __restrict__ p0, const float* __restrict__ p1,
                                                      float* __restrict__ p, Dims d, int Wh) {
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= d.W || y >= d.H) return;
    const int c = (x + y + z) & 1;
    p[pidx(d, x, y, z)] = (c ? p1 : p0)[split_q(Wh, d.H, x >> 1, y, z)];
}

// interleaved counts -> split n_c, both colours (slots past a row's end hold 0); blockIdx.z = 2*z + colour
__global__ __launch_bounds__(BX* BY) void k_split_counts(const uint8_t* __restrict__ nc, uint8_t* __restrict__ n0,
                                                         uint8_t* __restrict__ n1, Dims d, int Wh) {
    const int xh = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y * BY + threadIdx.y;
    const int z = blockIdx.z >> 1, c = blockIdx.z & 1;
    if (xh >= Wh || y >= d.H) return;
    const int x = 2 * xh + ((c + y + z) & 1);
    (c ? n1 : n0)[(z * d.H + y) * Wh + xh] = x < d.W ? nc[cidx(d, x, y, z)] : (uint8_t)0;
}

// One colour of an SOR iteration on split storage: pc (this colour) updated in place from po.
// 1-D grid; each XCD (block b runs on XCD b % 8) receives a contiguous z-slab of tiles so the
// plane above and below a tile are usually in the same L2.
__global__ __launch_bounds__(BX* BY) void k_sor_split(const uint8_t* __restrict__ nn, const float* __restrict__ bb,
                                                      const float* __restrict__ mean32, float* __restrict__ pc,
                                                      const float* __restrict__ po, Dims d, int Wh, int gx, int gy,
                                                      float omega, float omw, int colour) {
    const int per = gridDim.x >> 3;
    const int bid = blockIdx.x;
    const int tile = (bid & 7) * per + (bid >> 3);
    const int bx = tile % gx;
    const int rest = tile / gx;
    const int by = rest % gy;
    const int z = rest / gy;
    const int xh = bx * BX + (threadIdx.x & (BX - 1));
    const int y = by * BY + (threadIdx.x >> 6);
    if (z >= d.D || xh >= Wh || y >= d.H) return;
    const int par = (colour + y + z) & 1;
    if (2 * xh + par >= d.W) return;
    const int oh = (z * d.H + y) * Wh + xh;
    const int n = nn[oh];
    if (!n) return;
    const int SP = split_row(Wh), SZ = SP * (d.H + 2);
    const int q = split_q(Wh, d.H, xh, y, z);
    const float s = po[q - 1 + par] + po[q + par] + po[q - SP] + po[q + SP] + po[q - SZ] + po[q + SZ];
    const float bv = bb[oh] - *mean32;
    const float pg = (s - bv) * rcp_count(n);
    pc[q] = omw * pc[q] + omega * pg;
}

// One red+black iteration per launch on split storage (sor_split_rb_tile); 1-D grid of tiles
// (SPLIT_TX slots x TY rows x zc planes), numbered so each XCD (block b runs on XCD b % 8) receives
// a contiguous run of tiles, which share their halos through that XCD's L2.
template <int TY, int NW, bool ZS = false>
__global__ __launch_bounds__(64 * NW) void k_sor_split_rb(const uint8_t* __restrict__ n0, const uint8_t* __restrict__ n1,
                                                          const float* __restrict__ b0, const float* __restrict__ b1,
                                                          const float* __restrict__ mean32,
                                                          const float* __restrict__ s0, const float* __restrict__ s1,
                                                          float* __restrict__ d0, float* __restrict__ d1, Dims d,
                                                          int Wh, int gx, int gy, int zc, float omega, float omw) {
    const int per = gridDim.x >> 3;
    const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int bx = tile % gx;
    const int rest = tile / gx;
    const int by = rest % gy;
    const int z0 = (rest / gy) * zc;
    if (z0 >= d.D) return;  // spare block (the grid is a multiple of 8): uniform exit before any barrier
    sor_split_rb_tile<SPLIT_TX, TY, NW, false, ZS>(n0, n1, b0, b1, *mean32, s0, s1, d0, d1, d, Wh, bx * SPLIT_TX,
                                                  by * TY, z0, min(z0 + zc, d.D), omega, omw);
}

// Iteration `it` of a gated pass (k_sor_split_rb otherwise): it returns at once when the previous
// iteration's largest update, gate[it - 1] (FP32 bits; 0 when that iteration did not run either),
// is at most `thresh`, or, past iteration 2 GATE_K, no smaller than gate[it - 1 - GATE_K] (the
// updates have stopped shrinking: the FP32 iterate's rounding floor; the refinement's next FP64
// residual takes over); otherwise it runs the iteration, folds its largest update into gate[it]
// and counts itself in *done, copied to the mapped host word *host_done for the host's progress
// check.  gate[it] is GATE_LANES words (the maximum is their maximum): each workgroup reduces its
// waves in LDS and adds one atomic to word blockIdx % GATE_LANES, so no address takes more than a
// few dozen atomics per launch (thousands of atomics on one word serialise and doubled the
// iteration's time).  All workgroups read the same words, written by the previous launch, so they
// all run or all return.  ZS: the first iteration of a pass (zero source).
template <int TY, int NW, bool ZS>
__global__ __launch_bounds__(64 * NW) void k_sor_split_rb_gated(
    const uint8_t* __restrict__ n0, const uint8_t* __restrict__ n1, const float* __restrict__ b0,
    const float* __restrict__ b1, const float* __restrict__ mean32, const float* __restrict__ s0,
    const float* __restrict__ s1, float* __restrict__ d0, float* __restrict__ d1, Dims d, int Wh, int gx, int gy,
    int zc, float omega, float omw, unsigned* __restrict__ gate, int it, unsigned thresh, int* __restrict__ done,
    unsigned long long* __restrict__ host_done) {
    auto word = [&](int i) {
        unsigned m = 0;
#pragma unroll
        for (int l = 0; l < GATE_LANES; l++) m = max(m, gate[i * GATE_LANES + l]);
        return m;
    };
    if (it > 0) {
        const unsigned g1 = word(it - 1);
        if (g1 <= thresh || (it > 2 * GATE_K && g1 >= word(it - 1 - GATE_K))) return;
    }
    const int per = gridDim.x >> 3;
    const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int bx = tile % gx;
    const int rest = tile / gx;
    const int by = rest % gy;
    const int z0 = (rest / gy) * zc;
    if (z0 >= d.D) return;  // spare block (the grid is a multiple of 8): uniform exit before any barrier
    unsigned dm = 0;
    sor_split_rb_tile<SPLIT_TX, TY, NW, true, ZS>(n0, n1, b0, b1, *mean32, s0, s1, d0, d1, d, Wh, bx * SPLIT_TX,
                                                  by * TY, z0, min(z0 + zc, d.D), omega, omw, &dm);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dm = max(dm, (unsigned)__shfl_xor((int)dm, o));
    __shared__ unsigned wave_max[NW];
    if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = dm;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned m = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) m = max(m, wave_max[w]);
        atomicMax(&gate[it * GATE_LANES + (blockIdx.x & (GATE_LANES - 1))], m);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int v = *done + 1;
        *done = v;
        *host_done = (unsigned long long)v;
    }
}

// Starting pressure extrapolated in time (zq3d_config.warm_start 2): p = p + beta (p - prev) and
// prev = the old p, over the padded array (ghosts stay 0); with valid == 0 only prev = p.
__global__ __launch_bounds__(256) void k_extrapolate_p64(double* __restrict__ p, double* __restrict__ prev, size_t n,
                                                         int valid, double beta) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double t = p[i];
        if (valid) p[i] = t + beta * (t - prev[i]);
        prev[i] = t;
    }
}

// Zeroes the gate words of a pass and its iteration count.
__global__ __launch_bounds__(256) void k_gate_reset(unsigned* __restrict__ gate, int n, int* __restrict__ done) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) gate[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) *done = 0;
}

// p64 += e after a gated pass: e is in the pair the last executed iteration wrote (*done iterations
// from pair p: odd -> alt).  e is not cleared (the next pass's first iteration does not read it).
__global__ __launch_bounds__(BX* BY) void k_update_p64_gated(const uint8_t* __restrict__ nc,
                                                             const float* __restrict__ p0, const float* __restrict__ p1,
                                                             const float* __restrict__ a0, const float* __restrict__ a1,
                                                             const int* __restrict__ done, double* __restrict__ p,
                                                             Dims d, int Wh) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    if (!nc[cidx(d, i, j, k)]) return;
    const bool odd = (*done) & 1;
    const bool c1 = (i + j + k) & 1;
    const float* e = odd ? (c1 ? a1 : a0) : (c1 ? p1 : p0);
    const int q = pidx(d, i, j, k);
    p[q] = p[q] + (double)e[split_q(Wh, d.H, i >> 1, j, k)];
}

// Two red+black iterations per launch on split storage (sor_split_rb2_tile); grid as k_sor_split_rb.
template <int TY, int NW>
__global__ __launch_bounds__(64 * NW) void k_sor_split_rb2(const uint8_t* __restrict__ n0, const uint8_t* __restrict__ n1,
                                                           const float* __restrict__ b0, const float* __restrict__ b1,
                                                           const float* __restrict__ mean32,
                                                           const float* __restrict__ s0, const float* __restrict__ s1,
                                                           float* __restrict__ d0, float* __restrict__ d1, Dims d,
                                                           int Wh, int gx, int gy, int zc, float omega, float omw) {
    const int per = gridDim.x >> 3;
    const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int bx = tile % gx;
    const int rest = tile / gx;
    const int by = rest % gy;
    const int z0 = (rest / gy) * zc;
    if (z0 >= d.D) return;  // spare block (the grid is a multiple of 8): uniform exit before any barrier
    sor_split_rb2_tile<SPLIT_TX, TY, NW>(n0, n1, b0, b1, *mean32, s0, s1, d0, d1, d, Wh, bx * SPLIT_TX, by * TY, z0,
                                         min(z0 + zc, d.D), omega, omw);
}

// Face correction u -= p_r - p_l (ZQ_SmokeSimulationUtils3D.cpp:2815-2861 open form, :2888-2922
// closed form); P = float (plain) or double (mixed precision; the face is rounded once).
// All three axes in one pass over the cells: cell (i, j, k) corrects its
// low u, v and w faces, and the cells of the last column / row / plane also the domain's high face.
// Every face is written once, with st_correct's value, so the result equals the three per-axis
// launches bit for bit while each pressure value is loaded once instead of three times.
template <class P, class F>
__global__ __launch_bounds__(BX* BY) void k_correct_cells(const uint8_t* __restrict__ occ, int open_mask,
                                                          const P* __restrict__ p, F* __restrict__ u,
                                                          F* __restrict__ v, F* __restrict__ w, Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    if (i >= d.W || j >= d.H) return;
    // the column marches up its chunk: the w face below plane k reads plane k-1's pressure and
    // occupancy, which this thread loaded for its u face one iteration earlier (a cache hit, not a
    // second trip to HBM as when neighbouring planes belong to other workgroups)
    const int k1 = min((int)(blockIdx.z + 1) * RES_RZ, d.D);
    for (int k = blockIdx.z * RES_RZ; k < k1; k++) {
        st_correct<0, P, F>(occ, open_mask, p, u, d, i, j, k);
        if (i == d.W - 1) st_correct<0, P, F>(occ, open_mask, p, u, d, i + 1, j, k);
        st_correct<1, P, F>(occ, open_mask, p, v, d, i, j, k);
        if (j == d.H - 1) st_correct<1, P, F>(occ, open_mask, p, v, d, i, j + 1, k);
        st_correct<2, P, F>(occ, open_mask, p, w, d, i, j, k);
        if (k == d.D - 1) st_correct<2, P, F>(occ, open_mask, p, w, d, i, j, k + 1);
    }
}

// Transposed correction: every face, solid and ghost pressures being 0 (the adjoint of the
// projection with respect to the faces the forward correction leaves alone is not zero: those
// faces enter the divergence).
template <int AXIS, class F>
__global__ __launch_bounds__(BX* BY) void k_correct_all(const double* __restrict__ p, F* __restrict__ face, Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    const int pstep = AXIS == 0 ? 1 : (AXIS == 1 ? d.W + 2 : (d.W + 2) * (d.H + 2));
    const int qh = pidx(d, i, j, k);
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    face[fi] = (F)((double)face[fi] - (p[qh] - p[qh - pstep]));
}

// Mixed-precision set-up on faces of type F: neighbour counts, FP64 divergence (FREE: only the
// faces between two fluid cells and open domain faces, i.e. those whose neighbour is counted),
// p64 = 0 on uncounted cells, per-block sum/count of the divergence over fluid cells.
template <class F, bool FREE>
__global__ __launch_bounds__(BX* BY) void k_proj_setup_mp(const F* __restrict__ u, const F* __restrict__ v,
                                                          const F* __restrict__ w, const uint8_t* __restrict__ occ,
                                                          int open_mask, uint8_t* __restrict__ nc,
                                                          double* __restrict__ b64, double* __restrict__ p64,
                                                          double* __restrict__ part, Dims d) {
    __shared__ double red[4];
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double dsum = 0.0, dcnt = 0.0;
    if (i < d.W && j < d.H) {
        const int o = cidx(d, i, j, k);
        const int HW = d.W * d.H;
        const bool fluid = !occ[o];
        bool c[6] = {false, false, false, false, false, false};
        if (fluid) {
            c[0] = (i > 0) ? !occ[o - 1] : ((open_mask & 1) != 0);
            c[1] = (i < d.W - 1) ? !occ[o + 1] : ((open_mask & 2) != 0);
            c[2] = (j > 0) ? !occ[o - d.W] : ((open_mask & 4) != 0);
            c[3] = (j < d.H - 1) ? !occ[o + d.W] : ((open_mask & 8) != 0);
            c[4] = (k > 0) ? !occ[o - HW] : ((open_mask & 16) != 0);
            c[5] = (k < d.D - 1) ? !occ[o + HW] : ((open_mask & 32) != 0);
        }
        const int n = (int)c[0] + (int)c[1] + (int)c[2] + (int)c[3] + (int)c[4] + (int)c[5];
        nc[o] = (uint8_t)n;
        double u0 = (double)u[uidx(d, i, j, k)], u1 = (double)u[uidx(d, i + 1, j, k)];
        double v0 = (double)v[vidx(d, i, j, k)], v1 = (double)v[vidx(d, i, j + 1, k)];
        double w0 = (double)w[widx(d, i, j, k)], w1 = (double)w[widx(d, i, j, k + 1)];
        if (FREE) {
            u0 = c[0] ? u0 : 0.0;
            u1 = c[1] ? u1 : 0.0;
            v0 = c[2] ? v0 : 0.0;
            v1 = c[3] ? v1 : 0.0;
            w0 = c[4] ? w0 : 0.0;
            w1 = c[5] ? w1 : 0.0;
        }
        const double dd = ((u1 - u0) + (v1 - v0)) + (w1 - w0);
        b64[o] = n ? dd : 0.0;
        if (!n) p64[pidx(d, i, j, k)] = 0.0;
        if (fluid) {
            dsum = dd;
            dcnt = 1.0;
        }
    }
    const double sm = block_reduce<false>(dsum, red);
    __syncthreads();
    const double ct = block_reduce<false>(dcnt, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[2 * block_linear()] = sm;
        part[2 * block_linear() + 1] = ct;
    }
}

// per-block max |b64| over counted cells
__global__ __launch_bounds__(BX* BY) void k_max_abs64(const uint8_t* __restrict__ nc, const double* __restrict__ b64,
                                                      double* __restrict__ part, Dims d) {
    __shared__ double red[4];
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double m = 0.0;
    if (i < d.W && j < d.H) {
        const int o = cidx(d, i, j, k);
        if (nc[o]) m = fabs(b64[o]);
    }
    m = block_reduce<true>(m, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = m;
}

// per-block sum/count of p64 over fluid cells
__global__ __launch_bounds__(BX* BY) void k_sum_p64(const uint8_t* __restrict__ occ, const double* __restrict__ p,
                                                    double* __restrict__ part, Dims d) {
    __shared__ double red[4];
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    double sm = 0.0, ct = 0.0;
    if (i < d.W && j < d.H && !occ[cidx(d, i, j, k)]) {
        sm = p[pidx(d, i, j, k)];
        ct = 1.0;
    }
    sm = block_reduce<false>(sm, red);
    __syncthreads();
    ct = block_reduce<false>(ct, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[2 * block_linear()] = sm;
        part[2 * block_linear() + 1] = ct;
    }
}

__global__ __launch_bounds__(BX* BY) void k_sub_p64(const uint8_t* __restrict__ occ, double* __restrict__ p,
                                                    const double* __restrict__ mean64, Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H || occ[cidx(d, i, j, k)]) return;
    const int q = pidx(d, i, j, k);
    p[q] = p[q] - *mean64;
}

// FP64 residual r = (b64 - mean) - (sum_nb p - n p); r32 feeds the FP32 inner solve; per-block max|r|.
// SPLIT: r32 goes to the colour-split right-hand sides r0/r1 (slot x >> 1 of row (y, z) of colour
// (x+y+z)&1) instead of the interleaved array r0; the arithmetic and the partials are the same.
template <bool SPLIT>
__global__ __launch_bounds__(BX* BY) void k_residual64(const uint8_t* __restrict__ nc, const double* __restrict__ b64,
                                                       const double* __restrict__ mean64,
                                                       const double* __restrict__ p, float* __restrict__ r0,
                                                       float* __restrict__ r1, double* __restrict__ part, Dims d,
                                                       int Wh) {
    __shared__ double red[4];
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
        if (SPLIT) (((i + j + k) & 1) ? r1 : r0)[(k * d.H + j) * Wh + (i >> 1)] = r;
        else r0[o] = r;
    }
    m = block_reduce<true>(m, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = m;
}

// k_residual64<true> marching z: each thread owns one (i, j) column of a chunk of RES_RZ planes and
// keeps the pressures of planes k-1, k, k+1 in registers, so each padded FP64 pressure is read once
// (the per-plane kernel reads every plane for three neighbouring planes' blocks).  The same operations
// per cell in the same order; the block maxima are over the chunk (a maximum is order-free).
__global__ __launch_bounds__(BX* BY) void k_residual64_z(const uint8_t* __restrict__ nc, const double* __restrict__ b64,
                                                         const double* __restrict__ mean64,
                                                         const double* __restrict__ p, float* __restrict__ r0,
                                                         float* __restrict__ r1, double* __restrict__ part, Dims d,
                                                         int Wh) {
    __shared__ double red[4];
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * RES_RZ, k1 = min(k0 + RES_RZ, d.D);
    double m = 0.0;
    if (i < d.W && j < d.H) {
        const int sy = d.W + 2, sz = (d.W + 2) * (d.H + 2);
        const double mean = *mean64;
        const int HW = d.W * d.H;
        int q = pidx(d, i, j, k0), o = cidx(d, i, j, k0);
        double pm = p[q - sz], pc = p[q];
        // this plane's loads, then each step issues the next plane's before computing: every load is
        // unconditional (padded pressure, all cells of b64 and nc) so it can be in flight early
        double xl = p[q - 1], xr = p[q + 1], yl = p[q - sy], yr = p[q + sy], pp = p[q + sz], bv = b64[o];
        int n = nc[o];
        for (int k = k0; k < k1; k++) {
            const bool more = k + 1 < k1;
            const int qn = more ? q + sz : q, on = more ? o + HW : o;
            const double xl2 = p[qn - 1], xr2 = p[qn + 1], yl2 = p[qn - sy], yr2 = p[qn + sy], pp2 = p[qn + sz];
            const double bv2 = b64[on];
            const int n2 = nc[on];
            const double sum = xl + xr + yl + yr + pm + pp;
            const double bb = bv - mean;
            const double rr = bb - (sum - (double)n * pc);
            float r = 0.0f;
            if (n) {
                r = (float)rr;
                m = fmax(m, fabs(rr));
            }
            (((i + j + k) & 1) ? r1 : r0)[(k * d.H + j) * Wh + (i >> 1)] = r;
            pm = pc;
            pc = pp;
            xl = xl2, xr = xr2, yl = yl2, yr = yr2, pp = pp2, bv = bv2, n = n2;
            q = qn, o = on;
        }
    }
    m = block_reduce<true>(m, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = m;
}

// p64 += e, and with ZERO e = 0, on counted cells; SPLIT: e lives in the padded colour-split arrays
// e0/e1.
template <bool SPLIT, bool ZERO = true>
__global__ __launch_bounds__(BX* BY) void k_update_p64(const uint8_t* __restrict__ nc, float* __restrict__ e0,
                                                       float* __restrict__ e1, double* __restrict__ p, Dims d,
                                                       int Wh) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const int q = pidx(d, i, j, k);
    float* ep = SPLIT ? &(((i + j + k) & 1) ? e1 : e0)[split_q(Wh, d.H, i >> 1, j, k)] : &e0[q];
    // the three loads together (the count no longer gates the other two), one round trip
    const int n = nc[cidx(d, i, j, k)];
    const double pv = p[q];
    const float ev = *ep;
    if (!n) return;
    p[q] = pv + (double)ev;
    if (ZERO) *ep = 0.0f;
}

template <class F>
__global__ __launch_bounds__(BX* BY) void k_max_div64(const F* __restrict__ u, const F* __restrict__ v,
                                                      const F* __restrict__ w, const uint8_t* __restrict__ nc,
                                                      const double* __restrict__ mean64, double* __restrict__ part,
                                                      Dims d) {
    __shared__ double red[4];
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
    m = block_reduce<true>(m, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = m;
}

// max |div - mean| over counted cells of the field the FP64 correction WOULD produce: each of a
// cell's six faces is evaluated as st_correct stores it (rounded once to F), then summed in FP64
// as k_max_div64 sums stored faces, so the value equals k_max_div64 after launch_correct64 bit for
// bit.  Reads the faces and the pressure only; nothing is written but the per-block maxima.
// A face after the correction: f - (p_high - p_low) rounded once to F where the correction applies
// (ok), else f as it is; the expression of st_corrected_face.
template <class F>
__device__ __forceinline__ F pred_face(F f, bool ok, double ph, double pl) {
    return ok ? (F)((double)f - (ph - pl)) : f;
}

// One (i, j) column per thread up a chunk of RES_RZ planes.  Per plane a thread evaluates three
// faces of its cell, the low u and v faces and the high w face; the high u face comes from the
// next lane (the lane at a block's or the grid's edge evaluates it), the high v face from the
// next row through LDS (likewise), the low w face is the plane below's high w face.  Every face
// is evaluated once, and six times fewer loads are issued than with a cell per thread evaluating
// all six.
template <class F>
__global__ __launch_bounds__(BX* BY) void k_max_div64_pred(const F* __restrict__ u, const F* __restrict__ v,
                                                           const F* __restrict__ w, const uint8_t* __restrict__ occ,
                                                           int open_mask, const uint8_t* __restrict__ nc,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ mean64,
                                                           double* __restrict__ part, Dims d) {
    static_assert(BX == 64, "a block row is one wave: the high u face comes from the next lane");
    __shared__ double red[4];
    __shared__ F vrow[BY][BX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = blockIdx.x * BX + tx;
    const int j = blockIdx.y * BY + ty;
    const int k0 = blockIdx.z * RES_RZ, k1 = min(k0 + RES_RZ, d.D);
    const bool in = i < d.W && j < d.H;
    const int sy = d.W + 2, sz = (d.W + 2) * (d.H + 2), W = d.W, HW = d.W * d.H;
    const double mean = *mean64;
    const bool right_here = tx == BX - 1 || i == d.W - 1;  // this lane evaluates its high u face
    const bool top_here = ty == BY - 1 || j == d.H - 1;    // this row evaluates its high v face
    double m = 0.0;
    // the centre cell and its low w face, carried up the column
    int q = pidx(d, min(i, d.W - 1), min(j, d.H - 1), k0);
    int o = cidx(d, min(i, d.W - 1), min(j, d.H - 1), k0);
    double pc = 0.0;
    int oc = 1;
    F wb = (F)0;
    if (in) {
        pc = p[q];
        oc = occ[o];
        const F f = w[widx(d, i, j, k0)];
        const bool ok = k0 == 0 ? ((open_mask & 16) && !oc) : (!occ[o - HW] && !oc);
        wb = pred_face(f, ok, pc, p[q - sz]);
    }
    for (int k = k0; k < k1; k++, q += sz, o += HW) {
        F ul = (F)0, vb = (F)0, wt = (F)0, ur = (F)0, vt = (F)0;
        double pu = 0.0;
        int ou = 1, n = 0;
        if (in) {
            n = nc[o];
            pu = p[q + sz];
            ou = k + 1 < d.D ? occ[o + HW] : 1;
            {  // low u face
                const bool ok = i == 0 ? ((open_mask & 1) && !oc) : (!occ[o - 1] && !oc);
                ul = pred_face(u[uidx(d, i, j, k)], ok, pc, p[q - 1]);
            }
            {  // low v face
                const bool ok = j == 0 ? ((open_mask & 4) && !oc) : (!occ[o - W] && !oc);
                vb = pred_face(v[vidx(d, i, j, k)], ok, pc, p[q - sy]);
            }
            {  // high w face
                const bool ok = k + 1 == d.D ? ((open_mask & 32) && !oc) : (!oc && !ou);
                wt = pred_face(w[widx(d, i, j, k + 1)], ok, pu, pc);
            }
            if (right_here) {
                const bool ok = i + 1 == d.W ? ((open_mask & 2) && !oc) : (!occ[o + 1] && !oc);
                ur = pred_face(u[uidx(d, i + 1, j, k)], ok, p[q + 1], pc);
            }
            if (top_here) {
                const bool ok = j + 1 == d.H ? ((open_mask & 8) && !oc) : (!occ[o + W] && !oc);
                vt = pred_face(v[vidx(d, i, j + 1, k)], ok, p[q + sy], pc);
            }
        }
        const F ul_next = __shfl_down(ul, 1, BX);
        vrow[ty][tx] = vb;
        __syncthreads();
        if (!right_here) ur = ul_next;
        if (!top_here) vt = vrow[ty + 1][tx];
        __syncthreads();  // vrow is rewritten at the next plane
        if (in && n) {
            const double du = (double)ur - (double)ul;
            const double dv = (double)vt - (double)vb;
            const double dw = (double)wt - (double)wb;
            m = fmax(m, fabs(((du + dv) + dw) - mean));
        }
        wb = wt;
        pc = pu;
        oc = ou;
    }
    m = block_reduce<true>(m, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = m;
}

int cell_blocks(const Dims& d) {
    const dim3 g = grid_for(d.W, d.H, d.D);
    return (int)(g.x * g.y * g.z);
}

// Fused-kernel launch shape: TY rows per block, ZC planes per chunk; nblocks 0 when it cannot run.
struct FusedShape {
    int ty, zc, ntiles_y, nchunks, nblocks;
    size_t lds;
};

bool shape_supported(int ty) { return ty == 2 || ty == 4 || ty == 8; }

FusedShape fused_shape(const Dims& d, SorShape want) {
    FusedShape f{};
    f.ty = want.ty ? want.ty : 4;
    f.zc = want.zc ? want.zc : 16;
    if (!shape_supported(f.ty) || f.zc < 1) return FusedShape{};
    f.lds = (size_t)4 * (f.ty + 4) * (d.W + 2) * sizeof(float);
    if (f.lds > 64 * 1024 || d.W > 512) return FusedShape{};
    f.ntiles_y = (d.H + f.ty - 1) / f.ty;
    f.nchunks = (d.D + f.zc - 1) / f.zc;
    const int tiles = f.ntiles_y * f.nchunks;
    f.nblocks = (tiles + 7) & ~7;  // multiple of 8 for the XCD numbering; spare blocks exit
    return f;
}

template <int TY>
void launch_fused(const FusedShape& f, const Dims& d, const uint8_t* nc, const float* b, const float* mean32,
                  const float* src, float* dst, float omega, float omw, hipStream_t s) {
    if (d.W <= 256)
        hipLaunchKernelGGL((k_sor_fused<TY, 1>), dim3(f.nblocks), dim3(256), f.lds, s, nc, b, mean32, src, dst, d, omega,
                           omw, f.zc, f.ntiles_y);
    else
        hipLaunchKernelGGL((k_sor_fused<TY, 2>), dim3(f.nblocks), dim3(256), f.lds, s, nc, b, mean32, src, dst, d, omega,
                           omw, f.zc, f.ntiles_y);
}

}  // namespace

int reduction_partials(const Dims& d) { return 2 * cell_blocks(d); }

void launch_proj_setup(const Dims& d, const float* u, const float* v, const float* w, const uint8_t* occ,
                       int open_mask, uint8_t* nc, float* b, double* b64, float* p_pad, double* p64_pad,
                       double* part, hipStream_t s) {
    if (b64)
        hipLaunchKernelGGL(k_proj_setup_z, grid_z(d), dim3(BX, BY), 0, s, u, v, w, occ, open_mask, nc, b64, p64_pad,
                           part, d);
    else
        hipLaunchKernelGGL(k_proj_setup, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, u, v, w, occ, open_mask, nc, b,
                           p_pad, part, d);
}

void launch_mean_finalize(const Dims& d, const double* part, float* mean32, double* mean64, int want, hipStream_t s) {
    hipLaunchKernelGGL(k_finish_mean, dim3(1), dim3(1024), 0, s, part, cell_blocks(d), mean32, mean64, want);
}

void launch_mean_finalize_mp(const Dims& d, const double* part, double* mean64, int want, hipStream_t s) {
    const dim3 g = grid_z(d);
    hipLaunchKernelGGL(k_finish_mean, dim3(1), dim3(1024), 0, s, part, (int)(g.x * g.y * g.z), nullptr, mean64, want);
}

size_t split_padded_size(const Dims& d) { return (size_t)split_row((d.W + 1) / 2) * (d.H + 2) * (d.D + 2); }
size_t split_cell_size(const Dims& d) { return (size_t)((d.W + 1) / 2) * d.H * d.D; }

template <int TY, int NW, bool TWO, bool ZS = false>
void launch_split_rb(const Dims& d, const SplitBufs& sp, const float* const* src, float* const* dst,
                     const float* mean32, float omega, float omw, int zc, hipStream_t s) {
    const int Wh = (d.W + 1) / 2;
    const int gx = (Wh + SPLIT_TX - 1) / SPLIT_TX, gy = (d.H + TY - 1) / TY, gz = (d.D + zc - 1) / zc;
    const int nb = (gx * gy * gz + 7) & ~7;
    if (TWO)
        hipLaunchKernelGGL((k_sor_split_rb2<TY, NW>), dim3(nb), dim3(64 * NW), 0, s, sp.n[0], sp.n[1], sp.b[0],
                           sp.b[1], mean32, src[0], src[1], dst[0], dst[1], d, Wh, gx, gy, zc, omega, omw);
    else
        hipLaunchKernelGGL((k_sor_split_rb<TY, NW, ZS>), dim3(nb), dim3(64 * NW), 0, s, sp.n[0], sp.n[1], sp.b[0],
                           sp.b[1], mean32, src[0], src[1], dst[0], dst[1], d, Wh, gx, gy, zc, omega, omw);
}

// split shape ty = -r (one iteration per launch) or -(SOR_TWO + r) (two iterations per launch)
inline int split_rows(SorShape sh) { return sh.ty <= -SOR_TWO ? -sh.ty - SOR_TWO : -sh.ty; }

int launch_split_sweeps(const Dims& d, const SplitBufs& sp, int from, const float* mean32, float omega, int iters,
                        SorShape shape, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, bool zero_start) {
    const float omw = 1.0f - omega;
    const int Wh = (d.W + 1) / 2;
    if (ev0) (void)hipEventRecord(ev0, s);
    if (shape.ty == SOR_SPLIT) {
        float* const* p = from ? sp.alt : sp.p;
        const int gx = (Wh + BX - 1) / BX, gy = (d.H + BY - 1) / BY;
        const int nb = (gx * gy * d.D + 7) & ~7;
        for (int it = 0; it < iters; it++)
            for (int c = 0; c < 2; c++)
                hipLaunchKernelGGL(k_sor_split, dim3(nb), dim3(BX * BY), 0, s, sp.n[c], sp.b[c], mean32, p[c],
                                   p[1 - c], d, Wh, gx, gy, omega, omw, c);
    } else {
        const int ty = split_rows(shape), zc = shape.zc;
        const bool two = shape.ty <= -SOR_TWO;
        for (int it = 0; it < iters;) {
            float* const* src = from ? sp.alt : sp.p;
            float* const* dst = from ? sp.p : sp.alt;
            if (two && it + 2 <= iters) {
                if (ty == 16) launch_split_rb<16, 8, true>(d, sp, src, dst, mean32, omega, omw, zc, s);
                else launch_split_rb<32, 16, true>(d, sp, src, dst, mean32, omega, omw, zc, s);
                it += 2;
            } else if (zero_start && it == 0 && split_zero_start(shape)) {
                if (ty == 4) launch_split_rb<4, 4, false, true>(d, sp, src, dst, mean32, omega, omw, zc, s);
                else if (ty == 8) launch_split_rb<8, 8, false, true>(d, sp, src, dst, mean32, omega, omw, zc, s);
                else if (ty == 16) launch_split_rb<16, 8, false, true>(d, sp, src, dst, mean32, omega, omw, zc, s);
                else launch_split_rb<32, 16, false, true>(d, sp, src, dst, mean32, omega, omw, zc, s);
                it += 1;
            } else {
                if (ty == 4) launch_split_rb<4, 4, false>(d, sp, src, dst, mean32, omega, omw, zc, s);
                else if (ty == 8) launch_split_rb<8, 8, false>(d, sp, src, dst, mean32, omega, omw, zc, s);
                else if (ty == 16) launch_split_rb<16, 8, false>(d, sp, src, dst, mean32, omega, omw, zc, s);
                else launch_split_rb<32, 16, false>(d, sp, src, dst, mean32, omega, omw, zc, s);
                it += 1;
            }
            from = 1 - from;
        }
    }
    if (ev1) (void)hipEventRecord(ev1, s);
    return from;
}

template <int TY, int NW>
void launch_gated_one(const Dims& d, const SplitBufs& sp, int it, const float* mean32, float omega, int zc,
                      const GateBufs& g, unsigned thresh, hipStream_t s) {
    const int Wh = (d.W + 1) / 2;
    const int gx = (Wh + SPLIT_TX - 1) / SPLIT_TX, gy = (d.H + TY - 1) / TY, gz = (d.D + zc - 1) / zc;
    const int nb = (gx * gy * gz + 7) & ~7;
    const float omw = 1.0f - omega;
    // iteration it reads pair (it & 1: alt) and writes the other
    const float* const* src = (it & 1) ? sp.alt : sp.p;
    float* const* dst = (it & 1) ? sp.p : sp.alt;
    if (it == 0)
        hipLaunchKernelGGL((k_sor_split_rb_gated<TY, NW, true>), dim3(nb), dim3(64 * NW), 0, s, sp.n[0], sp.n[1],
                           sp.b[0], sp.b[1], mean32, src[0], src[1], dst[0], dst[1], d, Wh, gx, gy, zc, omega, omw,
                           g.gate, it, thresh, g.done, g.host_done);
    else
        hipLaunchKernelGGL((k_sor_split_rb_gated<TY, NW, false>), dim3(nb), dim3(64 * NW), 0, s, sp.n[0], sp.n[1],
                           sp.b[0], sp.b[1], mean32, src[0], src[1], dst[0], dst[1], d, Wh, gx, gy, zc, omega, omw,
                           g.gate, it, thresh, g.done, g.host_done);
}

void launch_extrapolate_p64(double* p64_pad, double* prev_pad, size_t n, int valid, float beta, hipStream_t s) {
    hipLaunchKernelGGL(k_extrapolate_p64, dim3((unsigned)std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                       p64_pad, prev_pad, n, valid, (double)beta);
}

void launch_gate_reset(const GateBufs& g, int n, hipStream_t s) {
    const int words = n * GATE_LANES;
    hipLaunchKernelGGL(k_gate_reset, dim3(std::min(64, (words + 255) / 256 + 1)), dim3(256), 0, s, g.gate, words,
                       g.done);
}

void launch_gated_sweeps(const Dims& d, const SplitBufs& sp, int it0, int count, const float* mean32, float omega,
                         SorShape shape, const GateBufs& g, unsigned thresh, hipStream_t s) {
    // one iteration per launch; the one-launch shapes use their rows and planes, others 32 x zc 16
    const bool one = is_split(shape) && shape.ty != SOR_SPLIT && shape.ty > -SOR_TWO;
    const int ty = one ? split_rows(shape) : 32;
    const int zc = one && shape.zc > 0 ? shape.zc : 16;
    for (int it = it0; it < it0 + count; it++) {
        if (ty == 4) launch_gated_one<4, 4>(d, sp, it, mean32, omega, zc, g, thresh, s);
        else if (ty == 8) launch_gated_one<8, 8>(d, sp, it, mean32, omega, zc, g, thresh, s);
        else if (ty == 16) launch_gated_one<16, 8>(d, sp, it, mean32, omega, zc, g, thresh, s);
        else launch_gated_one<32, 16>(d, sp, it, mean32, omega, zc, g, thresh, s);
    }
}

void launch_update_p64_gated(const Dims& d, const uint8_t* nc, const SplitBufs& sp, const int* done, double* p64_pad,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_update_p64_gated, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, nc, sp.p[0], sp.p[1],
                       sp.alt[0], sp.alt[1], done, p64_pad, d, (d.W + 1) / 2);
}

void launch_split_counts(const Dims& d, const uint8_t* nc, const SplitBufs& sp, hipStream_t s) {
    const int Wh = (d.W + 1) / 2;
    hipLaunchKernelGGL(k_split_counts, dim3((Wh + BX - 1) / BX, (d.H + BY - 1) / BY, 2 * d.D), dim3(BX, BY), 0, s,
                       nc, sp.n[0], sp.n[1], d, Wh);
}

float* launch_sor(const Dims& d, const uint8_t* nc, const float* b, const float* mean32, float* p_pad, float* p_alt,
                  float omega, int iters, hipStream_t s, SorShape shape, const SplitBufs* sp, hipEvent_t ev0,
                  hipEvent_t ev1) {
    const float omw = 1.0f - omega;
    if (is_split(shape) && sp) {
        const int Wh = (d.W + 1) / 2;
        hipLaunchKernelGGL(k_split_in, dim3((Wh + BX - 1) / BX, (d.H + BY - 1) / BY, 2 * d.D), dim3(BX, BY), 0, s, nc,
                           b, p_pad, sp->n[0], sp->n[1], sp->b[0], sp->b[1], sp->p[0], sp->p[1], d, Wh);
        const int res = launch_split_sweeps(d, *sp, 0, mean32, omega, iters, shape, s, ev0, ev1);
        float* const* r = res ? sp->alt : sp->p;
        hipLaunchKernelGGL(k_split_out, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, r[0], r[1], p_pad, d, Wh);
        return p_pad;
    }
    if (ev0) (void)hipEventRecord(ev0, s);
    const FusedShape f = fused_shape(d, shape);
    if (f.nblocks > 0 && p_alt) {
        float* src = p_pad;
        float* dst = p_alt;
        for (int it = 0; it < iters; it++) {
            if (f.ty == 2) launch_fused<2>(f, d, nc, b, mean32, src, dst, omega, omw, s);
            else if (f.ty == 8) launch_fused<8>(f, d, nc, b, mean32, src, dst, omega, omw, s);
            else launch_fused<4>(f, d, nc, b, mean32, src, dst, omega, omw, s);
            float* t = src;
            src = dst;
            dst = t;
        }
        if (ev1) (void)hipEventRecord(ev1, s);
        return src;
    }
    const dim3 g = grid_for((d.W + 1) / 2, d.H, d.D);
    for (int it = 0; it < iters; it++) {
        hipLaunchKernelGGL(k_sor_half, g, dim3(BX, BY), 0, s, nc, b, mean32, p_pad, d, omega, omw, 0);
        hipLaunchKernelGGL(k_sor_half, g, dim3(BX, BY), 0, s, nc, b, mean32, p_pad, d, omega, omw, 1);
    }
    if (ev1) (void)hipEventRecord(ev1, s);
    return p_pad;
}

const char* sor_kernel_name(const Dims& d, SorShape shape) {
    if (shape.ty == SOR_SPLIT) return "k_sor_split";
    if (shape.ty <= -SOR_TWO) return "k_sor_split_rb2";
    if (is_split(shape)) return "k_sor_split_rb";
    return fused_shape(d, shape).nblocks > 0 ? "k_sor_fused" : "k_sor_half";
}

long long sor_launches(const Dims& d, SorShape shape, int iters) {
    if (shape.ty == SOR_SPLIT) return 2LL * iters;
    if (shape.ty <= -SOR_TWO) return (iters + 1) / 2;
    if (is_split(shape)) return iters;
    return (fused_shape(d, shape).nblocks > 0 ? 1LL : 2LL) * iters;
}

bool split_buffers_fit(const Dims& d) {
    return split_padded_size(d) * 4 <= 0x7FFFFFFFull && split_cell_size(d) * 4 <= 0x7FFFFFFFull;
}

int sor_shape_candidates(const Dims& d, SorShape* out, int max) {
    static const int tys[] = {2, 4, 8};
    static const int zcs[] = {8, 16, 32, 64};
    int n = 0;
    for (int ty : tys)
        for (int zc : zcs) {
            const SorShape sh{ty, zc};
            if (fused_shape(d, sh).nblocks > 0 && n < max) out[n++] = sh;
        }
    if (n < max) out[n++] = SorShape{SOR_SPLIT, 0};
    static const int rb_rows[] = {4, 8, 16, 32};
    static const int rb_zc[] = {8, 16, 32};
    if (split_buffers_fit(d))
        for (int ty : rb_rows)
            for (int zc : rb_zc)
                if (n < max) out[n++] = SorShape{-ty, zc};
    static const int rb2_rows[] = {16, 32};
    static const int rb2_zc[] = {8, 16, 32};
    for (int ty : rb2_rows)
        for (int zc : rb2_zc)
            if (n < max) out[n++] = SorShape{-(SOR_TWO + ty), zc};
    return n;
}

void launch_correct(const Dims& d, const uint8_t* occ, int open_mask, const float* p_pad, float* u, float* v,
                    float* w, hipStream_t s) {
    hipLaunchKernelGGL((k_correct_cells<float, float>), grid_z(d), dim3(BX, BY), 0, s, occ, open_mask, p_pad, u, v,
                       w, d);
}

template <class F>
void launch_correct64(const Dims& d, const uint8_t* occ, int open_mask, const double* p64_pad, F* u, F* v, F* w,
                      hipStream_t s) {
    hipLaunchKernelGGL((k_correct_cells<double, F>), grid_z(d), dim3(BX, BY), 0, s, occ, open_mask, p64_pad, u, v,
                       w, d);
}

template <class F>
void launch_correct_all64(const Dims& d, const double* p64_pad, F* u, F* v, F* w, hipStream_t s) {
    hipLaunchKernelGGL((k_correct_all<0, F>), grid_for(d.W + 1, d.H, d.D), dim3(BX, BY), 0, s, p64_pad, u, d);
    hipLaunchKernelGGL((k_correct_all<1, F>), grid_for(d.W, d.H + 1, d.D), dim3(BX, BY), 0, s, p64_pad, v, d);
    hipLaunchKernelGGL((k_correct_all<2, F>), grid_for(d.W, d.H, d.D + 1), dim3(BX, BY), 0, s, p64_pad, w, d);
}

void launch_proj_setup64(const Dims& d, const double* u, const double* v, const double* w, const uint8_t* occ,
                         int open_mask, int free_only, uint8_t* nc, double* b64, double* p64_pad, double* part,
                         hipStream_t s) {
    if (free_only)
        hipLaunchKernelGGL((k_proj_setup_mp<double, true>), grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, u, v, w, occ,
                           open_mask, nc, b64, p64_pad, part, d);
    else
        hipLaunchKernelGGL((k_proj_setup_mp<double, false>), grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, u, v, w,
                           occ, open_mask, nc, b64, p64_pad, part, d);
}

void launch_proj_setup_free(const Dims& d, const float* u, const float* v, const float* w, const uint8_t* occ,
                            int open_mask, uint8_t* nc, double* b64, double* p64_pad, double* part, hipStream_t s) {
    hipLaunchKernelGGL((k_proj_setup_mp<float, true>), grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, u, v, w, occ,
                       open_mask, nc, b64, p64_pad, part, d);
}

void launch_max_abs64(const Dims& d, const uint8_t* nc, const double* b64, double* part, unsigned long long* out,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_max_abs64, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, nc, b64, part, d);
    hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(1024), 0, s, part, cell_blocks(d), out);
}

void launch_remove_mean_p64(const Dims& d, const uint8_t* occ, double* p64_pad, double* part, double* mean64,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_sum_p64, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, occ, p64_pad, part, d);
    hipLaunchKernelGGL(k_finish_mean, dim3(1), dim3(1024), 0, s, part, cell_blocks(d), nullptr, mean64, 1);
    hipLaunchKernelGGL(k_sub_p64, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, occ, p64_pad, mean64, d);
}

void launch_residual64(const Dims& d, const uint8_t* nc, const double* b64, const double* mean64,
                       const double* p64_pad, float* r32, double* part, unsigned long long* rmax, hipStream_t s) {
    hipLaunchKernelGGL(k_residual64<false>, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, nc, b64, mean64, p64_pad,
                       r32, nullptr, part, d, 0);
    hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(1024), 0, s, part, cell_blocks(d), rmax);
}

void launch_residual64_split(const Dims& d, const uint8_t* nc, const double* b64, const double* mean64,
                             const double* p64_pad, const SplitBufs& sp, double* part, unsigned long long* rmax,
                             hipStream_t s) {
    const dim3 g = grid_z(d);
    hipLaunchKernelGGL(k_residual64_z, g, dim3(BX, BY), 0, s, nc, b64, mean64, p64_pad, sp.b[0], sp.b[1], part, d,
                       (d.W + 1) / 2);
    hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(1024), 0, s, part, (int)(g.x * g.y * g.z), rmax);
}

void launch_update_p64(const Dims& d, const uint8_t* nc, float* e_pad, double* p64_pad, hipStream_t s) {
    hipLaunchKernelGGL(k_update_p64<false>, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, nc, e_pad, nullptr, p64_pad,
                       d, 0);
}

void launch_update_p64_split(const Dims& d, const uint8_t* nc, float* const* e, double* p64_pad, hipStream_t s,
                             bool zero_e) {
    if (zero_e)
        hipLaunchKernelGGL((k_update_p64<true, true>), grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, nc, e[0], e[1],
                           p64_pad, d, (d.W + 1) / 2);
    else
        hipLaunchKernelGGL((k_update_p64<true, false>), grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, nc, e[0], e[1],
                           p64_pad, d, (d.W + 1) / 2);
}

template <class F>
void launch_max_div64(const Dims& d, const F* u, const F* v, const F* w, const uint8_t* nc, const double* mean64,
                      double* part, unsigned long long* dmax, hipStream_t s) {
    hipLaunchKernelGGL(k_max_div64<F>, grid_for(d.W, d.H, d.D), dim3(BX, BY), 0, s, u, v, w, nc, mean64, part, d);
    hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(1024), 0, s, part, cell_blocks(d), dmax);
}

template <class F>
void launch_max_div64_pred(const Dims& d, const F* u, const F* v, const F* w, const uint8_t* occ, int open_mask,
                           const uint8_t* nc, const double* p64_pad, const double* mean64, double* part,
                           unsigned long long* dmax, hipStream_t s) {
    const dim3 g = grid_z(d);
    hipLaunchKernelGGL(k_max_div64_pred<F>, g, dim3(BX, BY), 0, s, u, v, w, occ, open_mask, nc, p64_pad, mean64, part,
                       d);
    hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(1024), 0, s, part, (int)(g.x * g.y * g.z), dmax);
}

#define ZQ3D_PROJ_INSTANTIATE(F)                                                                                       \
    template void launch_correct64<F>(const Dims&, const uint8_t*, int, const double*, F*, F*, F*, hipStream_t);       \
    template void launch_correct_all64<F>(const Dims&, const double*, F*, F*, F*, hipStream_t);                        \
    template void launch_max_div64<F>(const Dims&, const F*, const F*, const F*, const uint8_t*, const double*,        \
                                      double*, unsigned long long*, hipStream_t);                                      \
    template void launch_max_div64_pred<F>(const Dims&, const F*, const F*, const F*, const uint8_t*, int,             \
                                           const uint8_t*, const double*, const double*, double*,                      \
                                           unsigned long long*, hipStream_t);
ZQ3D_PROJ_INSTANTIATE(float)
ZQ3D_PROJ_INSTANTIATE(double)
#undef ZQ3D_PROJ_INSTANTIATE

}  // namespace zq3d
