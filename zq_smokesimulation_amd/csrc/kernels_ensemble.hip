// kernels_ensemble.hip -- one launch per stage for a whole ensemble of independent smoke grids.
//
// No reference counterpart.  The members have their own sizes and parameters.  Every stage is a
// persistent launch over a packed tile table: the tiles of all members' grids (64 x 4 cells of
// one z plane, the blocks of the single-grid kernels) are numbered member after member; a
// workgroup takes tiles blockIdx.x, blockIdx.x + gridDim.x, ...; finds the member by binary search
// over the table's prefix sums; and runs the same per-cell function (zq3d_stencils.h) that the
// single-grid kernel runs for that cell.  Each member therefore steps bit for bit as a zq3d_sim
// of the same configuration (plain FP32 projection).  The red-black sweep takes its tile shape
// (TBX x TBY cells) as a template parameter chosen by the ensemble autotuner.
#include "zq3d_ensemble.h"
#include "zq3d_sor_split.h"
#include "zq3d_stencils.h"

namespace zq3d {

namespace {

constexpr int NT = 256;

// member s owning tile `tile` of a table with prefix sums start[0..B]
__device__ __forceinline__ int tile_member(const int* __restrict__ start, int B, int tile) {
    int lo = 0, hi = B - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Runs f(member, local tile, i, j, k) over the tiles of table T for thread layout TBX x TBY.
template <int TBX, int TBY, class F>
__device__ __forceinline__ void for_tiles(const EnsTable& T, int B, F&& f) {
    const int total = T.start[B];
    const int tx = threadIdx.x % TBX, ty = threadIdx.x / TBX;
    for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
        const int s = tile_member(T.start, B, tile);
        const int local = tile - T.start[s];
        const int gx = T.gx[s], gy = T.gy[s];
        const int bx = local % gx, by = (local / gx) % gy, bz = local / (gx * gy);
        f(s, local, bx * TBX + tx, by * TBY + ty, bz);
    }
}

__global__ __launch_bounds__(NT) void ke_vorticity(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        // four planes of the member's N values in its vel4 slot (the advection fills it again)
        st_vorticity_to<float>(m.u, m.v, m.w, VortPlanesOut<float>{(float*)m.vel4, m.d.W * m.d.H * m.d.D}, m.d,
                               m.two_h, i, j, k);
    });
}

__global__ __launch_bounds__(NT) void ke_force_cells(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        st_force_cell_from<float, true>(VortPlanesIn<float>{(const float*)m.vel4, m.d.W * m.d.H * m.d.D}, m.T, m.R,
                                        (const float*)nullptr, m.scratch, m.d, m.ch, m.beta, m.alpha, m.Tamb, m.aT,
                                        m.aR, i, j, k);
    });
}

template <int AXIS>
__global__ __launch_bounds__(NT) void ke_force_faces(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        float* face = AXIS == 0 ? m.u : (AXIS == 1 ? m.v : m.w);
        st_force_face<float, AXIS, true>(face, m.occ, m.scratch, m.d, m.dt, m.dth, m.aV, i, j, k);
    });
}

// MAC -> cell velocity of the current faces
__global__ __launch_bounds__(NT) void ke_mac_to_cell4(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        st_mac_to_cell4<float>(m.u, m.v, m.w, m.vel4, m.d, i, j, k, m.occ);
    });
}

__global__ __launch_bounds__(NT) void ke_advect_cells(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        st_advect_cell<float>(m.vel4, m.occ, m.scratch, m.d, StepScale<float>{m.h, m.inv_h}, m.steps, m.dtt, i, j, k);
    });
}

template <int AXIS>
__global__ __launch_bounds__(NT) void ke_cells_to_faces(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        float* face = AXIS == 0 ? m.u : (AXIS == 1 ? m.v : m.w);
        // the reference's stage always advects with is_open = true (ZQ_SmokeSimulationUtils3D.cpp:2163)
        st_cells_to_face<float, AXIS>(m.scratch, m.occ, face, m.d, 1, i, j, k);
    });
}

__global__ __launch_bounds__(NT) void ke_advect_scalar(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        const RegularVelCached<float> V{m.vel4, m.d, {}};
        st_advect_scalar<float, RegularVelCached<float>>(V, m.occ, m.T, m.R, m.T2, m.R2, m.d,
                                                         StepScale<float>{m.h, m.inv_h}, m.steps, m.dtt, i, j, k);
    });
}

__global__ __launch_bounds__(NT) void ke_reinject(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        for (int q = 0; q < m.n_box; q++) {
            const float* b = m.box[q];
            st_reinject_box<float>(m.T, m.R, m.d, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], i, j, k);
        }
        for (int q = 0; q < m.n_sph; q++) {
            const float* b = m.sph[q];
            st_reinject_sphere<float>(m.T, m.R, m.d, b[0], b[1], b[2], b[3], b[4], b[5], i, j, k);
        }
    });
}

// set-up of the plain projection; the per-tile (sum, count) of the divergence goes to the member's
// partials at the tile's index, as the single-grid set-up writes them per block
__global__ __launch_bounds__(NT) void ke_proj_setup(const EnsMember* __restrict__ M, EnsTable T, int B) {
    __shared__ double red[4];
    for_tiles<64, 4>(T, B, [&](int s, int local, int i, int j, int k) {
        const EnsMember& m = M[s];
        double dsum = 0.0, dcnt = 0.0;
        st_proj_setup_plain(m.u, m.v, m.w, m.occ, m.open_mask, m.nc, m.b, m.p, m.d, i, j, k, dsum, dcnt);
        const double sm = block_reduce<false>(dsum, red);
        __syncthreads();
        const double ct = block_reduce<false>(dcnt, red);
        if (threadIdx.x == 0) {
            m.part[2 * local] = sm;
            m.part[2 * local + 1] = ct;
        }
        __syncthreads();
    });
}

// one workgroup of 1024 threads per member
__global__ __launch_bounds__(1024) void ke_finish_mean(const EnsMember* __restrict__ M, EnsTable T) {
    __shared__ double ss[1024], sc[1024];
    const int s = blockIdx.x;
    const EnsMember& m = M[s];
    st_fold_mean(m.part, T.start[s + 1] - T.start[s], m.mean32, nullptr, m.closed_mean, ss, sc);
}

// one colour of iteration `it` of the red-black SOR; members whose own iteration count is reached
// are left alone
template <int TBX, int TBY>
__global__ __launch_bounds__(NT) void ke_sor_half(const EnsMember* __restrict__ M, EnsTable T, int B, int colour,
                                                 int it) {
    for_tiles<TBX, TBY>(T, B, [&](int s, int, int t, int j, int k) {
        const EnsMember& m = M[s];
        if (it >= m.iters) return;
        st_sor_half(m.nc, m.b, m.mean32, m.p, m.d, m.omega, m.omw, colour, t, j, k);
    });
}

// every member's counts, right-hand side and pressure -> its colour-split arrays (pressure pair 0)
__global__ __launch_bounds__(NT) void ke_split_in(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        if (i >= m.d.W || j >= m.d.H) return;
        const int Wh = (m.d.W + 1) / 2, c = (i + j + k) & 1;
        const int oh = (k * m.d.H + j) * Wh + (i >> 1);
        m.sn[c][oh] = m.nc[cidx(m.d, i, j, k)];
        m.sb[c][oh] = m.b[cidx(m.d, i, j, k)];
        m.sp[c][split_q(Wh, m.d.H, i >> 1, j, k)] = m.p[pidx(m.d, i, j, k)];
    });
}

// the pair holding every member's result (its own iteration count's parity) -> interleaved pressure
__global__ __launch_bounds__(NT) void ke_split_out(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        if (i >= m.d.W || j >= m.d.H) return;
        const int Wh = (m.d.W + 1) / 2, c = (i + j + k) & 1;
        m.p[pidx(m.d, i, j, k)] = m.sp[2 * (m.iters & 1) + c][split_q(Wh, m.d.H, i >> 1, j, k)];
    });
}

// iteration `it` of every member's SOR in one launch: one workgroup per tile of the table (members'
// tiles numbered member after member, handed out so each XCD gets a contiguous run); iteration it
// reads pair it & 1 and writes the other; members whose own iteration count is reached are left
// alone
// 8-wave workgroups are held to six waves per SIMD (80 VGPRs), so three share a CU: the LDS rings of
// three fit in 160 KB and the ensemble's tiles are too small to keep a CU busy two at a time.
template <int TX, int TY, int NW>
__global__ __launch_bounds__(64 * NW, NW == 8 ? 6 : 1) void ke_sor_split_rb(const EnsMember* __restrict__ M,
                                                                            EnsTable T, int B, int zc, int it) {
    const int per = gridDim.x >> 3;
    const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (tile >= T.start[B]) return;  // spare block: uniform exit before any barrier
    const int s = tile_member(T.start, B, tile);
    const EnsMember& m = M[s];
    if (it >= m.iters) return;
    const int local = tile - T.start[s];
    const int gx = T.gx[s], gy = T.gy[s];
    const int bx = local % gx, by = (local / gx) % gy, z0 = (local / (gx * gy)) * zc;
    const int src = it & 1;
    sor_split_rb_tile<TX, TY, NW>(m.sn[0], m.sn[1], m.sb[0], m.sb[1], *m.mean32, m.sp[2 * src], m.sp[2 * src + 1],
                                  m.sp[2 - 2 * src], m.sp[3 - 2 * src], m.d, (m.d.W + 1) / 2, bx * TX, by * TY, z0,
                                  min(z0 + zc, m.d.D), m.omega, m.omw);
}

template <int AXIS>
__global__ __launch_bounds__(NT) void ke_correct(const EnsMember* __restrict__ M, EnsTable T, int B) {
    for_tiles<64, 4>(T, B, [&](int s, int, int i, int j, int k) {
        const EnsMember& m = M[s];
        float* face = AXIS == 0 ? m.u : (AXIS == 1 ? m.v : m.w);
        st_correct<AXIS, float, float>(m.occ, m.open_mask, m.p, face, m.d, i, j, k);
    });
}

__global__ __launch_bounds__(NT) void ke_copy(float* __restrict__ dst, const float* __restrict__ src, size_t n) {
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t x = (size_t)blockIdx.x * NT + threadIdx.x; x < n; x += stride) dst[x] = src[x];
}

__global__ __launch_bounds__(NT) void ke_zero(float* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t x = (size_t)blockIdx.x * NT + threadIdx.x; x < n; x += stride) dst[x] = 0.0f;
}

inline unsigned grid_of(const EnsTable& T, int blocks) {
    return (unsigned)(blocks > 0 && blocks < T.total ? blocks : (T.total > 0 ? T.total : 1));
}

}  // namespace

// tile layouts of the sweep kernel (TBX x TBY) the autotuner may pick
const int kEnsSweepShapes[kEnsSweepShapeCount][2] = {{64, 4}, {32, 8}, {16, 16}};
const int kEnsRbShapes[kEnsRbShapeCount][3] = {{64, 8, 8},  {64, 8, 16},  {64, 16, 8}, {64, 16, 16},
                                               {64, 16, 32}, {64, 32, 8}, {64, 32, 16}};

void ens_launch_step(const EnsLaunch& L, hipStream_t s) {
    const dim3 nt(NT);
    const EnsMember* M = L.members;
    const int B = L.count;
    const unsigned gc = grid_of(L.cells, L.blocks), gu = grid_of(L.ufaces, L.blocks), gv = grid_of(L.vfaces, L.blocks),
                   gw = grid_of(L.wfaces, L.blocks);
    // force + attenuation
    hipLaunchKernelGGL(ke_vorticity, dim3(gc), nt, 0, s, M, L.cells, B);
    hipLaunchKernelGGL(ke_force_cells, dim3(gc), nt, 0, s, M, L.cells, B);
    hipLaunchKernelGGL(ke_force_faces<0>, dim3(gu), nt, 0, s, M, L.ufaces, B);
    hipLaunchKernelGGL(ke_force_faces<1>, dim3(gv), nt, 0, s, M, L.vfaces, B);
    hipLaunchKernelGGL(ke_force_faces<2>, dim3(gw), nt, 0, s, M, L.wfaces, B);
    // velocity advection (inRegular_outRegular)
    hipLaunchKernelGGL(ke_mac_to_cell4, dim3(gc), nt, 0, s, M, L.cells, B);
    hipLaunchKernelGGL(ke_advect_cells, dim3(gc), nt, 0, s, M, L.cells, B);
    hipLaunchKernelGGL(ke_cells_to_faces<0>, dim3(gu), nt, 0, s, M, L.ufaces, B);
    hipLaunchKernelGGL(ke_cells_to_faces<1>, dim3(gv), nt, 0, s, M, L.vfaces, B);
    hipLaunchKernelGGL(ke_cells_to_faces<2>, dim3(gw), nt, 0, s, M, L.wfaces, B);
    // plain projection
    ens_launch_setup(L, s);
    if (L.sweep_rb) ens_launch_split_in(L, s);
    ens_launch_sweeps(L, 0, L.max_iters, s);
    if (L.sweep_rb) hipLaunchKernelGGL(ke_split_out, dim3(gc), nt, 0, s, M, L.cells, B);
    hipLaunchKernelGGL(ke_correct<0>, dim3(gu), nt, 0, s, M, L.ufaces, B);
    hipLaunchKernelGGL(ke_correct<1>, dim3(gv), nt, 0, s, M, L.vfaces, B);
    hipLaunchKernelGGL(ke_correct<2>, dim3(gw), nt, 0, s, M, L.wfaces, B);
    // scalar advection through the projected velocity, then the sources
    hipLaunchKernelGGL(ke_mac_to_cell4, dim3(gc), nt, 0, s, M, L.cells, B);
    hipLaunchKernelGGL(ke_advect_scalar, dim3(gc), nt, 0, s, M, L.cells, B);
    hipLaunchKernelGGL(ke_copy, dim3(2048), nt, 0, s, L.T_pool, L.T2_pool, L.cell_pool_size);
    hipLaunchKernelGGL(ke_copy, dim3(2048), nt, 0, s, L.R_pool, L.R2_pool, L.cell_pool_size);
    hipLaunchKernelGGL(ke_reinject, dim3(gc), nt, 0, s, M, L.cells, B);
}

void ens_launch_setup(const EnsLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL(ke_zero, dim3(2048), dim3(NT), 0, s, L.p_pool, L.p_pool_size);
    hipLaunchKernelGGL(ke_proj_setup, dim3(grid_of(L.cells, L.blocks)), dim3(NT), 0, s, L.members, L.cells, L.count);
    hipLaunchKernelGGL(ke_finish_mean, dim3(L.count), dim3(1024), 0, s, L.members, L.cells);
}

void ens_launch_probe(const EnsLaunch& L, hipStream_t s) {
    const unsigned gc = grid_of(L.cells, L.blocks);
    hipLaunchKernelGGL(ke_mac_to_cell4, dim3(gc), dim3(NT), 0, s, L.members, L.cells, L.count);
    hipLaunchKernelGGL(ke_advect_cells, dim3(gc), dim3(NT), 0, s, L.members, L.cells, L.count);
}

void ens_launch_split_in(const EnsLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL(ke_split_in, dim3(grid_of(L.cells, L.blocks)), dim3(NT), 0, s, L.members, L.cells, L.count);
}

void ens_launch_sweeps(const EnsLaunch& L, int it0, int it1, hipStream_t s) {
    const dim3 nt(NT);
    if (L.sweep_rb) {
        const int r = L.sweep_rb - 1;
        const int ty = kEnsRbShapes[r][1], zc = kEnsRbShapes[r][2];  // tiles of 64 slots
        const EnsTable& T = L.rb[r];
        const unsigned nb = (unsigned)((T.total + 7) & ~7);  // a multiple of 8 for the XCD numbering
        const EnsMember* M = L.members;
        for (int it = it0; it < it1; it++) {
            if (ty == 8)
                hipLaunchKernelGGL((ke_sor_split_rb<64, 8, 8>), dim3(nb), dim3(512), 0, s, M, T, L.count, zc, it);
            else if (ty == 16)
                hipLaunchKernelGGL((ke_sor_split_rb<64, 16, 8>), dim3(nb), dim3(512), 0, s, M, T, L.count, zc, it);
            else
                hipLaunchKernelGGL((ke_sor_split_rb<64, 32, 16>), dim3(nb), dim3(1024), 0, s, M, T, L.count, zc, it);
        }
        return;
    }
    const EnsTable& H = L.half[L.sweep_shape];
    const unsigned gh = grid_of(H, L.sweep_blocks);
    for (int it = it0; it < it1; it++)
        for (int c = 0; c < 2; c++) {
            switch (L.sweep_shape) {
            case 1: hipLaunchKernelGGL((ke_sor_half<32, 8>), dim3(gh), nt, 0, s, L.members, H, L.count, c, it); break;
            case 2: hipLaunchKernelGGL((ke_sor_half<16, 16>), dim3(gh), nt, 0, s, L.members, H, L.count, c, it); break;
            default: hipLaunchKernelGGL((ke_sor_half<64, 4>), dim3(gh), nt, 0, s, L.members, H, L.count, c, it); break;
            }
        }
}

}  // namespace zq3d
