// zq3d_octree.h -- launchers of the octree projection with the red-black solver (the reference's
// OPEN_OCTREE_POISSON_SOR / CLOSED_OCTREE_POISSON_SOR: ZQ_SmokeOctreeGrid3D.cpp, solver kernels as
// CPU_PoissonSolver3D_OpenOctreePoisson.cpp states them) on device pointers.
//
// The fine grid (W, H, D), all multiples of 8, is covered by leaves of four levels: fine cells
// (level 0) and aligned 2^3, 4^3 and 8^3 blocks (levels 1-3), chosen from the taxicab distance to the
// solid cells.  One byte per fine cell (`lev`) names the level of the leaf holding it (4: solid);
// per level L the pressure and divergence live in dense (W >> L, H >> L, D >> L) arrays, x fastest.
// A projection: divergence of every level, red-black Gauss-Seidel over the levels, the face
// correction between leaves, then every level-L leaf's interior faces projected as a regular block.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zq3d_device.h"

namespace zq3d {

constexpr uint8_t OCT_SOLID = 4;  // lev value of a solid fine cell

struct OctreeBufs {
    uint8_t* lev;               // level of the leaf holding each fine cell (OCT_SOLID: solid)
    uint8_t* leaf[4];           // leaf flags per level (leaf[0] unused: lev == 0)
    int32_t* dist;              // taxicab distance to the nearest solid cell
    int32_t* mn[4];             // per level L >= 1: the smallest distance in each block
    float* p[4];                // pressure per level
    float* d[4];                // divergence per level
    uint32_t* list[4][2];       // per level and colour: the linear level index of every cell solved
    uint32_t* count;            // 8 list lengths (level-major, colour-minor), then 8 overflow totals
};

// Per list k = 2 L + colour: the neighbour entries of its cells (the reference's
// _selectInfos_for_CUDAOpenPoissonRedBlack3 lists, built on the device) and their right-hand sides,
// rebuilt with the leaves in grow-only device buffers owned here.  head holds each cell's first six
// entries transposed (entry j of cell i at j * n + i); the rest of a cell's entries sit in ovf at
// ovf_base[k] + off[k][i].
struct OctreeLists {
    uint32_t counts[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // cells per level and colour
    uint32_t* head[8] = {};
    uint32_t* tot[8] = {};    // entries per cell
    uint32_t* ovc[8] = {};    // entries beyond the head per cell (n + 1 values, last 0)
    uint32_t* off[8] = {};    // exclusive sums of ovc
    float* rhs[8] = {};       // right-hand side per cell, in list order
    size_t ovf_base[8] = {};
    void* dyn = nullptr;
    size_t dyn_cap = 0;
    uint32_t* ovf = nullptr;
    size_t ovf_cap = 0;
    void* scan = nullptr;
    size_t scan_cap = 0;
};

// bytes of one OctreeBufs allocation for the grid, and the carving of `mem` into it
size_t octree_bytes(const Dims& g);
void octree_layout(const Dims& g, char* mem, OctreeBufs* B);

// distance field, leaf flags and the level map from the occupancy (thresholds t1, t2, t3)
void launch_octree_build(const Dims& g, const uint8_t* occ, int t1, int t2, int t3, const OctreeBufs& B,
                         hipStream_t s);
// level map and leaf flags from caller-given leaf flags of the four levels (0/1 bytes)
void launch_octree_from_leaves(const Dims& g, const uint8_t* l0, const uint8_t* l1, const uint8_t* l2,
                               const uint8_t* l3, const OctreeBufs& B, hipStream_t s);
// the cell lists and neighbour entries from the level map (synchronises the stream twice: the list
// lengths and entry totals size the launches and the entry buffer)
hipError_t octree_build_lists(const Dims& g, const OctreeBufs& B, OctreeLists* Ls, hipStream_t s);
void octree_free_lists(OctreeLists* Ls);
// divergence of every level, zero pressures
void launch_octree_prepare(const Dims& g, const float* u, const float* v, const float* w, const OctreeBufs& B,
                           hipStream_t s);
// the right-hand sides (mean != nullptr: the closed form div - mean * 8^L), then `iters` passes of
// level 0 red, level 0 black, ..., level 3 black, one launch per phase.  Returns the launches made.
int launch_octree_sweeps(const Dims& g, const OctreeBufs& B, const OctreeLists& Ls, int iters, const float* mean,
                         hipStream_t s);
// face correction from the level pressures
void launch_octree_adjust(const Dims& g, const OctreeBufs& B, float* u, float* v, float* w, hipStream_t s);
// every level-L leaf's interior faces (L = 1, 2, 3) through the regular-block projection
void launch_octree_sub_project(const Dims& g, const OctreeBufs& B, float* u, float* v, float* w, hipStream_t s);

}  // namespace zq3d
