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
// divergence of every level, zero pressures
void launch_octree_prepare(const Dims& g, const float* u, const float* v, const float* w, const OctreeBufs& B,
                           hipStream_t s);
// `iters` passes of level 0 red, level 0 black, ..., level 3 black; mean != nullptr: the closed
// right-hand side div - mean * 8^L
void launch_octree_sweeps(const Dims& g, const OctreeBufs& B, int iters, const float* mean, hipStream_t s);
// face correction from the level pressures
void launch_octree_adjust(const Dims& g, const OctreeBufs& B, float* u, float* v, float* w, hipStream_t s);
// every level-L leaf's interior faces (L = 1, 2, 3) through the regular-block projection
void launch_octree_sub_project(const Dims& g, const OctreeBufs& B, float* u, float* v, float* w, hipStream_t s);

}  // namespace zq3d
