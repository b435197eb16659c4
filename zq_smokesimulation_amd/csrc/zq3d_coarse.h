// zq3d_coarse.h -- launchers of the coarse + sub-grid projection (the reference's default solver
// COARSE_OPEN_POISSON_SOR, ZQ_SmokeSimulationUtils3D.cpp:3688-3795) on device pointers.
//
// The fine grid (W, H, D) is split into s^3 blocks; the coarse grid (Wc, Hc, Dc) = (W/s, H/s, D/s)
// keeps one cell per block and its faces in the fine layout (x fastest, u (Wc+1) Hc Dc, ...).
// A projection is: coarse face open fractions (setup), fine -> coarse face averages, red-black SOR
// of the face-fraction Poisson problem on the coarse grid, the coarse correction added to the fine
// faces on coarse planes, then every block's interior faces projected with its boundary faces
// held (regular blocks: fast separable solve; blocks with solid cells: FP64 CG in the workgroup).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zq3d_device.h"

namespace zq3d {

constexpr int COARSE_MAX_S = 8;  // largest block edge the sub-projection kernels are built for

struct CoarseDims {
    Dims f;        // fine grid
    int s;         // block edge
    int Wc, Hc, Dc;
    size_t Nc() const { return (size_t)Wc * Hc * Dc; }
    size_t Nu() const { return (size_t)(Wc + 1) * Hc * Dc; }
    size_t Nv() const { return (size_t)Wc * (Hc + 1) * Dc; }
    size_t Nw() const { return (size_t)Wc * Hc * (Dc + 1); }
    size_t Np() const { return (size_t)(Wc + 2) * (Hc + 2) * (Dc + 2); }
};
inline CoarseDims coarse_dims(const Dims& f, int s) { return CoarseDims{f, s, f.W / s, f.H / s, f.D / s}; }

// Device arrays of one coarse grid.
struct CoarseBufs {
    float *rU, *rV, *rW;     // face open fractions
    float *cu, *cv, *cw;     // averaged coarse faces
    float* inv;              // 1/diagonal of the face-fraction operator per coarse cell (0: not updated)
    float* div;              // coarse divergence
    float* p;                // padded coarse pressure (zero ghost layer)
    float* p2;               // second pressure the multi-iteration sweep ping-pongs with
    uint8_t* cocc;           // coarse occupancy
    uint8_t* regular;        // 1: the block holds no solid fine cell
    float *gU, *gV, *gW;     // guide faces (guiding control)
    float* vol;              // fluid volume fraction 1 - (solid fine cells)/s^3 per coarse cell
    double* mean;            // closed domains: mean divergence per unit fluid volume
};

// fractions, coarse occupancy, 1/diagonal (open_mask: open domain faces; flux: the flux form's
// diagonal), volume fractions and block flags from the fine occupancy
void launch_coarse_setup(const CoarseDims& g, const uint8_t* occ, int open_mask, const CoarseBufs& B, hipStream_t s,
                         int flux = 0);
// coarse faces = mean of their s x s fine sub-faces; coarse divergence; p = p2 = 0
void launch_coarse_restrict(const CoarseDims& g, const float* u, const float* v, const float* w, const CoarseBufs& B,
                            hipStream_t s);
// Sweep schedule of the coarse SOR: k == 0 launches each colour separately; k > 0 runs k red-black
// iterations per launch on slabs of zc z-planes (halo planes recomputed).  All give the same bits.
struct CoarseSweep {
    int k, zc;
};
bool coarse_sweep_valid(const CoarseDims& g, CoarseSweep sw);
// schedules valid for the grid (LDS of a slab within 160 KiB); returns the count
int coarse_sweep_candidates(const CoarseDims& g, CoarseSweep* out, int max);
// closed domains: div -= mean * vol on every non-solid coarse cell, mean = sum(div) / sum(vol) over
// them in FP64 (one workgroup, fixed summation order, so repeatable)
void launch_coarse_remove_mean(const CoarseDims& g, const CoarseBufs& B, hipStream_t s);
// `iters` red-black iterations from B.p (colour 0 first); returns B.p or B.p2, whichever holds the result
const float* launch_coarse_sor(const CoarseDims& g, const CoarseBufs& B, float omega, int iters, CoarseSweep sw,
                               hipStream_t s);
// adds the coarse correction (pressure p; with guide != 0 also guide_coeff * (guide - coarse face))
// to the fine faces on coarse planes (fine cells beside them fluid)
void launch_coarse_prolong(const CoarseDims& g, const CoarseBufs& B, const float* p, int open_mask, const uint8_t* occ,
                           float* u, float* v, float* w, hipStream_t s, int guide = 0, float guide_coeff = 0.0f);
// projects every block's interior faces; s in [2, COARSE_MAX_S]
void launch_sub_project(const CoarseDims& g, const CoarseBufs& B, const uint8_t* occ, float* u, float* v, float* w,
                        hipStream_t s);
// u -= grad p on the coarse faces with fraction > 0 (domain faces: open ones), in place
void launch_coarse_correct(const CoarseDims& g, const CoarseBufs& B, const float* p, int open_mask, float* u, float* v,
                           float* w, hipStream_t s);
// coarse divergence, 1/diagonal and p = p2 = 0 from caller-given coarse faces, occupancy and fractions
void launch_coarse_prepare(const CoarseDims& g, const float* u, const float* v, const float* w, int open_mask,
                           const CoarseBufs& B, hipStream_t s);
// the regular-block projection of the interior faces of every bs^3 block (bs = 2, 4 or 8) whose mask
// byte (a (W/bs, H/bs, D/bs) array) is set; w_first: see k_sub_regular
void launch_sub_regular_mask(const Dims& f, int bs, const uint8_t* mask, float* u, float* v, float* w, int w_first,
                             hipStream_t s);
// out (rows x n) = M (rows x cols) * in (cols x n), each output summed over k = 0..cols-1 in order
void launch_batched_matvec(int rows, int cols, int n, const float* M, const float* in, float* out, hipStream_t s);

}  // namespace zq3d
