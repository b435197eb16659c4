// zq3d_ensemble.h -- device-side description of an ensemble of smoke grids and the launchers of
// kernels_ensemble.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "zq3d_device.h"

namespace zq3d {

constexpr int kEnsMaxSources = 4;

// One member: device pointers into the ensemble's pools and its step parameters, precomputed on
// the host exactly as the single-grid launchers compute them.
struct EnsMember {
    Dims d;
    float *u, *v, *w, *T, *R, *T2, *R2;
    uint8_t *occ, *nc;
    float4* vel4;    // cell velocity; also the vorticity scratch of the force stage
    float* scratch;  // 3N: cell force, then advected cell velocity
    float *b, *p;    // divergence, padded pressure
    // colour-split SOR storage: counts and right-hand side per colour, padded pressure per
    // (pair, colour) at sp[2 * pair + colour]; the one-launch iteration ping-pongs between pairs
    uint8_t* sn[2];
    float* sb[2];
    float* sp[4];
    double* part;    // (sum, count) per set-up tile
    float* mean32;
    float h, dt, dth, two_h, ch, dtt;
    float inv_h;     // 1/h when h is a power of two (backtrace steps multiply), else 0 (they divide)
    int steps;
    float beta, alpha, Tamb, aV, aT, aR;
    int open_mask, closed_mean, iters;
    float omega, omw;
    int n_box, n_sph;
    float box[kEnsMaxSources][8];  // cx, cy, cz, hx, hy, hz (cells), rho, T
    float sph[kEnsMaxSources][6];  // x0, y0, z0 (cells), r^2 (cells^2), rho, T
};

// Packed tile table of one grid kind: member s owns tiles [start[s], start[s+1]), a gx x gy x gz
// grid of tiles.
struct EnsTable {
    const int* start;  // count + 1
    const int* gx;
    const int* gy;
    int total;
};

constexpr int kEnsSweepShapeCount = 3;
extern const int kEnsSweepShapes[kEnsSweepShapeCount][2];  // TBX x TBY of the sweep kernel
// slots x rows x planes per workgroup of the one-launch red-black iteration on colour-split storage
constexpr int kEnsRbShapeCount = 7;
extern const int kEnsRbShapes[kEnsRbShapeCount][3];

struct EnsLaunch {
    const EnsMember* members;  // device array
    int count;
    EnsTable cells, ufaces, vfaces, wfaces;
    EnsTable half[kEnsSweepShapeCount];  // half-row grids of the sweep, one per tile shape
    int sweep_shape;                     // index into kEnsSweepShapes
    EnsTable rb[kEnsRbShapeCount];       // tile tables of the one-launch iteration, one per shape
    int sweep_rb;                        // 0: half sweeps of sweep_shape; r > 0: kEnsRbShapes[r - 1]
    int blocks;                          // workgroups of the persistent launches (<= 0: one per tile)
    int sweep_blocks;
    int max_iters;
    float *p_pool, *T_pool, *T2_pool, *R_pool, *R2_pool;
    size_t p_pool_size, cell_pool_size;
};

// all stages of one step of every member
void ens_launch_step(const EnsLaunch& L, hipStream_t s);
// projection set-up of every member (pressure zeroed, counts, divergence, mean)
void ens_launch_setup(const EnsLaunch& L, hipStream_t s);
// mac -> cell velocity and cell advection of every member: writes only scratch, so it can be
// repeated to time launch shapes
void ens_launch_probe(const EnsLaunch& L, hipStream_t s);
// red-black iterations [it0, it1) of every member's SOR (on colour-split storage when sweep_rb)
void ens_launch_sweeps(const EnsLaunch& L, int it0, int it1, hipStream_t s);
// interleaved counts, right-hand side and pressure -> colour-split storage (pair 0)
void ens_launch_split_in(const EnsLaunch& L, hipStream_t s);

}  // namespace zq3d
