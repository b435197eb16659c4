/*
 * zq_smoke3d.h -- C ABI of the MI355X (gfx950) 3D smoke step.
 *
 * Part 1 replaces, symbol for symbol, the extern "C" entry points the reference's stage
 * functions call (ZQ_SmokeSimulationUtils3D.cpp:2088,2112,2157,2163,3835): same names, argument
 * order and meaning, enum values, host pointers, synchronous, elapsed milliseconds returned.
 * Under C++ they are declared inside the reference's namespaces, so `ZQ_CUDA_Advection3D::
 * Advect_Velocity(...)` in caller code resolves to the same unmangled symbol.
 * Part 2 is the red-black SOR pressure solve the reference takes from ZQlibCuda (absent from
 * the reference tree; signatures inferred from the calls at ZQ_SmokeSimulationUtils3D.cpp:2354-2363,
 * 2424-2437) plus an extended variant.
 * Part 3 keeps a whole simulation resident in HBM behind an opaque handle (no reference
 * counterpart: the reference copies every field host<->device at every stage).
 *
 * Errors: the reference aborts the process on a CUDA error (checkCudaErrors).  Here the Part 1/2
 * functions return a negative value instead and Part 3 returns a nonzero ZQ3D_ERR_*; the message
 * is available from zq3d_last_error() on the calling thread.
 */
#ifndef ZQ_SMOKE3D_H
#define ZQ_SMOKE3D_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

/* ====================== Part 1: reference stage entry points ====================== */

#ifdef __cplusplus
namespace ZQ_CUDA_Advection3D {
#endif
/* ZQ_CUDA_Advection3D.h:7-21 (values differ from ZQ_SmokeSimulationPara3D's enums) */
enum AdvectVelocityType {
    ADV_VEL_INREG_OUTREG = 0,
    ADV_VEL_INREG_OUTREG_BFECC,
    ADV_VEL_INMAC_OUTMAC,
    ADV_VEL_INMAC_OUTMAC_BFECC,
    ADV_VEL_INREG_OUTMAC,
    ADV_VEL_INREG_OUTMAC_BFECC
};
enum AdvectScalarType { ADV_SCA_MAC, ADV_SCA_MAC_BFECC, ADV_SCA_REG, ADV_SCA_REG_BFECC };
#ifdef __cplusplus
extern "C" {
#endif
/* replaces ZQ_CUDA_Advection3D.h:23 -- stores the grid and step constants used by the next calls */
void ZQ_Cuda_Prepare_Advection(const unsigned int width, const unsigned int height, const unsigned int depth,
                               const float voxelSize, const unsigned int steps, const float deltatt);
/* replaces ZQ_CUDA_Advection3D.h:31 -- u,v,w updated in place, all six modes (the BFECC ones as
 * ZQ_CUDA_Advection3D.cu:1110-1182, 1240-1336, 1389-1474, including the MAC variant leaving the
 * corrected field on faces its final apply skips). */
float Advect_Velocity(float* u, float* v, float* w, const bool* occupy, bool is_open, enum AdvectVelocityType type);
/* replaces ZQ_CUDA_Advection3D.h:43-44.  The BFECC modes leave the outputs zero, as the
 * reference's switch does (ZQ_CUDA_Advection3D.cu:1782-1794). */
float Advect_Scalar(const float* u, const float* v, const float* w, const bool* occupy,
                    const float* input_temperature, const float* input_density, float* output_temperature,
                    float* output_density, enum AdvectScalarType type);
#ifdef __cplusplus
}
}
namespace ZQ_CUDA_AddForce3D {
#endif
/* ZQ_CUDA_AddForce3D.h:6-10 */
enum AddForceType { ADD_FORCE_ENTIRE, ADD_FORCE_SLICE };
#ifdef __cplusplus
extern "C" {
#endif
/* replaces ZQ_CUDA_AddForce3D.h:12-14 -- buoyancy beta*(T-Tamb) on y plus vorticity confinement;
 * both types run the same computation, as in the reference (ZQ_CUDA_AddForce3D.cu:633-641). */
float AddForce3D(float* mac_u, float* mac_v, float* mac_w, const bool* occupy, const float* temperature,
                 const float deltah, const float deltat, const float buoyCoeff, const float confineCoeff,
                 const float Tamb, const int width, const int height, const int depth, enum AddForceType type);
#ifdef __cplusplus
}
}
namespace ZQ_CUDA_Attenuation3D {
extern "C" {
#endif
/* replaces ZQ_CUDA_Attenuation3D.h:6-8 */
float Attenuation3D(float* mac_u, float* mac_v, float* mac_w, float* temperature, float* density, const bool* occupy,
                    const float velAtten, const float tempAtten, const float densityAtten, const int width,
                    const int height, const int depth);
#ifdef __cplusplus
}
}
#endif

/* ====================== Part 2: red-black SOR projection ====================== */

#ifdef __cplusplus
namespace ZQ_CUDA_PoissonSolver3D {
extern "C" {
#endif
/* replaces the ZQlibCuda call at ZQ_SmokeSimulationUtils3D.cpp:2354-2363: p = 0 outside all six
 * domain faces, maxIter red-black Gauss-Seidel iterations (omega = 1) from p = 0, then u -= grad p. */
float SolveOpenPoissonRedBlackwithOccupy3D_MAC(float* mac_u, float* mac_v, float* mac_w, const bool* occupy,
                                               const int width, const int height, const int depth, const int maxIter);
/* replaces the ZQlibCuda call at ZQ_SmokeSimulationUtils3D.cpp:2424-2437: all domain faces are
 * walls; div_per_volume (the caller's mean divergence over fluid cells) is subtracted from every
 * fluid cell's divergence so the Neumann problem is compatible. */
float SolveClosedPoissonRedBlackwithOccupy3D_MAC(float* mac_u, float* mac_v, float* mac_w, const bool* occupy,
                                                 const float div_per_volume, const int width, const int height,
                                                 const int depth, const int maxIter);
/* replaces the ZQlibCuda call at ZQ_SmokeSimulationUtils3D.cpp:2630-2632 (the coarse grid of
 * ZQ_SolveCoarseOpenPoisson3DSOR): the face-fraction Poisson problem of ZQ_InitMatOpenPoisson3D
 * (:476-592; weight = open fraction of the face, p = 0 outside all six domain faces, cells with
 * occupy set skipped), maxIter red-black Gauss-Seidel iterations from p = 0 (colour (i+j+k)&1,
 * colour 0 first, neighbour sum x-, x+, y-, y+, z-, z+), then u -= p_r - p_l on every face with
 * fraction > 0.  Grid sizes are the coarse grid's; faces in the MAC layout. */
float SolveOpenPoissonRedBlackwithFaceRatio3D_MAC(float* mac_u, float* mac_v, float* mac_w, const bool* occupy,
                                                 const float* unoccupyU, const float* unoccupyV,
                                                 const float* unoccupyW, const int width, const int height,
                                                 const int depth, const int maxIter);
/* replace the ZQlibCuda calls at ZQ_SmokeOctreeGrid3D.cpp:2897-2907 and 3000-3010 (signature as
 * CPU_SolveOpenOctreePoissonRedBlack3_MAC, CPU_PoissonSolver3D_OpenOctreePoisson.h:5-14): maxIter
 * passes of red-black Gauss-Seidel over the four leaf levels from p = 0, then the face correction
 * between leaves (no sub-projection: the caller's SubProjection follows).  W, H, D multiples of 8.
 * The index / neighbour lists are what _selectInfos_for_CUDAOpenPoissonRedBlack3 builds from the
 * leaf flags; the same neighbours are derived from the flags on the device, so the lists are not
 * read.  The closed form subtracts div_per_volume times each leaf's volume from its divergence. */
float SolveOpenOctreePoissonRedBlack3_3D_MAC(
    float* mac_u, float* mac_v, float* mac_w, const bool* leaf0, const bool* leaf1, const bool* leaf2,
    const bool* leaf3, const int width, const int height, const int depth, const int maxIter,
    const int level0_num_red, const int* level0_index_red, const int level0_info_len_red,
    const int* level0_neighborinfo_red, const int level0_num_black, const int* level0_index_black,
    const int level0_info_len_black, const int* level0_neighborinfo_black, const int level1_num_red,
    const int* level1_index_red, const int level1_info_len_red, const int* level1_neighborinfo_red,
    const int level1_num_black, const int* level1_index_black, const int level1_info_len_black,
    const int* level1_neighborinfo_black, const int level2_num_red, const int* level2_index_red,
    const int level2_info_len_red, const int* level2_neighborinfo_red, const int level2_num_black,
    const int* level2_index_black, const int level2_info_len_black, const int* level2_neighborinfo_black,
    const int level3_num_red, const int* level3_index_red, const int level3_info_len_red,
    const int* level3_neighborinfo_red, const int level3_num_black, const int* level3_index_black,
    const int level3_info_len_black, const int* level3_neighborinfo_black);
float SolveClosedOctreePoissonRedBlack3_3D_MAC(
    float* mac_u, float* mac_v, float* mac_w, const bool* leaf0, const bool* leaf1, const bool* leaf2,
    const bool* leaf3, const float div_per_volume, const int width, const int height, const int depth,
    const int maxIter, const int level0_num_red, const int* level0_index_red, const int level0_info_len_red,
    const int* level0_neighborinfo_red, const int level0_num_black, const int* level0_index_black,
    const int level0_info_len_black, const int* level0_neighborinfo_black, const int level1_num_red,
    const int* level1_index_red, const int level1_info_len_red, const int* level1_neighborinfo_red,
    const int level1_num_black, const int* level1_index_black, const int level1_info_len_black,
    const int* level1_neighborinfo_black, const int level2_num_red, const int* level2_index_red,
    const int level2_info_len_red, const int* level2_neighborinfo_red, const int level2_num_black,
    const int* level2_index_black, const int level2_info_len_black, const int* level2_neighborinfo_black,
    const int level3_num_red, const int* level3_index_red, const int level3_info_len_red,
    const int* level3_neighborinfo_red, const int level3_num_black, const int* level3_index_black,
    const int level3_info_len_black, const int* level3_neighborinfo_black);
#ifdef __cplusplus
}
}
namespace ZQ_CUDA_MatrixMul {
extern "C" {
#endif
/* replaces ZQlibCuda's ZQ_CUDA_MatrixMul::MatrixMul, which SubProjection3DCuda wraps
 * (ZQ_SmokeSimulationUtils3D.cpp:13-16; SampleMatrixMul/main.cpp:18): C (h1 x w2) = A (h1 x w1) *
 * B (w1 x w2), row-major host arrays, each entry summed over k = 0..w1-1 in order. */
float MatrixMul(const float* A, const float* B, const int h1, const int w1, const int w2, float* C);
#ifdef __cplusplus
}
}
extern "C" {
#endif

/* domain faces across which p = 0 (open); the others are walls */
enum { ZQ3D_OPEN_XLO = 1, ZQ3D_OPEN_XHI = 2, ZQ3D_OPEN_YLO = 4, ZQ3D_OPEN_YHI = 8, ZQ3D_OPEN_ZLO = 16,
       ZQ3D_OPEN_ZHI = 32, ZQ3D_OPEN_ALL = 63 };

typedef struct zq3d_sor_opts {
    int open_mask;        /* ZQ3D_OPEN_* bits */
    int closed_mean;      /* subtract the mean divergence over fluid cells (FP64 on the device) */
    float omega;          /* SOR relaxation, 1 = Gauss-Seidel */
    int iters;            /* plain: iterations; mixed: FP32 sweeps per refinement step */
    int mixed_precision;  /* 0: FP32 SOR from p = 0.  1: FP64 residual + pressure, FP32 sweeps */
    int max_outer;        /* mixed: refinement steps at most (>= 1) */
    double tol;           /* mixed: stop once max|r| <= tol and the corrected faces (each rounded to FP32
                             once) have max|div - mean| <= tol */
    int pass_rule;        /* mixed: how long a refinement pass runs.  0: `iters` sweeps.  1: fitted, once a
                             pass has lowered max|r| (still above tol) the next runs only the sweeps that
                             pass's per-sweep decay needs to reach tol, times 1.25 (1 .. iters).  2: gated,
                             each pass sweeps until an iteration's largest update |new - old| is at most
                             0.75 tol omega / 6, or is no smaller than 16 iterations earlier (at most
                             `iters` <= 8192 sweeps); the iterations run back to back on the device,
                             which skips the rest of the pass */
} zq3d_sor_opts;

typedef struct zq3d_proj_stats {
    int outer;            /* refinement steps run (mixed) */
    int sweeps;           /* red-black iterations run in total (over all passes) */
    double max_residual;  /* mixed: final max|r| (FP64) */
    double max_div;       /* max |div - mean| over counted cells after the correction, FP64 */
    double mean;          /* subtracted mean divergence */
    int converged;        /* mixed: 1 when max_div <= tol, 0 when max_outer passes ended first;
                             -1 for the fixed-count solvers (no tolerance) */
} zq3d_proj_stats;

/* No reference counterpart (extends the two calls above): host pointers, omega, per-face open
 * boundaries and the mixed-precision mode.  stats may be NULL. */
float ZQ3D_SolvePoissonRedBlack_ex(float* mac_u, float* mac_v, float* mac_w, const bool* occupy, const int width,
                                   const int height, const int depth, const zq3d_sor_opts* opts,
                                   zq3d_proj_stats* stats);

/* No reference counterpart.  For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): ZQ3D_SolvePoissonRedBlack_ex on
 * DEVICE pointers (faces in the reference layout, occupancy one byte per cell, nonzero = solid),
 * stream-ordered on `stream` (a hipStream_t, NULL = the default stream), with the same kernels and
 * results bit for bit.  It calls no hipMalloc, hipFree or allocating library routine: its scratch
 * (counts, right-hand side, two sweep pressures, reduction partials; in the mixed mode also the FP64
 * pressure and right-hand side) lives in one static device array of ZQ3D_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends.  That bounds the grid: about 80^3 cells in the mixed mode and
 * 100^3 in FP32; larger grids are rejected with ZQ3D_ERR_ARG and a message giving the bytes needed.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race and must be serialised by the caller.  With stats == NULL and mixed_precision == 0
 * the call only enqueues work; stats, or the mixed mode's convergence test, read results back and
 * synchronize the stream.  Returns ZQ3D_OK or a ZQ3D_ERR_* code (message in zq3d_last_error). */
#define ZQ3D_NOALLOC_SCRATCH_BYTES (16u << 20)
int ZQ3D_SolvePoissonRedBlack_noalloc(float* mac_u, float* mac_v, float* mac_w, const bool* occupy,
                                      const int width, const int height, const int depth,
                                      const zq3d_sor_opts* opts, zq3d_proj_stats* stats, void* stream);

/* No reference counterpart (extends AddForce3D): also applies the density term
 * f_y -= gravityCoeff*rho (the reference's gravityCoeff, declared unused there) and adds an
 * optional cell-centred force field (3 floats per cell) before the face scatter.  density and
 * extra_force may be NULL (density is required when gravityCoeff != 0). */
float ZQ3D_AddForce3D_ex(float* mac_u, float* mac_v, float* mac_w, const bool* occupy, const float* temperature,
                         const float* density, const float* extra_force, const float deltah, const float deltat,
                         const float buoyCoeff, const float gravityCoeff, const float confineCoeff, const float Tamb,
                         const int width, const int height, const int depth);

/* replaces the reference's MAC->cell conversion (ZQ_CUDA_MACtoRegular.cu:9-30 on the device, CPU
 * twin MACtoRegular4 at ZQ_SmokeSimulationUtils3D.cpp:355-374): vel4 gets (u,v,w,0) per cell. */
float ZQ3D_MACtoRegular4(const float* mac_u, const float* mac_v, const float* mac_w, float* vel4, const int width,
                         const int height, const int depth);

/* ====================== Part 3: device-resident simulation ====================== */

enum {
    ZQ3D_OK = 0,
    ZQ3D_ERR_ARG = 1,      /* bad argument (sizes, NULL, mode) */
    ZQ3D_ERR_HIP = 2,      /* HIP runtime error */
    ZQ3D_ERR_STATE = 3     /* call not valid in the current state */
};

/* field ids for upload / download / device pointers */
enum {
    ZQ3D_FIELD_U = 0,      /* (W+1)*H*D float */
    ZQ3D_FIELD_V = 1,      /* W*(H+1)*D float */
    ZQ3D_FIELD_W = 2,      /* W*H*(D+1) float */
    ZQ3D_FIELD_T = 3,      /* W*H*D float */
    ZQ3D_FIELD_RHO = 4,    /* W*H*D float */
    ZQ3D_FIELD_OCC = 5,    /* W*H*D bytes, nonzero = solid */
    ZQ3D_FIELD_P64 = 6,    /* (W+2)*(H+2)*(D+2) double, padded pressure of the mixed-precision solve */
    ZQ3D_FIELD_VEL4 = 7    /* W*H*D float4: last MAC->cell conversion */
};

typedef struct zq3d_config {
    int struct_size;       /* sizeof(zq3d_config) */
    int width, height, depth;
    float voxel_size;      /* h; the reference uses 1/width (ZQ_SmokeSimulationUtils3D.cpp:323) */
    float step_time;       /* dt (stepTime) */
    int steps;             /* advection substeps (steps) */
    float Tamb;            /* ambient temperature */
    float buoyancy;        /* beta (buoyancyCoeff) */
    float gravity;         /* alpha: f_y -= alpha*rho (gravityCoeff, unused by the reference) */
    float confine;         /* eps (confineCoeff) */
    float vel_atten, temp_atten, dens_atten;   /* per-step factors exp(-dt*coeff) */
    int adv_vel_type;      /* enum AdvectVelocityType value */
    int adv_sca_type;      /* enum AdvectScalarType value */
    zq3d_sor_opts sor;     /* projection */
    int warm_start;        /* mixed: the projection's starting pressure.  0: zero.  1: the previous step's.
                              2: extrapolated in time from the previous two steps' pressures,
                              p1 + warm_extrapolate (p1 - p2) (p1 alone on the first step after a reset or
                              an upload of the pressure); the adjoint rollout starts each step from the
                              previous step's pressure */
    /* 0: projection on the full grid (TOTAL_*_POISSON_SOR).  s in [1, 8] dividing width, height and
     * depth: the reference's coarse + sub-grid projection (COARSE_OPEN_POISSON_SOR, subSize = s,
     * ZQ_SmokeSimulationUtils3D.cpp:3688-3795): sor.iters face-fraction SOR iterations on the
     * (W/s, H/s, D/s) grid with sor.omega and sor.open_mask, the coarse correction, then each s^3
     * block's interior faces projected (s > 1).  Requires sor.mixed_precision = 0.  With
     * sor.closed_mean the divergence's mean per unit fluid volume is removed first (closed domains). */
    int sub_size;
    /* sub_size > 0: 1 selects the flux form of the coarse solve (the *_FLUX_SOR solvers): an open
     * domain face weighs by its own open fraction, as in the KKT system of ZQ_InitMatOpenFlux3D
     * (ZQ_SmokeSimulationUtils3D.cpp:749-936), instead of 1 (ZQ_InitMatOpenPoisson3D, :476-592). */
    int coarse_flux;
    /* 1: the octree projection with the red-black solver (OPEN_OCTREE_POISSON_SOR /
     * CLOSED_OCTREE_POISSON_SOR, ZQ_SmokeOctreeGrid3D.cpp): leaves of 1, 2^3, 4^3 and 8^3 cells chosen
     * from the taxicab distance to the solid cells with thresholds octree_thresh (octree_thresh1..3,
     * defaults 2, 4, 16), sor.iters Gauss-Seidel passes over the levels from p = 0, the face
     * correction, then each 2^3 / 4^3 / 8^3 leaf's interior faces projected.  Every domain face beside
     * a leaf is open (a ghost pressure 0 four cells out; sor.open_mask is not used -- closed domains
     * carry a wall shell in the occupancy); with sor.closed_mean the mean divergence per fluid cell
     * times each leaf's volume is removed.  Requires width, height, depth multiples of 8, sub_size 0
     * and sor.mixed_precision 0. */
    int octree;
    int octree_thresh[3];
    /* warm_start 2: the extrapolation factor (default 0.5; 1 is linear extrapolation, which
     * amplifies the previous solves' errors more and measured erratic on start-up transients) */
    float warm_extrapolate;
} zq3d_config;

typedef struct zq3d_sim zq3d_sim;

/* fills defaults equal to ZQ_SmokeSimulationPara3D's constructor (ZQ_SmokeSimulationUtils3D.cpp:18-72)
 * mapped to the TOTAL_OPEN_POISSON_SOR solver */
int zq3d_config_default(zq3d_config* cfg, int width, int height, int depth);
int zq3d_create(const zq3d_config* cfg, zq3d_sim** out);
void zq3d_destroy(zq3d_sim* sim);
/* stream is a hipStream_t (NULL = the default stream) */
int zq3d_set_stream(zq3d_sim* sim, void* stream);
int zq3d_synchronize(zq3d_sim* sim);
int zq3d_upload(zq3d_sim* sim, int field, const void* host, size_t bytes);
int zq3d_download(zq3d_sim* sim, int field, void* host, size_t bytes);
int zq3d_device_ptr(zq3d_sim* sim, int field, void** dev, size_t* bytes);
/* box sources: cx,cy,cz,half_x,half_y,half_z (normalized), rho, T per source (ZQ_SmokeSource3D) */
int zq3d_set_box_sources(zq3d_sim* sim, int count, const float* boxes8);
/* sphere sources: cx,cy,cz (normalized), radius (units of width), rho, T */
int zq3d_set_sphere_sources(zq3d_sim* sim, int count, const float* spheres6);
/* solid sphere obstacle: occupancy = shell_mask walls + sphere, faces of solid voxels = velocity */
int zq3d_stamp_sphere_obstacle(zq3d_sim* sim, float cx, float cy, float cz, float radius, float su, float sv,
                               float sw, int shell_mask);

/* stages, in the order of ZQ_SmokeSimulation3D::UpdateOneFrame (ZQ_SmokeSimulation3D.cpp:156-243).
 * extra_force: NULL or a device pointer to 3 floats per cell added to the cell force. */
int zq3d_add_force_attenuate(zq3d_sim* sim, const float* extra_force_dev);
int zq3d_advect_velocity(zq3d_sim* sim);
int zq3d_project(zq3d_sim* sim, zq3d_proj_stats* stats);
int zq3d_advect_scalar(zq3d_sim* sim);
int zq3d_reinject(zq3d_sim* sim);
/* all stages of one step */
int zq3d_step(zq3d_sim* sim, const float* extra_force_dev, zq3d_proj_stats* stats);
/* MAC -> cell float4 into ZQ3D_FIELD_VEL4 (the reference's output conversion) */
int zq3d_mac_to_regular(zq3d_sim* sim);

/* Stage timing with HIP events on the simulation's stream (the reference prints per-stage
 * QueryPerformanceCounter / cudaEvent times, ZQ_SmokeSimulation3D.cpp:158-226).  on == 1: each stage
 * records an event pair and waits for it, so the stage numbers are exact kernel time, and every
 * red-black sweep loop is timed as well.  on == 2: only the sweep loops are timed, by event pairs
 * that nothing waits for until zq3d_get_timing, so the timing adds no synchronisation to a step. */
typedef struct zq3d_timing {
    double ms_force;            /* force + attenuation */
    double ms_advect_velocity;
    double ms_project;          /* whole projection, SOR included */
    double ms_sor;              /* red-black sweep launches only */
    double ms_advect_scalar;
    double ms_reinject;
    long long sor_launches;     /* single-colour sweep kernels timed in ms_sor */
    long long steps;
} zq3d_timing;
/* on = 0 (off), 1 (stages and sweeps) or 2 (sweeps only); enables and resets the counters */
int zq3d_set_profiling(zq3d_sim* sim, int on);
int zq3d_get_timing(zq3d_sim* sim, zq3d_timing* out);

/* Sweep-algorithm and tile-shape autotuner (no reference counterpart): each candidate valid for
 * the grid is timed for `iters` iterations with HIP events on the simulation's stream and
 * occupancy; the fastest is kept for this simulation.  Candidates (ty, zc): ty > 0 is the one-pass
 * kernel on interleaved storage (ty rows per workgroup, zc planes per chunk); ty == -1 is two
 * single-colour sweeps per iteration on colour-split storage; ty == -r (r = 4, 8, 16, 32) is one
 * launch per iteration on colour-split storage (r rows by zc = 8, 16, 32 planes per workgroup);
 * ty == -(100 + r) (r = 16, 32) is two iterations per launch on colour-split storage.  All produce
 * the same bits.
 * Clobbers the projection scratch (not the fields). */
#define ZQ3D_TUNE_MAX 32
typedef struct zq3d_tune_report {
    int count;                       /* candidates timed */
    int best;                        /* index of the chosen one (-1: the grid is too wide for the fused kernel) */
    int ty[ZQ3D_TUNE_MAX];
    int zc[ZQ3D_TUNE_MAX];
    float us_per_iter[ZQ3D_TUNE_MAX];
} zq3d_tune_report;
int zq3d_autotune_sor(zq3d_sim* sim, int iters, zq3d_tune_report* report);
/* explicit shape, one of the autotuner's candidates; (0, 0) restores the default */
int zq3d_set_sor_shape(zq3d_sim* sim, int ty, int zc);
/* Coarse + sub-grid projection (sub_size > 0; no reference counterpart): schedule of the coarse
 * red-black sweeps.  k == 0: one launch per colour; k > 0: k iterations per launch, each workgroup
 * owning zc z-planes of the coarse grid and recomputing a halo of 2k planes in LDS.  zq3d_create
 * times the valid schedules and keeps the fastest; all produce the same bits. */
int zq3d_set_coarse_sweep(zq3d_sim* sim, int k, int zc);
int zq3d_get_coarse_sweep(zq3d_sim* sim, int* k, int* zc);
/* Guiding control of the coarse + sub-grid projection (ZQ_ApplyGuidingVelocity3D,
 * ZQ_SmokeSimulationUtils3D.cpp:3139-3170, applied between the coarse solve and the coarse -> fine
 * correction, :3748-3755): while guide faces are set, every coarse face's delta gets
 * + coeff * (guide - averaged coarse face).  gu, gv, gw: host arrays of the coarse grid's faces
 * ((Wc+1)*Hc*Dc, Wc*(Hc+1)*Dc, Wc*Hc*(Dc+1) floats: the first three arrays of a .guidedat file at the
 * coarse resolution); gu == NULL turns guiding off.  Requires sub_size > 0 or octree != 0.  With
 * octree != 0, gu != NULL turns on the special-region projections of ZQ_SmokeOctreeGrid3D::
 * ApplyGuidingVelocity (ZQ_SmokeOctreeGrid3D.cpp:546-720, the reference's CUDA branch, run between
 * the octree's face correction and its sub-projection): each 6-connected region of fine cells in
 * level-0 and level-1 leaves projected alone by the closed red-black solver on its bounding box,
 * with min(sor.iters, region cells) iterations; that branch never adds the guide, so coeff and the
 * arrays' values are not read.  Host work per projection: the level map and faces cross PCIe. */
int zq3d_set_guide(zq3d_sim* sim, float coeff, const float* gu, const float* gv, const float* gw);
/* name of the sweep kernel the current shape runs ("k_sor_split_rb", "k_sor_split", "k_sor_fused"
 * or "k_sor_half") */
const char* zq3d_sor_kernel(zq3d_sim* sim);

/* ====================== Part 4: host-side scene set-up (no GPU) ====================== */

/* A reference config file parsed by ZQ_SmokeSimulationPara3D::LoadFromFile
 * (ZQ_SmokeSimulationUtils3D.cpp:78-352, same keys and quirks, plus the extension keys
 * buoyancyCoeff gravityCoeff sphereSource omega mixedPrecision sorTol maxOuter openFaces wallFaces
 * tempNoise deviceResident warmStart). */
#define ZQ3D_MAX_SOURCES 16
typedef struct zq3d_scene {
    int struct_size;              /* sizeof(zq3d_scene) */
    zq3d_config config;           /* grid, physics and projection as a device-resident config */
    int frame_count;
    int solver;                   /* GlobalSolverType3D value */
    int wall_faces;               /* domain faces given a one-cell solid shell */
    int internal_object;          /* 0 none, 1 sphere, 2 box, 3 oval, 4 special1 */
    int use_moving_object;
    int export_type;              /* 0 .di3, 1 .zqwv, 2 .zqci */
    int device_resident;
    int sub_size, max_iter, inner_iter, steps;
    float vel_atten_coeff, temp_atten_coeff, dens_atten_coeff;
    int adv_vel_type_para;        /* ZQ_SmokeSimulationPara3D::AdvectVelocityType value */
    int adv_sca_type_para;        /* ZQ_SmokeSimulationPara3D::AdvectScalarType value */
    int add_force_type;           /* ZQ_SmokeSimulationPara3D::AddForceType value */
    int n_box_sources;
    float box_sources[ZQ3D_MAX_SOURCES][8];
    int n_sphere_sources;
    float sphere_sources[ZQ3D_MAX_SOURCES][6];
    float temp_noise;
    unsigned noise_seed;
    char moving_object_file[200];
} zq3d_scene;
/* 0 on success; nonzero when the file is missing or LoadFromFile rejects it */
int zq3d_scene_load(const char* path, zq3d_scene* out);
/* ZQ_InitFineGrid3D for the config at path: occupancy (W*H*D bytes) of the initial grid */
int zq3d_scene_init_occupancy(const char* path, uint8_t* occ, size_t bytes);
/* ZQ_ApplyMovingObject3D applied after `frames` object updates of the config's moving-object file,
 * into caller arrays (occ W*H*D bytes, faces in the reference layout, updated in place) */
int zq3d_scene_apply_moving_objects(const char* path, int frames, uint8_t* occ, float* u, float* v, float* w);
/* reads a .di3 density file (the layout ZQ_SmokeGrid3D::ExportDensity writes, ZQ_SmokeGrid3D.cpp:55-81:
 * 16-byte NUL-terminated type tag "float", int32 width, height, depth, channels, a flag byte, FP32 data
 * x fastest): dims4 receives width, height, depth, channels; data (capacity floats) receives the
 * samples, or with data == NULL only the header is read and the file's length is checked against it.
 * 0 on success; nonzero for a missing, malformed or short file or too small a buffer. */
int zq3d_di3_read(const char* path, int* dims4, float* data, size_t capacity);

/* ====================== Part 5: adjoint rollout (gradient-based control) ====================== */

/* No reference counterpart.  A rollout of nsteps steps of the device-resident simulation (force +
 * attenuation with a per-step cell-centred control force added to the cell force, velocity
 * advection, mixed-precision projection, scalar advection, reinjection) and its hand-written
 * adjoint.  The forward keeps a checkpoint (state and the projection's warm start) every
 * checkpoint_every steps; the backward recomputes each segment from its checkpoint, storing the
 * segment's intermediate states, and runs the adjoint steps in reverse.
 * precision 4: every field FP32 (the product kernels); precision 8: every field FP64, the FD check
 * mode (the inner SOR sweeps stay FP32, the FP64 refinement converges to cfg->sor.tol).
 * Requires cfg->sor.mixed_precision, cfg->steps <= 32 and a non-BFECC advection mode.
 * Data pointers are device pointers of the chosen precision (occupancy: bytes), stream-ordered on
 * the rollout's stream.  The adjoint projection of step s is warm-started from the previous adjoint
 * step and solved to rtol * max|rhs|, where rtol is set by zq3d_adj_set_adjoint_tol or, when that is
 * 0 (the default), equals the forward projection's own tol / max|rhs| at step s (the same relative
 * accuracy as the forward). */
typedef struct zq3d_adj zq3d_adj;
typedef struct zq3d_adj_stats {
    long long forward_sweeps;    /* SOR sweeps of the last call's forward projections (recomputation included) */
    long long adjoint_sweeps;    /* SOR sweeps of the adjoint projections */
    int recomputed_steps;        /* forward steps rerun by the last backward */
    double max_div;              /* largest post-projection max|div| of the last forward */
    double ms;                   /* wall time of the call on the stream (HIP events) */
} zq3d_adj_stats;
int zq3d_adj_create(const zq3d_config* cfg, int precision, int nsteps, int checkpoint_every, zq3d_adj** out);
void zq3d_adj_destroy(zq3d_adj* a);
/* hipStream_t (NULL = the default stream) */
int zq3d_adj_set_stream(zq3d_adj* a, void* stream);
int zq3d_adj_set_sources(zq3d_adj* a, int n_box, const float* boxes8, int n_sphere, const float* spheres6);
int zq3d_adj_set_adjoint_tol(zq3d_adj* a, double rtol);
/* on != 0: keep every forward step's final FP64 pressure (nsteps x (W+2)(H+2)(D+2) doubles,
 * allocated on the first call) so that the backward's recomputed steps apply the forward's own
 * pressure instead of solving their projections again: the same correction, so the same states and
 * gradient, for O(nsteps) pressure memory beside the O(nsteps / checkpoint_every) checkpoints.  The
 * store is filled by the next zq3d_adj_forward and used by the zq3d_adj_backward that follows it
 * (whose controls are those of that forward); until then recomputed steps solve as without it. */
int zq3d_adj_set_pressure_store(zq3d_adj* a, int on);
/* sweep algorithm / tile shape of every projection of the rollout, one of zq3d_autotune_sor's
 * candidates for this grid (default (-1, 0)); all give the same bits */
int zq3d_adj_set_sor_shape(zq3d_adj* a, int ty, int zc);
/* initial state: ZQ3D_FIELD_U/V/W/T/RHO (device, chosen precision), ZQ3D_FIELD_OCC (device bytes) or
 * ZQ3D_FIELD_P64 (device doubles, padded: the forward projection's warm start; default 0) */
int zq3d_adj_set_initial(zq3d_adj* a, int field, const void* dev);
/* runs the rollout from the initial state.  controls: NULL or nsteps x (W*H*D) x 3 (per cell x, y, z);
 * rho_out: NULL or W*H*D, the density after the last step */
int zq3d_adj_forward(zq3d_adj* a, const void* controls, zq3d_adj_stats* st, void* rho_out);
/* given dL/d(density after the last step), writes dL/d(controls) (nsteps x W*H*D x 3); controls must
 * be those of the preceding forward */
int zq3d_adj_backward(zq3d_adj* a, const void* controls, const void* rho_bar, void* controls_bar,
                      zq3d_adj_stats* st);
/* The final density (rho_out, W*H*D) for `controls` without disturbing the latest forward: the
 * rollout restarts from that forward's checkpoint at or before from_step and runs to the end,
 * saving no checkpoint, tape or pressure.  The controls of the steps before that checkpoint must be
 * the forward's own (its saved states embody them).  For finite-difference checks and line
 * searches: a perturbation of the controls of step s needs only from_step = s. */
int zq3d_adj_evaluate(zq3d_adj* a, int from_step, const void* controls, zq3d_adj_stats* st, void* rho_out);
/* number of zq3d_adj_forward calls so far (-1 for a NULL rollout).  A backward always differentiates
 * the latest forward (its checkpoints, tape and pressure store): a caller holding an earlier forward's
 * result compares the count it saw with the current one before asking for that result's gradient. */
long long zq3d_adj_generation(zq3d_adj* a);
/* the recorded states of step `step` (absolute index) of the segment the rollout holds on tape (the
 * last segment after a forward; after a backward, the first): ZQ3D_FIELD_U/V/W the projected faces of
 * that step, ZQ3D_FIELD_T/RHO the temperature and density after its force stage; device memory of
 * the chosen precision, sizes as for zq3d_adj_set_initial.  For checking a backward's recomputation
 * against the forward. */
int zq3d_adj_get_tape(zq3d_adj* a, int step, int field, void* dev);
/* after a forward: the final state field; after a backward: with grad != 0, dL/d(initial field) */
int zq3d_adj_get(zq3d_adj* a, int field, int grad, void* dev);

/* ====================== Part 6: ensembles of uneven grids ====================== */

/* No reference counterpart.  count independent simulations, each with its own zq3d_config (grid
 * size and physics; ADV_VEL_INREG_OUTREG, ADV_SCA_REG and the plain FP32 projection with its own
 * sweep count), stepped together: every stage is one persistent launch over all members' tiles,
 * and the step (about 20 launches plus one or two per SOR iteration) is captured once in a hipGraph
 * and replayed.  Each member evolves bit for bit as a zq3d_sim of the same config.  Runs on its own
 * HIP stream (default priority). */
typedef struct zq3d_ens zq3d_ens;
int zq3d_ens_create(int count, const zq3d_config* configs, zq3d_ens** out);
void zq3d_ens_destroy(zq3d_ens* e);
/* at most 4 box and 4 sphere sources per member (formats of zq3d_set_box/sphere_sources) */
int zq3d_ens_set_sources(zq3d_ens* e, int member, int n_box, const float* boxes8, int n_sphere,
                         const float* spheres6);
/* ZQ3D_FIELD_U/V/W/T/RHO/OCC of one member, host memory, sizes as for zq3d_upload */
int zq3d_ens_upload(zq3d_ens* e, int member, int field, const void* host, size_t bytes);
int zq3d_ens_download(zq3d_ens* e, int member, int field, void* host, size_t bytes);
/* `steps` steps of every member; use_graph != 0 replays the captured step (captured on first use
 * and again after an autotune) */
int zq3d_ens_step(zq3d_ens* e, int steps, int use_graph);
int zq3d_ens_synchronize(zq3d_ens* e);
/* total cells of all members */
long long zq3d_ens_cells(zq3d_ens* e);
/* the ensemble's hipStream_t */
int zq3d_ens_stream(zq3d_ens* e, void** stream);
/* Launch-shape autotuner: times the sweep kernels -- kind 0: single-colour half sweeps for each
 * tile shape (tbx x tby cells) and persistent launch size (blocks; 0 = one workgroup per tile);
 * kind 1: one launch per red-black iteration on colour-split storage, tiles of tbx slots x tby rows
 * x `blocks` planes (the default before any autotune is 64 x 16 x 16) -- and the cell kernels for
 * each launch size, on the members' own grids and occupancy, and keeps the fastest.  All give the
 * same bits.  Fields are unchanged. */
#define ZQ3D_ENS_TUNE_MAX 24
typedef struct zq3d_ens_tune {
    int count;
    int kind[ZQ3D_ENS_TUNE_MAX], tbx[ZQ3D_ENS_TUNE_MAX], tby[ZQ3D_ENS_TUNE_MAX], blocks[ZQ3D_ENS_TUNE_MAX];
    float us_per_iter[ZQ3D_ENS_TUNE_MAX];
    int best_kind, best_tbx, best_tby, best_sweep_blocks;  /* kind 1: best_sweep_blocks = planes */
    float best_us_per_iter;
    int best_blocks;       /* launch size of the other stages */
    float probe_ms;        /* cell advection probe at that size */
} zq3d_ens_tune;
int zq3d_ens_autotune(zq3d_ens* e, int iters, zq3d_ens_tune* report);
/* selects one of the autotuner's sweep candidates (kind, tbx, tby, blocks as reported) */
int zq3d_ens_set_sweep(zq3d_ens* e, int kind, int tbx, int tby, int blocks);

/* message of the last failed call on this thread ("" if none) */
const char* zq3d_last_error(void);
/* frees the device workspace cached by the Part 1/2 host-pointer functions */
void zq3d_release_workspace(void);

#ifdef __cplusplus
}
#endif
#endif
