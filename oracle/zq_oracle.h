/*
 * zq_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the zuoqing1988/ZQ_SmokeSimulation 3D smoke step
 * (the path behind ZQ_SmokeSimulation3D / ZQ_SmokeGrid3D).  It exists to check
 * the HIP kernels and to provide the timed CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Layout (ZQ_SmokeGrid3D.cpp:5-18): x fastest.
 *   cell   (i,j,k) : (k*H + j)*W + i               N  = W*H*D
 *   u face (i,j,k) : (k*H + j)*(W+1) + i, i<=W     Nu = (W+1)*H*D
 *   v face (i,j,k) : (k*(H+1) + j)*W + i, j<=H     Nv = W*(H+1)*D
 *   w face (i,j,k) : (k*H + j)*W + i,     k<=D     Nw = W*H*(D+1)
 * occupancy is one byte per cell, nonzero = solid.
 *
 * Parity status of each entry point is written beside it ("pinned by": the
 * tests that fix it; "parity unpinned": what nothing in the reference fixes).
 */
#ifndef ZQ_ORACLE_H
#define ZQ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Advection modes, values of the CUDA-side enums (ZQ_CUDA_Advection3D.h:7-21). */
enum {
    ORC_ADV_VEL_INREG_OUTREG = 0,
    ORC_ADV_VEL_INREG_OUTREG_BFECC = 1,
    ORC_ADV_VEL_INMAC_OUTMAC = 2,
    ORC_ADV_VEL_INMAC_OUTMAC_BFECC = 3,
    ORC_ADV_VEL_INREG_OUTMAC = 4,
    ORC_ADV_VEL_INREG_OUTMAC_BFECC = 5
};
enum {
    ORC_ADV_SCA_MAC = 0,
    ORC_ADV_SCA_MAC_BFECC = 1,
    ORC_ADV_SCA_REG = 2,
    ORC_ADV_SCA_REG_BFECC = 3
};

/* Domain-face bits for open (p = 0 outside) boundaries. */
enum { ORC_OPEN_XLO = 1, ORC_OPEN_XHI = 2, ORC_OPEN_YLO = 4, ORC_OPEN_YHI = 8, ORC_OPEN_ZLO = 16, ORC_OPEN_ZHI = 32 };

/* ---- MAC <-> cell-centred velocity (ZQ_CUDA_MACtoRegular.cu:9-99) ---- */
void orc_mac_to_cell4(int W, int H, int D, const float* u, const float* v, const float* w, float* vel4);
void orc_cell4_to_mac(int W, int H, int D, const float* vel4, float* u, float* v, float* w);

/* ---- trilinear, clamp-to-edge sample in texel space (texel centres at integers) ---- */
float orc_sample3(const float* g, int nx, int ny, int nz, float x, float y, float z);

/* ---- force stage (ZQ_CUDA_AddForce3D.cu:11-420, 510-556) ----
 * vort4 (optional, may be NULL): per cell (wx,wy,wz,|w|); gradn3 (optional): normalized grad|w|;
 * force3 (optional): cell force.  extra_force (optional, 3 floats per cell): added to the cell
 * force before the face scatter (control input, not in the reference). */
void orc_add_force(int W, int H, int D, float* u, float* v, float* w, const uint8_t* occ,
                   const float* T, const float* rho, float h, float dt, float beta, float alpha,
                   float eps, float Tamb, const float* extra_force,
                   float* vort4, float* gradn3, float* force3);

/* ---- attenuation (ZQ_CUDA_Attenuation3D.cu:8-116) ---- */
void orc_attenuation(int W, int H, int D, float* u, float* v, float* w, float* T, float* rho,
                     const uint8_t* occ, float aV, float aT, float aR);

/* ---- semi-Lagrangian advection (ZQ_CUDA_Advection3D.cu:76-650, 1061-1108, 1184-1238, 1338-1387, 1476-1505) ----
 * All six velocity modes (BFECC: :1110-1182, 1240-1336, 1389-1474).  Pinned by: zero velocity and
 * uniform translation of linear fields (identity / exact shift, BFECC included), BFECC's lower error
 * than the plain mode on a translated smooth field, the MAC variant's kept phi~ on skipped faces.
 * Returns 0 on success, -1 for an unsupported mode. */
int orc_advect_velocity(int W, int H, int D, float h, int steps, float dtt, float* u, float* v, float* w,
                        const uint8_t* occ, int is_open, int type);
int orc_advect_scalar(int W, int H, int D, float h, int steps, float dtt, const float* u, const float* v,
                      const float* w, const uint8_t* occ, const float* Tin, const float* Rin,
                      float* Tout, float* Rout, int type);
/* Backtrace end point (lastpos) from a start position, regular or MAC velocity sampling. */
void orc_backtrace(int W, int H, int D, float h, int steps, float dtt, const float* u, const float* v,
                   const float* w, const uint8_t* occ, int mac_velocity, float px, float py, float pz,
                   float* out3);

/* ---- projection: divergence, red-black SOR, gradient correction ----
 * Parity unpinned against ZQlibCuda, which is absent: ordering, omega, initial guess and solid
 * handling below are this project's own definition.  What is pinned is the converged projection
 * (FP64 CG on the same operator, divergence-free and curl-free-correction checks, idempotence,
 * symmetry). */
void orc_divergence(int W, int H, int D, const float* u, const float* v, const float* w,
                    const uint8_t* occ, float* div);
/* neighbour count per cell (0 = not updated), following the open-face mask */
void orc_sor_count(int W, int H, int D, const uint8_t* occ, int open_mask, uint8_t* ncount);
/* K red-black SOR iterations on padded p ((W+2)(H+2)(D+2)), rhs b (unpadded). */
void orc_sor(int W, int H, int D, const uint8_t* ncount, const float* b, float* p_pad, float omega, int iters);
/* orc_sor that stops early: before iteration it >= 1 it returns when the largest update |new - old|
 * of iteration it-1 is <= thresh, or (it > 32) no smaller than that of iteration it-17 (the updates
 * no longer shrink: the FP32 iterate's rounding floor).  Sizes compare as FP32 bit patterns, so a NaN
 * update never stops it.  Returns the iterations run. */
int orc_sor_gated(int W, int H, int D, const uint8_t* ncount, const float* b, float* p_pad, float omega, int iters,
                  float thresh);
/* the gated passes' threshold: (float)(0.75 * tol * omega / 6) */
float orc_gate_threshold(double tol, float omega);
/* full FP32 projection: div -> b = div - mean -> SOR(iters) from p=0 -> correct faces.
 * p_pad_out (optional) receives the final padded pressure. */
void orc_project(int W, int H, int D, float* u, float* v, float* w, const uint8_t* occ, int open_mask,
                 float mean, float omega, int iters, float* p_pad_out);
/* mixed precision: FP32 SOR inner sweeps, FP64 residual and pressure, iterative refinement.
 * p64_pad (in/out, padded, FP64) is the warm start.  Returns the number of outer iterations;
 * stats[0] = total inner sweeps, stats[1] = final max|r|, stats[2] = max|div| after correction.
 * pass_rule != 0: pass lengths from orc_fitted_pass_length (else inner_iters each). */
int orc_project_mp(int W, int H, int D, float* u, float* v, float* w, const uint8_t* occ, int open_mask,
                   int subtract_mean, float omega, int inner_iters, int max_outer, double tol,
                   double* p64_pad, double* stats, int pass_rule);
/* Sweeps of the next refinement pass when the pass lengths are fitted: r is the max|r| the last pass
 * left, rprev the one before it, nprev that pass's sweeps.  Once a pass has lowered max|r| (still
 * above tol), the next pass runs the sweeps that pass's per-sweep decay (r/rprev)^(1/nprev) needs to
 * reach tol, times 1.25, at least 1 and at most iters; otherwise iters. */
int orc_fitted_pass_length(int iters, int fit, double r, double rprev, int nprev, double tol);
/* Starting pressure extrapolated in time (the device's warm_start 2): over n padded cells,
 * p = p + beta (p - prev) and prev = the old p; valid == 0: prev = p only. */
void orc_extrapolate_warm_start(size_t n, double* p64_pad, double* prev_pad, int valid, double beta);
/* FP64 conjugate gradient on the same operator (pins the converged projection).
 * Writes corrected faces into u,v,w (FP32) and returns iterations. */
int orc_project_cg64(int W, int H, int D, float* u, float* v, float* w, const uint8_t* occ, int open_mask,
                     int subtract_mean, double tol, int max_iter, double* p64_pad);
/* max |div| over cells with a nonzero neighbour count, FP64 */
double orc_max_abs_div(int W, int H, int D, const float* u, const float* v, const float* w,
                       const uint8_t* occ, int open_mask, double mean);
/* FP32 sequential mean divergence over fluid cells (ZQ_SmokeSimulationUtils3D.cpp:2387-2421) */
float orc_closed_mean_divergence(int W, int H, int D, const float* u, const float* v, const float* w,
                                 const uint8_t* occ, int* count_out);

/* FP32 divergences summed in FP64 over fluid cells / count, as float (device-resident path) */
float orc_mean_div32_f64(int W, int H, int D, const float* u, const float* v, const float* w, const uint8_t* occ);

/* ---- grid set-up, sources, obstacles ---- */
/* ZQ_SmokeSimulationUtils3D.cpp:1523-1664: solid shell on the domain faces in shell_mask (the
 * reference's closed solvers use all six = 63); internal object 0 none,1 sphere,2 box,3 oval,4 special1 */
void orc_init_fine_grid(int W, int H, int D, int shell_mask, int intobj, uint8_t* occ, float* u, float* v, float* w);
/* ZQ_SmokeSimulationUtils3D.cpp:3842-3882 (box sources; randDensity off). box = cx,cy,cz,hx,hy,hz,rho,T */
void orc_reinject_box(int W, int H, int D, const float* box, float* T, float* rho);
/* extension: sphere source at normalized centre, radius in units of W */
void orc_reinject_sphere(int W, int H, int D, float cx, float cy, float cz, float r, float rho_src, float T_src,
                         float* T, float* rho);
/* extension: analytic solid sphere (voxel-centre test), faces of solid voxels set to (su,sv,sw),
 * then the shell of shell_mask (semantics of ZQ_SmokeSimulationUtils3D.cpp:1933-2046). */
void orc_stamp_sphere(int W, int H, int D, float cx, float cy, float cz, float r, float su, float sv, float sw,
                      int shell_mask, uint8_t* occ, float* u, float* v, float* w);
/* OpenMP threads the oracle's parallel loops use */
int orc_num_threads(void);
void orc_set_num_threads(int n);
/* counter-based hash noise in [-amp, amp] per cell (identical in the product) */
void orc_noise(int W, int H, int D, uint32_t seed, float amp, float* out);

/* ---- coarse + sub-grid projection (COARSE_OPEN_POISSON_SOR, ZQ_SmokeSimulationUtils3D.cpp:3688-3795),
 * in zq_oracle_coarse.cpp.  s divides W, H and D; coarse grid Wc = W/s etc., faces in the same layout.
 * Pinned by: the face fractions, averages and coarse -> fine correction are integer/indexing work
 * checked against brute-force numpy restatements and closed forms; the coarse solver reduces bit for
 * bit to orc_sor with unit weights (s = 1) and converges to the dense solve of the same system;
 * the sub-block operator matches the reference's mat.txt (s = 2).  Parity unpinned: the coarse
 * solver's sweep order and omega (ZQlibCuda, absent) and the sub-block product's summation order
 * (ZQ_CUDA_MatrixMul, absent). */
void orc_coarse_setup(int W, int H, int D, int s, const uint8_t* occ, float* rU, float* rV, float* rW,
                      uint8_t* cocc, uint8_t* regular);
float orc_coarse_inv_diag(int Wc, int Hc, int Dc, int i, int j, int k, const float* rU, const float* rV,
                          const float* rW, int open_mask, int flux);
void orc_fine_to_coarse(int W, int H, int D, int s, const float* u, const float* v, const float* w, float* cu,
                        float* cv, float* cw);
void orc_coarse_project(int Wc, int Hc, int Dc, const float* cu, const float* cv, const float* cw,
                        const uint8_t* cocc, const float* rU, const float* rV, const float* rW, int open_mask,
                        float omega, int iters, float* p, float* du, float* dv, float* dw, const float* vol,
                        int flux);
void orc_coarse_to_fine(int W, int H, int D, int s, const float* du, const float* dv, const float* dw,
                        const uint8_t* occ, float* u, float* v, float* w);
/* M: R x F row-major float, R = 3 s^2 (s-1), F = 3 s^2 (s+1) */
void orc_sub_project_regular(int W, int H, int D, int s, const float* M, const uint8_t* regular, float* u, float* v,
                             float* w);

/* ---- one forward step (ZQ_SmokeSimulation3D.cpp:156-243 order) ---- */
typedef struct orc_step_params {
    int W, H, D;
    float h, dt;
    int steps;             /* advection substeps */
    float Tamb, beta, alpha, eps;
    float aV, aT, aR;      /* attenuation factors */
    int adv_vel_type, adv_sca_type;
    int open_mask;         /* open domain faces */
    int closed_mean;       /* subtract mean divergence (closed domains) */
    float omega;
    int sor_iters;         /* plain: iterations; mixed: inner sweeps per outer */
    int mixed;             /* 0 plain FP32 SOR from p=0; 1 mixed precision */
    int max_outer;
    double tol;
    int n_box; const float* boxes;              /* 8 floats each */
    int n_sphere; const float* spheres;         /* cx,cy,cz,r,rho,T each */
    int pass_rule;        /* mixed: pass lengths, as orc_project_mp */
} orc_step_params;
/* p64_pad may be NULL in plain mode; in mixed mode it carries the warm start */
int orc_step(const orc_step_params* P, float* u, float* v, float* w, float* T, float* rho, const uint8_t* occ,
             const float* extra_force, double* p64_pad, double* stats);

/* ---- octree projection with the red-black solver (OPEN_OCTREE_POISSON_SOR, CLOSED_OCTREE_POISSON_SOR,
 * ZQ_SmokeOctreeGrid3D.cpp), in zq_oracle_octree.cpp.  W, H, D are multiples of 8; level L holds
 * (W >> L, H >> L, D >> L) cells, x fastest.  Pinned by: the distance field against a brute-force
 * taxicab distance, the leaf flags against their closed form, the neighbour lists against brute-force
 * face adjacency, the converged solve against a direct sparse solve of the same system and against
 * zero net flux through every leaf after the correction.  Parity unpinned: ZQlibCuda's closed octree
 * solver (absent) -- the closed right-hand side div - mean * volume is defined here. */
void orc_oct_distance(int W, int H, int D, const uint8_t* occ, int32_t* dist);
void orc_oct_leaves(int W, int H, int D, const uint8_t* occ, const int32_t* dist, int t1, int t2, int t3, uint8_t* l0,
                    uint8_t* l1, uint8_t* l2, uint8_t* l3);
int orc_oct_neighbors(int W, int H, int D, const uint8_t* l0, const uint8_t* l1, const uint8_t* l2, const uint8_t* l3,
                      int L, int x, int y, int z, int32_t* out);
void orc_oct_divergence(int W, int H, int D, const float* u, const float* v, const float* w, float* d0, float* d1,
                        float* d2, float* d3);
float orc_oct_closed_mean(int W, int H, int D, const uint8_t* occ, const float* u, const float* v, const float* w);
void orc_oct_solve(int W, int H, int D, const uint8_t* l0, const uint8_t* l1, const uint8_t* l2, const uint8_t* l3,
                   const float* d0, const float* d1, const float* d2, const float* d3, int iters, int closed,
                   float mean, float* p0, float* p1, float* p2, float* p3);
void orc_oct_adjust(int W, int H, int D, const uint8_t* l0, const uint8_t* l1, const uint8_t* l2, const uint8_t* l3,
                    const float* p0, const float* p1, const float* p2, const float* p3, float* u, float* v, float* w);
/* The special-region projections of octree guiding (ZQ_SmokeOctreeGrid3D::ApplyGuidingVelocity with
 * guiding_specialRegion_useCUDA, the reference's build: ZQ_SmokeOctreeGrid3D.h:37, .cpp:546-720).
 * Regions are labelled over the fine cells in level-0 or level-1 leaves (specialRegion_mask,
 * :5563-5591) exactly as _find_special_label labels them (:5169-5320): a pass seeds the first
 * unlabelled cell of the lowest plane that has one and the first unlabelled cell of row 0 of every
 * plane above, all with the pass's label, then fills from the last seed, labelling the cells it
 * reaches along y and x but not those it reaches only along z, which seed later passes.  Each is
 * projected alone, in order, by the closed red-black solver on its bounding box (orc_project, omega 1,
 * open_mask 0): the box's cells outside the region occupied, the region's mean divergence (float,
 * summed in scan order as u1 - u0 + v1 - v0 + w1 - w0) subtracted, min(iters, cells) iterations; the
 * box's faces are written back except the u faces of its high x side (:700-712).  The guide's blended
 * deltas (:265-545) are not applied on this path of the reference (only its PCG branch adds them), so
 * no guide is read.  labels (optional, W*H*D): 0 outside the regions, else the region number.
 * Returns the number of regions, or -1 where the reference's fill does not end (two cells above each
 * other reached only along z queue each other until its W*H*D queue overflows). */
int orc_oct_special_regions(int W, int H, int D, const uint8_t* l0, const uint8_t* l1, int iters, float* u, float* v,
                            float* w, int32_t* labels);

#ifdef __cplusplus
}
#endif
#endif
