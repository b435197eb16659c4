// zq3d_kernels.h -- host-side launchers of the gfx950 kernels (device pointers, stream-ordered).
// Every launcher checks nothing itself: callers validate shapes first (zq3d_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zq3d_device.h"

namespace zq3d {

struct ForceParams {
    float h, dt, beta, alpha, eps, Tamb;
};

// Launchers templated on R are instantiated for float (the product path) and double (the FP64
// check mode of the adjoint); parameters stay float, as the reference passes them.

// ---- force + attenuation (ZQ_CUDA_AddForce3D.cu, ZQ_CUDA_Attenuation3D.cu) ----
// vort: N 4-wide scratch, fcell: 3N scratch.  extra (nullable): 3 values per cell.
// atten != 0 fuses the attenuation stage (faces *= aV on the same mask, T *= aT, rho *= aR).
// vort_planes: the vorticity records go to vort as four planes of N values (wx, wy, wz, |w|) instead
// of one 4-wide record per cell (same bytes, same results; the force stage's |w| gradient then reads
// dense neighbours).  Callers that read vort afterwards as records leave it false.
template <class R>
void launch_force_atten(const Dims& d, R* u, R* v, R* w, R* T, R* rho, const uint8_t* occ, const R* extra,
                        const ForceParams& fp, int atten, float aV, float aT, float aR, vec4_t<R>* vort, R* fcell,
                        hipStream_t s, bool vort_planes = false);
// (wx, wy, wz, |w|) per cell, zero on the border cells
template <class R>
void launch_vorticity(const Dims& d, const R* u, const R* v, const R* w, vec4_t<R>* vort, float h, hipStream_t s);
void launch_attenuation(const Dims& d, float* u, float* v, float* w, float* T, float* rho, const uint8_t* occ,
                        float aV, float aT, float aR, hipStream_t s);

// ---- MAC <-> cell (ZQ_CUDA_MACtoRegular.cu) ----
// with occ, the fourth component holds each cell's occupancy (the grids the forward backtrace samples)
template <class R>
void launch_mac_to_cell4(const Dims& d, const R* u, const R* v, const R* w, vec4_t<R>* vel4, hipStream_t s,
                         const uint8_t* occ = nullptr);

// ---- advection (ZQ_CUDA_Advection3D.cu) ----
struct AdvParams {
    float h, dtt;
    int steps;
};
// mode: AdvectVelocityType (0 inReg_outReg, 1 its BFECC, 2 inMAC_outMAC, 3 its BFECC, 4 inReg_outMAC,
// 5 its BFECC).  scratch: vel4 (N), faces (Nu+Nv+Nw values); bfecc: bfecc_scratch_values(d) values,
// needed by the BFECC modes only (nullable otherwise).
size_t bfecc_scratch_values(const Dims& d);
template <class R>
void launch_advect_velocity(const Dims& d, R* u, R* v, R* w, const uint8_t* occ, int is_open, int mode,
                            const AdvParams& ap, vec4_t<R>* vel4, R* scratch_faces, hipStream_t s,
                            R* bfecc = nullptr);
// mac != 0: MAC velocity sampling (ADV_SCA_MAC); else cell-centred (ADV_SCA_REG).
template <class R>
void launch_advect_scalar(const Dims& d, const R* u, const R* v, const R* w, const uint8_t* occ, const R* Tin,
                          const R* Rin, R* Tout, R* Rout, int mac, const AdvParams& ap, vec4_t<R>* vel4,
                          hipStream_t s);

// ---- projection ----
// Doubles needed by the reduction partials of a grid (per-block sums / maxima).
int reduction_partials(const Dims& d);
// Neighbour counts, FP32 divergence b (or FP64 b64), zero p at uncounted cells; when part is given,
// per-block FP64 sum/count of the divergence over fluid cells (closed-domain mean).
void launch_proj_setup(const Dims& d, const float* u, const float* v, const float* w, const uint8_t* occ,
                       int open_mask, uint8_t* nc, float* b, double* b64, float* p_pad, double* p64_pad,
                       double* part, hipStream_t s);
// FP64 variant on FP64 faces (always b64/p64); free_only != 0 sums only the faces the correction
// updates (between two fluid cells, and open domain faces): the right-hand side of the adjoint.
void launch_proj_setup64(const Dims& d, const double* u, const double* v, const double* w, const uint8_t* occ,
                         int open_mask, int free_only, uint8_t* nc, double* b64, double* p64_pad, double* part,
                         hipStream_t s);
// FP32-face variant of the adjoint right-hand side (free faces only) into b64
void launch_proj_setup_free(const Dims& d, const float* u, const float* v, const float* w, const uint8_t* occ,
                            int open_mask, uint8_t* nc, double* b64, double* p64_pad, double* part, hipStream_t s);
// mean32/mean64 = sum/count of the setup partials (0 when want == 0)
void launch_mean_finalize(const Dims& d, const double* part, float* mean32, double* mean64, int want, hipStream_t s);
// the same fold over the mixed-precision set-up's partials (launch_proj_setup with b64)
void launch_mean_finalize_mp(const Dims& d, const double* part, double* mean64, int want, hipStream_t s);
// Sweep algorithm / tile shape: ty rows per block (2, 4 or 8) and zc planes per chunk of the fused
// one-pass kernel on interleaved storage; ty == SOR_SPLIT selects single-colour sweeps on
// colour-split storage; ty == -r (r = 4, 8 or 16) selects one launch per red-black iteration on
// colour-split storage with r rows by zc planes per workgroup; {0, 0} selects the built-in default.
constexpr int SOR_SPLIT = -1;
constexpr int SOR_TWO = 100;  // ty == -(SOR_TWO + r), r = 16 or 32: two iterations per launch on split storage
struct SorShape {
    int ty, zc;
};
inline bool is_split(SorShape s) { return s.ty <= SOR_SPLIT; }
// the one-launch split shapes, whose first iteration can take the source pair as zeros unread
inline bool split_zero_start(SorShape s) { return s.ty < SOR_SPLIT && s.ty > -SOR_TWO; }
// whether the one-launch split kernels (buffer access with 32-bit offsets) can run on the grid: every
// colour-split array under 2 GiB (cubes up to about 1000 cells on a side)
bool split_buffers_fit(const Dims& d);
// Colour-split scratch: per colour a padded pressure (split_padded_size floats), a second padded
// pressure the one-launch iteration ping-pongs with, right-hand side and counts (split_cell_size
// each).  Padding slots and the slots past a row's end must hold 0 in p and alt.
struct SplitBufs {
    float* p[2];
    float* b[2];
    uint8_t* n[2];
    float* alt[2];
};
// Gated passes (zq3d_sor_opts.pass_rule 2): per-iteration largest-update words, the executed
// iteration count, and the mapped host word the count is copied to.
constexpr int GATE_SLOTS = 8192;  // iterations per gated pass at most
constexpr int GATE_K = 16;        // a gated pass also ends when the updates stop shrinking over GATE_K iterations
constexpr int GATE_LANES = 8;     // words per iteration (their maximum is the iteration's largest update)
struct GateBufs {
    unsigned* gate;               // GATE_SLOTS x GATE_LANES words
    int* done;
    unsigned long long* host_done;  // device address of a mapped host word
};
// warm_start 2: p64 += beta (p64 - prev), prev = the old p64 (valid == 0: prev = p64 only), n padded cells
void launch_extrapolate_p64(double* p64_pad, double* prev_pad, size_t n, int valid, float beta, hipStream_t s);
// zero gate[0..n) and *done
void launch_gate_reset(const GateBufs& g, int n, hipStream_t s);
// iterations it0 .. it0+count-1 of a gated pass on split storage: iteration it reads pair (it & 1 ?
// alt : p), writes the other, and returns at once once an iteration's largest update was <= thresh
// (FP32 bits); iteration 0 takes the source as zero
void launch_gated_sweeps(const Dims& d, const SplitBufs& sp, int it0, int count, const float* mean32, float omega,
                         SorShape shape, const GateBufs& g, unsigned thresh, hipStream_t s);
// p64 += e from the pair the gated pass ended in
void launch_update_p64_gated(const Dims& d, const uint8_t* nc, const SplitBufs& sp, const int* done, double* p64_pad,
                             hipStream_t s);
size_t split_padded_size(const Dims& d);
size_t split_cell_size(const Dims& d);
// iters red-black SOR iterations starting from p_pad.  With p_alt (same shape, ghost layer zero)
// the fused one-pass kernel ping-pongs between the two; the split algorithm converts to split
// storage, sweeps, and converts back into p_pad.  Returns the buffer holding the result.
float* launch_sor(const Dims& d, const uint8_t* nc, const float* b, const float* mean32, float* p_pad, float* p_alt,
                  float omega, int iters, hipStream_t s, SorShape shape = SorShape{0, 0},
                  const SplitBufs* split = nullptr, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
// iters red-black iterations on the split arrays as they stand (no layout conversion), starting
// from the pressure pair `from` (0: sp.p, 1: sp.alt) with the split algorithm `shape`; returns the
// pair holding the result.  Events (nullable) bracket the sweeps.  zero_start (split_zero_start
// shapes only): the start is all zeros and pair `from` is not read, so it need not hold zeros.
int launch_split_sweeps(const Dims& d, const SplitBufs& sp, int from, const float* mean32, float omega, int iters,
                        SorShape shape, hipStream_t s, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr,
                        bool zero_start = false);
// split counts n_0 / n_1 from the interleaved counts
void launch_split_counts(const Dims& d, const uint8_t* nc, const SplitBufs& sp, hipStream_t s);
// name of the sweep kernel a shape runs
const char* sor_kernel_name(const Dims& d, SorShape shape);
// sweep-kernel launches of `iters` SOR iterations
long long sor_launches(const Dims& d, SorShape shape, int iters);
// fused-kernel shapes valid for this grid (LDS ring within 64 KiB); returns the count
int sor_shape_candidates(const Dims& d, SorShape* out, int max);
void launch_correct(const Dims& d, const uint8_t* occ, int open_mask, const float* p_pad, float* u, float* v,
                    float* w, hipStream_t s);
template <class F>
void launch_correct64(const Dims& d, const uint8_t* occ, int open_mask, const double* p64_pad, F* u, F* v, F* w,
                      hipStream_t s);
// u -= p_r - p_l on every face (the adjoint of the projection: the faces the forward correction
// leaves alone carry the pressure's sensitivity too)
template <class F>
void launch_correct_all64(const Dims& d, const double* p64_pad, F* u, F* v, F* w, hipStream_t s);
// r = (b64 - mean64) - (sum_nb p - n p); writes r32 and max|r| (bits of a non-negative double) to *rmax
void launch_residual64(const Dims& d, const uint8_t* nc, const double* b64, const double* mean64,
                       const double* p64_pad, float* r32, double* part, unsigned long long* rmax, hipStream_t s);
// launch_residual64 writing r32 into the colour-split right-hand sides sp.b[0] / sp.b[1]
void launch_residual64_split(const Dims& d, const uint8_t* nc, const double* b64, const double* mean64,
                             const double* p64_pad, const SplitBufs& sp, double* part, unsigned long long* rmax,
                             hipStream_t s);
// p64 += e, e = 0 on counted cells
void launch_update_p64(const Dims& d, const uint8_t* nc, float* e_pad, double* p64_pad, hipStream_t s);
// the same with e in a colour-split pressure pair e[0] / e[1]; zero_e false leaves e as it is (the
// next pass starts with a zero_start sweep)
void launch_update_p64_split(const Dims& d, const uint8_t* nc, float* const* e, double* p64_pad, hipStream_t s,
                             bool zero_e = true);
// max |div64 - mean64| over counted cells into *dmax (bits of a non-negative double)
template <class F>
void launch_max_div64(const Dims& d, const F* u, const F* v, const F* w, const uint8_t* nc, const double* mean64,
                      double* part, unsigned long long* dmax, hipStream_t s);
// max |div64 - mean64| over counted cells of the faces launch_correct64 would produce from p64_pad
// (equal bit for bit to launch_max_div64 after launch_correct64), faces left unchanged
template <class F>
void launch_max_div64_pred(const Dims& d, const F* u, const F* v, const F* w, const uint8_t* occ, int open_mask,
                           const uint8_t* nc, const double* p64_pad, const double* mean64, double* part,
                           unsigned long long* dmax, hipStream_t s);
// max |b64| over counted cells into *out (bits of a non-negative double)
void launch_max_abs64(const Dims& d, const uint8_t* nc, const double* b64, double* part, unsigned long long* out,
                      hipStream_t s);
// mean of p64 over fluid cells, subtracted from those cells (fixes the free constant of a
// closed-domain pressure); part holds the partials, mean64 receives the mean
void launch_remove_mean_p64(const Dims& d, const uint8_t* occ, double* p64_pad, double* part, double* mean64,
                            hipStream_t s);

// ---- sources, obstacles, fills ----
template <class R>
void launch_reinject_box(const Dims& d, const float* box8, R* T, R* rho, hipStream_t s);
template <class R>
void launch_reinject_sphere(const Dims& d, const float* sph6, R* T, R* rho, hipStream_t s);
void launch_stamp_sphere(const Dims& d, float cx, float cy, float cz, float r, float su, float sv, float sw,
                         int shell_mask, uint8_t* occ, float* u, float* v, float* w, hipStream_t s);
void launch_fill_f32(float* p, size_t n, float val, hipStream_t s);
void launch_fill_f64(double* p, size_t n, double val, hipStream_t s);
void launch_fill_u64(unsigned long long* p, size_t n, unsigned long long val, hipStream_t s);
void launch_fill_u8(uint8_t* p, size_t n, uint8_t val, hipStream_t s);

}  // namespace zq3d
