// zq3d_internal.h -- pieces of the C-ABI implementation shared by zq3d_api.cpp (Parts 1-3) and
// zq3d_adjoint.cpp (Part 5): error reporting, sizes, projection drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "zq3d_coarse.h"
#include "zq3d_kernels.h"
#include "zq3d_octree.h"
#include "zq_smoke3d.h"

namespace zq3d {
namespace detail {

extern thread_local std::string g_err;
void set_err(const char* where, const char* what);
int check_dims(int W, int H, int D, const char* where);
int check_sor_opts(const zq3d_sor_opts& o, const char* where);

#define ZQ_HIP(call, where)                                                   \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            zq3d::detail::set_err(where, hipGetErrorString(e_));              \
            return ZQ3D_ERR_HIP;                                              \
        }                                                                     \
    } while (0)

struct Sizes {
    size_t N, Nu, Nv, Nw, Np;
    explicit Sizes(const Dims& d)
        : N((size_t)d.W * d.H * d.D),
          Nu((size_t)(d.W + 1) * d.H * d.D),
          Nv((size_t)d.W * (d.H + 1) * d.D),
          Nw((size_t)d.W * d.H * (d.D + 1)),
          Np((size_t)(d.W + 2) * (d.H + 2) * (d.D + 2)) {}
};

// Small per-call scalars in one device block.
struct Scalars {
    double mean64;
    double zero64;
    double aux64;   // the removed pressure mean of an adjoint projection
    unsigned long long rmax;
    unsigned long long dmax;
    unsigned long long done_sink;  // stands in for the mapped host word when there is none
    float mean32;
    float zero32;
    int done;                 // iterations a gated pass ran
    unsigned gate[GATE_SLOTS * GATE_LANES];  // per iteration of a gated pass: largest update (FP32 bits)
};

// Collects the first HIP error of a sequence of calls.
struct Errs {
    hipError_t e = hipSuccess;
    void operator()(hipError_t r) {
        if (e == hipSuccess) e = r;
    }
};

// Event timing of the red-black sweep loops (null when profiling is off).  Every sweep loop is
// bracketed by its own event pair from a pool; nothing waits for the events until resolve(), so
// timing adds no synchronisation to the stream.
struct SorProf {
    std::vector<hipEvent_t> ev;  // pairs (2i, 2i + 1)
    int used = 0;                // pairs recorded since the last resolve
    double ms = 0.0;
    long long launches = 0;
    static constexpr int kPool = 4096;
    // the next pair (created on first use); a full pool is resolved first
    hipError_t next(hipEvent_t* a, hipEvent_t* b) {
        if (used == kPool)
            if (hipError_t e = resolve()) return e;
        if ((int)ev.size() < 2 * (used + 1)) {
            for (int k = 0; k < 2; k++) {
                hipEvent_t x = nullptr;
                if (hipError_t e = hipEventCreate(&x)) return e;
                ev.push_back(x);
            }
        }
        *a = ev[2 * used];
        *b = ev[2 * used + 1];
        used++;
        return hipSuccess;
    }
    // adds the recorded pairs' elapsed times to ms (waits for the last pair)
    hipError_t resolve() {
        if (used == 0) return hipSuccess;
        if (hipError_t e = hipEventSynchronize(ev[2 * used - 1])) return e;
        for (int i = 0; i < used; i++) {
            float t = 0.0f;
            if (hipError_t e = hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1])) return e;
            ms += t;
        }
        used = 0;
        return hipSuccess;
    }
    void release() {
        for (hipEvent_t x : ev) (void)hipEventDestroy(x);
        ev.clear();
        used = 0;
    }
};

// Pinned, device-visible host slots the reduction kernels write the scalars the host reads back
// (slot 0: max|r| of a refinement pass, slot 1: a max|div| or max|rhs|), read after a stream
// synchronisation with no copy; null when the owner has none (then the device scalars in Scalars
// are copied back).
struct Readback {
    unsigned long long* host = nullptr;
    unsigned long long* dev = nullptr;
    hipError_t alloc() {
        if (hipError_t e = hipHostMalloc((void**)&host, 4 * sizeof(unsigned long long), hipHostMallocMapped |
                                         hipHostMallocCoherent))
            return e;
        return hipHostGetDevicePointer((void**)&dev, host, 0);
    }
    void release() {
        if (host) (void)hipHostFree(host);
        host = dev = nullptr;
    }
};

// Device buffers of a projection.  p/p_alt: padded FP32 pressure (plain) or correction e (mixed),
// ping-ponged by the fused sweep kernel; both keep a zero ghost layer.  part: reduction partials.
struct ProjBufs {
    uint8_t* nc;
    float* b;       // FP32 divergence (plain) / FP32 residual (mixed)
    float* p;
    float* p_alt;
    double* b64;
    double* p64;
    double* part;
    Scalars* sc;
    SorShape shape;  // sweep algorithm / tile shape ({0, 0}: default)
    SplitBufs split; // colour-split scratch (used when shape.ty == SOR_SPLIT)
    Readback rb{};   // host read-back slots (optional)
    hipEvent_t gate_ev[2] = {nullptr, nullptr};  // progress checks of a gated pass (optional)
    // where a reduction result of `slot` goes: the pinned slot, or the device scalar
    unsigned long long* target(int slot, unsigned long long* dev_scalar) const {
        return rb.dev ? rb.dev + slot : dev_scalar;
    }
    // the gate words of a gated pass; the executed count goes to pinned slot 2 when there is one
    GateBufs gate() const { return GateBufs{sc->gate, &sc->done, rb.dev ? rb.dev + 2 : &sc->done_sink}; }
    // split storage present (the gated passes need it)
    bool has_split() const { return split.p[0] && split.p[1] && split.alt[0] && split.alt[1] && split.b[0]; }
};

// the value a reduction left in `slot` (pinned) or in dev_scalar (copied back); waits for the stream
int read_result(const ProjBufs& B, int slot, const unsigned long long* dev_scalar, double* out, hipStream_t s,
                const char* where);

// launch_sor, timed when pf is set (the sweeps are the only work between the two events)
hipError_t timed_sor(const Dims& d, const uint8_t* nc, const float* b, const float* mean32, float* p, float* p_alt,
                     float omega, int iters, hipStream_t s, SorProf* pf, float** result, SorShape shape,
                     const SplitBufs* split);
int read_double_bits(const unsigned long long* dev, double* out, hipStream_t s, const char* where);
// Divergence check of a forward projection's refinement: the faces (float, or double when fp64) the
// correction will update in place, with its occupancy and open faces.
struct DivCheck {
    const void* u;
    const void* v;
    const void* w;
    const uint8_t* occ;
    int open_mask;
    bool fp64;
};
// Mixed-precision refinement of B.p64: FP64 residual, FP32 SOR on the correction e (zero start),
// p64 += e, until max|r| <= tol or o.max_outer passes (pass lengths: fitted_pass_length).  With chk,
// a residual within tol is not yet the end: the corrected faces' max|div - mean| is predicted (the
// correction rounds every face to the face type once more) and the passes go on while it exceeds
// tol and each pass still shrinks it below 0.9 of the previous prediction (past that, the faces'
// rounding dominates); out->dmax receives that value, or -1 when the last pass left it unevaluated.
// With the colour-split sweep algorithm the residual and e stay in split storage across passes
// (counts split once per call).
// Largest-update threshold of a gated pass: an iteration whose updates all satisfy |new - old| <=
// this bound ends the pass.  The update of a cell with n neighbours is omega * r / n, r its
// residual at that moment, so the bound asks |r| <= GATE_THETA * tol at every cell; measured on the
// plume at 128^3 the FP64 residual after such an iteration was 0.6-1.3 times 6 max|update| / omega.
constexpr double GATE_THETA = 0.75;
float gate_threshold(double tol, float omega);
struct RefineOut {
    int outer = 0;       // passes run
    long long sweeps = 0;  // red-black iterations run over all passes
    double rmax = 0.0;   // last max|r|
    double dmax = -1.0;  // last predicted max|div - mean| (-1: not evaluated after the last pass)
};
int refine_mixed(const Dims& d, const zq3d_sor_opts& o, double tol, const ProjBufs& B, hipStream_t s, SorProf* pf,
                 const char* where, const DivCheck* chk, RefineOut* out);
// Sweeps of the next refinement pass (zq3d_sor_opts.pass_rule): r is the max|r| the last pass left,
// rprev the one before it and nprev that pass's sweeps.  The same function in the same operations
// as the oracle's, so both choose the same lengths from the same residuals.
int fitted_pass_length(int iters, int fit, double r, double rprev, int nprev, double tol);

// Plain FP32 projection on device pointers: p = 0, setup, SOR, correct.  mean_mode: 0 = value
// `mean` given, 1 = FP64 device mean over fluid cells.
int project_plain(const Dims& d, float* u, float* v, float* w, const uint8_t* occ, const zq3d_sor_opts& o,
                  float mean, int mean_mode, const ProjBufs& B, zq3d_proj_stats* st, hipStream_t s,
                  SorProf* pf = nullptr);
// Mixed-precision projection: FP64 b, p and residual; FP32 inner SOR on the correction e.
// p64 holds the warm start (zeroed by the caller for a cold start).
int project_mixed(const Dims& d, float* u, float* v, float* w, const uint8_t* occ, const zq3d_sor_opts& o,
                  const ProjBufs& B, zq3d_proj_stats* st, hipStream_t s, SorProf* pf = nullptr);

// Coarse grid of the coarse + sub-grid projection (one device allocation).
struct CoarseState {
    CoarseDims cd{};
    CoarseBufs B{};
    CoarseSweep sweep{0, 0};  // schedule of the coarse SOR (timed at creation)
    int flux = 0;             // the flux form's diagonal (zq3d_config.coarse_flux)
    int guide = 0;            // guide faces set (guiding control)
    float guide_coeff = 0.0f;
    void* mem = nullptr;
};
int coarse_create(const Dims& d, int s, CoarseState* out, const char* where);
void coarse_destroy(CoarseState* c);
// host guide faces of the coarse grid (gu == NULL: guiding off)
int coarse_set_guide(CoarseState* c, float coeff, const float* gu, const float* gv, const float* gw, hipStream_t s,
                     const char* where);
// Coarse + sub-grid projection of the faces in place (o.iters coarse iterations with o.omega and
// o.open_mask).  B provides the stats scratch (nc, b, p, part, sc); stats may be null.
int project_coarse(const CoarseState& c, float* u, float* v, float* w, const uint8_t* occ, const zq3d_sor_opts& o,
                   const ProjBufs& B, zq3d_proj_stats* st, hipStream_t s, SorProf* pf);

// Octree of the octree projection (one device allocation).
struct OctreeState {
    Dims d{};
    int thresh[3] = {2, 4, 16};  // octree_thresh1..3
    OctreeBufs B{};
    OctreeLists lists;
    void* mem = nullptr;
    bool guiding = false;  // octree guiding: the special-region projections after the face correction
};
int octree_create(const Dims& d, const int thresh[3], OctreeState* out, const char* where);
void octree_destroy(OctreeState* o);
// Octree projection of the faces in place: the leaves from occ, o.iters red-black passes over the
// levels, the face correction, with t.guiding the special-region projections (octree_special_regions),
// the leaf sub-projection; o.closed_mean subtracts the mean divergence per fine cell times each
// leaf's volume.  B provides the mean and stats scratch; stats may be null.  Synchronises the stream
// twice (the cell lists' lengths and entry totals size the sweep launches), more with guiding.
// Octree guiding (ZQ_SmokeOctreeGrid3D::ApplyGuidingVelocity, the reference's build with
// guiding_specialRegion_useCUDA, ZQ_SmokeOctreeGrid3D.h:37, .cpp:546-720): the 6-connected regions of
// fine cells in level-0 and level-1 leaves, numbered in scan order, each projected alone by the closed
// red-black solver (SolveClosedPoissonRedBlackwithOccupy3D_MAC) on its bounding box with the region's
// mean divergence and min(iters, cells) iterations.  On that path the reference computes the guide's
// blended deltas but never adds them (only its PCG branch does), so no guide values are read.  The
// labelling and boxes are host work on the downloaded level map and faces, as in the reference.
int project_octree(OctreeState& t, float* u, float* v, float* w, const uint8_t* occ, const zq3d_sor_opts& o,
                   const ProjBufs& B, zq3d_proj_stats* st, hipStream_t s, SorProf* pf);

}  // namespace detail
}  // namespace zq3d
