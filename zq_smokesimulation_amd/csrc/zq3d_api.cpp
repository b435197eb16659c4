// zq3d_api.cpp -- C ABI of the gfx950 smoke step (include/zq_smoke3d.h).
//
// Part 1/2 keep the reference's calling convention (host pointers, synchronous, elapsed ms) but
// reuse a grow-only device workspace instead of the reference's cudaMalloc/cudaFree per call
// (e.g. ZQ_CUDA_AddForce3D.cu:616-651).  Part 3 keeps all fields resident in HBM.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "zq3d_device.h"
#include "zq3d_internal.h"
#include "zq3d_kernels.h"
#include "zq_smoke3d.h"

using namespace zq3d;
using namespace zq3d::detail;

namespace zq3d {
namespace detail {

thread_local std::string g_err;

void set_err(const char* where, const char* what) {
    g_err = std::string(where) + ": " + what;
}

constexpr long long kMaxCells = 1LL << 30;

int check_dims(int W, int H, int D, const char* where) {
    if (W < 2 || H < 2 || D < 2) {
        set_err(where, "grid dimensions must be >= 2");
        return ZQ3D_ERR_ARG;
    }
    const long long pad = (long long)(W + 2) * (H + 2) * (D + 2);
    if (pad >= kMaxCells || (long long)(W + 1) * (H + 1) * (D + 1) >= kMaxCells) {
        set_err(where, "grid too large for 32-bit indexing");
        return ZQ3D_ERR_ARG;
    }
    return ZQ3D_OK;
}

}  // namespace detail
}  // namespace zq3d

namespace {

// ---------------- grow-only workspace for the host-pointer entry points ----------------
enum Slot { S_U, S_V, S_W, S_OCC, S_T, S_R, S_T2, S_R2, S_VEL4, S_SCR, S_NC, S_B, S_P, S_P2, S_B64, S_P64, S_PART, S_SCAL,
            S_BFECC, S_SPLIT, S_COUNT };

struct Workspace {
    void* ptr[S_COUNT] = {};
    size_t cap[S_COUNT] = {};
    std::mutex mu;
    hipError_t ensure(int slot, size_t bytes) {
        if (cap[slot] >= bytes) return hipSuccess;
        if (ptr[slot]) (void)hipFree(ptr[slot]);
        ptr[slot] = nullptr;
        cap[slot] = 0;
        hipError_t e = hipMalloc(&ptr[slot], bytes);
        if (e == hipSuccess) cap[slot] = bytes;
        return e;
    }
    void release() {
        for (int i = 0; i < S_COUNT; i++) {
            if (ptr[i]) (void)hipFree(ptr[i]);
            ptr[i] = nullptr;
            cap[i] = 0;
        }
    }
};

Workspace& ws() {
    static Workspace w;
    return w;
}

// ZQ_Cuda_Prepare_Advection state (the reference keeps it in module globals + __constant__,
// ZQ_CUDA_Advection3D.cu:22-34,1652-1668)
struct AdvState {
    unsigned W = 0, H = 0, D = 0, steps = 0;
    float h = 0.0f, dtt = 0.0f;
};
AdvState g_adv;

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer() {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, 0);
    }
    float stop() {
        (void)hipEventRecord(b, 0);
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    ~Timer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

}  // namespace

namespace zq3d {
namespace detail {

namespace {
hipError_t account_sweeps(const Dims& d, SorShape shape, int iters, SorProf* pf) {
    pf->launches += sor_launches(d, shape, iters);
    return hipGetLastError();
}
}  // namespace

// launch_sor, timed when pf is set (the sweeps are the only work between the two events)
hipError_t timed_sor(const Dims& d, const uint8_t* nc, const float* b, const float* mean32, float* p, float* p_alt,
                     float omega, int iters, hipStream_t s, SorProf* pf, float** result, SorShape shape,
                     const SplitBufs* split) {
    if (!pf) {
        *result = launch_sor(d, nc, b, mean32, p, p_alt, omega, iters, s, shape, split);
        return hipGetLastError();
    }
    // the events bracket the sweep launches only (not the split-layout conversions)
    hipEvent_t ea = nullptr, eb = nullptr;
    if (hipError_t e = pf->next(&ea, &eb)) return e;
    *result = launch_sor(d, nc, b, mean32, p, p_alt, omega, iters, s, shape, split, ea, eb);
    return account_sweeps(d, shape, iters, pf);
}

float gate_threshold(double tol, float omega) { return (float)(GATE_THETA * tol * (double)omega / 6.0); }

namespace {

// launches per progress check of a gated pass: the iterations already queued when the check sees
// the pass end return at once (a few microseconds each)
constexpr int GATE_CHUNK = 8;

unsigned float_bits(float x) {
    unsigned u;
    std::memcpy(&u, &x, 4);
    return u;
}

// One gated pass on split storage: up to o.iters iterations from a zero correction, queued in
// chunks; with pinned read-back and events the host checks after each chunk (one chunk behind, so
// the device never waits) whether an iteration returned at once, and stops queueing.  Then p64 += e.
// The executed count is in B.sc->done (and pinned slot 2) once the stream passes the update.
int gated_pass(const Dims& d, const zq3d_sor_opts& o, double tol, const ProjBufs& B, hipStream_t s, SorProf* pf,
               const char* where) {
    const GateBufs g = B.gate();
    const unsigned thresh = float_bits(gate_threshold(tol, o.omega));
    const int n = o.iters;
    launch_gate_reset(g, n, s);
    hipEvent_t ea = nullptr, eb = nullptr;
    if (pf) ZQ_HIP(pf->next(&ea, &eb), where);
    if (ea) ZQ_HIP(hipEventRecord(ea, s), where);
    const bool poll = B.rb.host && B.gate_ev[0] && B.gate_ev[1];
    int launched = 0;
    for (int k = 0; launched < n; k++) {
        const int m = std::min(GATE_CHUNK, n - launched);
        launch_gated_sweeps(d, B.split, launched, m, &B.sc->zero32, o.omega, B.shape, g, thresh, s);
        launched += m;
        if (!poll) continue;
        ZQ_HIP(hipEventRecord(B.gate_ev[k & 1], s), where);
        if (k >= 1) {
            ZQ_HIP(hipEventSynchronize(B.gate_ev[(k - 1) & 1]), where);
            const long long ran = (long long)((volatile unsigned long long*)B.rb.host)[2];
            if (ran < launched - m) break;  // an iteration up to the previous chunk returned at once
        }
    }
    if (eb) ZQ_HIP(hipEventRecord(eb, s), where);
    ZQ_HIP(hipGetLastError(), where);
    launch_update_p64_gated(d, B.nc, B.split, g.done, B.p64, s);
    return ZQ3D_OK;
}

// iterations the last gated pass ran (the stream has passed it: called after a read_result)
int gated_count(const ProjBufs& B, hipStream_t s, const char* where, long long* ran) {
    if (B.rb.host) {
        *ran = (long long)((volatile unsigned long long*)B.rb.host)[2];
        return ZQ3D_OK;
    }
    int v = 0;
    ZQ_HIP(hipMemcpyAsync(&v, &B.sc->done, sizeof(v), hipMemcpyDeviceToHost, s), where);
    ZQ_HIP(hipStreamSynchronize(s), where);
    *ran = v;
    return ZQ3D_OK;
}

}  // namespace

int fitted_pass_length(int iters, int fit, double r, double rprev, int nprev, double tol) {
    if (!fit || !(rprev > 0.0) || !(r < rprev) || !(r > tol) || nprev < 1) return iters;
    // per-sweep decay over the last pass, rho = (r / rprev)^(1/nprev); sweeps to tol at that rate,
    // times 1.25 because the decay slows as the smooth error modes take over (on the 256^3 closed-box
    // plume 1.25 ran fastest of 1.0, 1.1, 1.25, 1.5 and 2: a short pass costs another pass's residual
    // and host wait, a long one overshoots)
    const double need = 1.25 * (double)nprev * (std::log(tol / r) / std::log(r / rprev));
    if (!(need < (double)iters)) return iters;
    const int n = (int)std::ceil(need);
    return n < 1 ? 1 : n;
}

int refine_mixed(const Dims& d, const zq3d_sor_opts& o, double tol, const ProjBufs& B, hipStream_t s, SorProf* pf,
                 const char* where, const DivCheck* chk, RefineOut* out) {
    const Sizes z(d);
    Scalars* sc = B.sc;
    const bool gated = o.pass_rule == 2;
    if (gated && !B.has_split()) {
        set_err(where, "pass_rule 2 needs the colour-split buffers");
        return ZQ3D_ERR_ARG;
    }
    if (gated && !split_buffers_fit(d)) {
        set_err(where, "pass_rule 2 needs colour-split arrays under 2 GiB (grids up to about 1000^3)");
        return ZQ3D_ERR_ARG;
    }
    const bool split = gated || is_split(B.shape);
    int pair = 0;  // split pressure pair holding the correction
    float* e = B.p;
    float* e_alt = B.p_alt;
    if (split) {
        launch_split_counts(d, B.nc, B.split, s);
        launch_fill_f32(B.split.p[0], split_padded_size(d), 0.0f, s);
        launch_fill_f32(B.split.p[1], split_padded_size(d), 0.0f, s);
    } else {
        launch_fill_f32(e, z.Np, 0.0f, s);
    }
    int outer = 0, nprev = 0;
    long long sweeps = 0;
    bool count_pending = false;  // a gated pass ran; its count is read after the next wait
    double rmax = 0.0, rprev = -1.0, dmax = -1.0, dprev = -1.0;
    for (;;) {
        if (split)
            launch_residual64_split(d, B.nc, B.b64, &sc->mean64, B.p64, B.split, B.part, B.target(0, &sc->rmax), s);
        else
            launch_residual64(d, B.nc, B.b64, &sc->mean64, B.p64, B.b, B.part, B.target(0, &sc->rmax), s);
        if (int rc = read_result(B, 0, &sc->rmax, &rmax, s, where)) return rc;
        if (count_pending) {
            long long ran = 0;
            if (int rc = gated_count(B, s, where, &ran)) return rc;
            sweeps += ran;
            if (pf) pf->launches += ran;  // one launch per gated iteration
            count_pending = false;
        }
        if (!(rmax > tol)) {
            if (!chk) break;
            // the residual is the divergence of the FP64-corrected faces; the correction rounds each
            // face to FP32 once more, so check the divergence the stored faces will have
            if (chk->fp64)
                launch_max_div64_pred(d, (const double*)chk->u, (const double*)chk->v, (const double*)chk->w,
                                      chk->occ, chk->open_mask, B.nc, B.p64, &sc->mean64, B.part,
                                      B.target(1, &sc->dmax), s);
            else
                launch_max_div64_pred(d, (const float*)chk->u, (const float*)chk->v, (const float*)chk->w,
                                      chk->occ, chk->open_mask, B.nc, B.p64, &sc->mean64, B.part,
                                      B.target(1, &sc->dmax), s);
            if (int rc = read_result(B, 1, &sc->dmax, &dmax, s, where)) return rc;
            if (!(dmax > tol)) break;
            // what is left is the faces' FP32 rounding once a pass no longer shrinks it by a tenth:
            // the bound is out of reach at this velocity scale (reported, not looped on)
            if (dprev >= 0.0 && !(dmax < 0.9 * dprev)) break;
            dprev = dmax;
        }
        if (outer >= o.max_outer) break;
        dmax = -1.0;
        if (gated) {
            if (int rc = gated_pass(d, o, tol, B, s, pf, where)) return rc;
            count_pending = true;
            outer++;
            continue;
        }
        const int n = fitted_pass_length(o.iters, o.pass_rule == 1, rmax, rprev, nprev, tol);
        if (split) {
            hipEvent_t ea = nullptr, eb = nullptr;
            if (pf) ZQ_HIP(pf->next(&ea, &eb), where);
            // the one-launch shapes start each pass from zero without reading the pair, so the update
            // need not clear it; the other split shapes read it and get it cleared
            const bool zs = split_zero_start(B.shape);
            pair = launch_split_sweeps(d, B.split, pair, &sc->zero32, o.omega, n, B.shape, s, ea, eb, zs);
            ZQ_HIP(pf ? account_sweeps(d, B.shape, n, pf) : hipGetLastError(), where);
            launch_update_p64_split(d, B.nc, pair ? B.split.alt : B.split.p, B.p64, s, !zs);
        } else {
            float* res = nullptr;
            ZQ_HIP(timed_sor(d, B.nc, B.b, &sc->zero32, e, e_alt, o.omega, n, s, pf, &res, B.shape, &B.split), where);
            if (res != e) std::swap(e, e_alt);
            launch_update_p64(d, B.nc, e, B.p64, s);  // also zeroes e for the next pass
        }
        rprev = rmax;
        nprev = n;
        sweeps += n;
        outer++;
    }
    out->outer = outer;
    out->sweeps = sweeps;
    out->rmax = rmax;
    out->dmax = dmax;
    return ZQ3D_OK;
}

int read_result(const ProjBufs& B, int slot, const unsigned long long* dev_scalar, double* out, hipStream_t s,
                const char* where) {
    if (!B.rb.host) return read_double_bits(dev_scalar, out, s, where);
    ZQ_HIP(hipStreamSynchronize(s), where);
    const unsigned long long bits = ((volatile unsigned long long*)B.rb.host)[slot];
    std::memcpy(out, &bits, sizeof(*out));
    return ZQ3D_OK;
}

int read_double_bits(const unsigned long long* dev, double* out, hipStream_t s, const char* where) {
    unsigned long long bits = 0;
    ZQ_HIP(hipMemcpyAsync(&bits, dev, sizeof(bits), hipMemcpyDeviceToHost, s), where);
    ZQ_HIP(hipStreamSynchronize(s), where);
    std::memcpy(out, &bits, sizeof(*out));
    return ZQ3D_OK;
}

// Plain FP32 projection on device pointers: p = 0, setup, SOR, correct.  mean_mode: 0 = value
// `mean` given, 1 = FP64 device mean over fluid cells.
int project_plain(const Dims& d, float* u, float* v, float* w, const uint8_t* occ, const zq3d_sor_opts& o,
                  float mean, int mean_mode, const ProjBufs& B, zq3d_proj_stats* st, hipStream_t s,
                  SorProf* pf) {
    const Sizes z(d);
    Scalars* sc = B.sc;
    launch_fill_f32(B.p, z.Np, 0.0f, s);
    launch_proj_setup(d, u, v, w, occ, o.open_mask, B.nc, B.b, nullptr, B.p, nullptr, mean_mode ? B.part : nullptr,
                      s);
    if (mean_mode) {
        launch_mean_finalize(d, B.part, &sc->mean32, &sc->mean64, 1, s);
    } else {
        launch_fill_f32(&sc->mean32, 1, mean, s);
        launch_fill_f64(&sc->mean64, 1, (double)mean, s);
    }
    float* res = nullptr;
    ZQ_HIP(timed_sor(d, B.nc, B.b, &sc->mean32, B.p, B.p_alt, o.omega, o.iters, s, pf, &res, B.shape, &B.split),
           "project");
    launch_correct(d, occ, o.open_mask, res, u, v, w, s);
    if (st) {
        launch_max_div64(d, u, v, w, B.nc, &sc->mean64, B.part, B.target(1, &sc->dmax), s);
        double md = 0.0, mean64 = 0.0;
        if (int rc = read_result(B, 1, &sc->dmax, &md, s, "project")) return rc;
        ZQ_HIP(hipMemcpyAsync(&mean64, &sc->mean64, sizeof(mean64), hipMemcpyDeviceToHost, s), "project");
        ZQ_HIP(hipStreamSynchronize(s), "project");
        st->outer = 0;
        st->sweeps = o.iters;
        st->max_residual = -1.0;
        st->max_div = md;
        st->mean = mean64;
        st->converged = -1;
    }
    return ZQ3D_OK;
}

// Mixed-precision projection: FP64 b, p and residual; FP32 inner SOR on the correction e.
// p64 holds the warm start (zeroed by the caller for a cold start).  Passes run until the faces the
// correction stores have max|div - mean| <= tol (or o.max_outer passes).
int project_mixed(const Dims& d, float* u, float* v, float* w, const uint8_t* occ, const zq3d_sor_opts& o,
                  const ProjBufs& B, zq3d_proj_stats* st, hipStream_t s, SorProf* pf) {
    Scalars* sc = B.sc;
    launch_fill_f32(&sc->zero32, 1, 0.0f, s);
    launch_proj_setup(d, u, v, w, occ, o.open_mask, B.nc, nullptr, B.b64, nullptr, B.p64, B.part, s);
    launch_mean_finalize_mp(d, B.part, &sc->mean64, o.closed_mean, s);
    const DivCheck chk{u, v, w, occ, o.open_mask, false};
    RefineOut ro;
    if (int rc = refine_mixed(d, o, o.tol, B, s, pf, "project_mixed", &chk, &ro)) return rc;
    double md = ro.dmax;
    launch_correct64(d, occ, o.open_mask, B.p64, u, v, w, s);
    if (!st) return ZQ3D_OK;
    if (md < 0.0) {  // the passes ran out before the residual met tol: measure the stored faces
        launch_max_div64(d, u, v, w, B.nc, &sc->mean64, B.part, B.target(1, &sc->dmax), s);
        if (int rc = read_result(B, 1, &sc->dmax, &md, s, "project_mixed")) return rc;
    }
    double mean64 = 0.0;
    ZQ_HIP(hipMemcpyAsync(&mean64, &sc->mean64, sizeof(mean64), hipMemcpyDeviceToHost, s), "project_mixed");
    ZQ_HIP(hipStreamSynchronize(s), "project_mixed");
    st->outer = ro.outer;
    st->sweeps = (int)ro.sweeps;
    st->max_residual = ro.rmax;
    st->max_div = md;
    st->mean = mean64;
    st->converged = md <= o.tol ? 1 : 0;
    return ZQ3D_OK;
}

int check_sor_opts(const zq3d_sor_opts& o, const char* where) {
    if (o.iters < 0 || o.max_outer < 0 || !(o.omega > 0.0f && o.omega < 2.0f) || (o.open_mask & ~63)) {
        set_err(where, "invalid SOR options (iters >= 0, 0 < omega < 2, open_mask in [0,63])");
        return ZQ3D_ERR_ARG;
    }
    if (o.pass_rule < 0 || o.pass_rule > 2 || (o.pass_rule == 2 && o.iters > GATE_SLOTS)) {
        set_err(where, "pass_rule must be 0, 1 or 2 (with 2, iters <= 8192)");
        return ZQ3D_ERR_ARG;
    }
    if (o.mixed_precision && (o.max_outer < 1 || o.iters < 1 || !(o.tol >= 0.0))) {
        set_err(where, "mixed_precision needs max_outer >= 1, iters >= 1 and tol >= 0 (with max_outer = 0 the "
                       "velocity would never be projected)");
        return ZQ3D_ERR_ARG;
    }
    return ZQ3D_OK;
}

}  // namespace detail
}  // namespace zq3d

// =============================== Part 1 ===============================

extern "C" void ZQ_Cuda_Prepare_Advection(const unsigned int width, const unsigned int height,
                                          const unsigned int depth, const float voxelSize, const unsigned int steps,
                                          const float deltatt) {
    g_adv.W = width;
    g_adv.H = height;
    g_adv.D = depth;
    g_adv.h = voxelSize;
    g_adv.steps = steps;
    g_adv.dtt = deltatt;
}

namespace {
int adv_ready(const char* where) {
    if (int rc = check_dims((int)g_adv.W, (int)g_adv.H, (int)g_adv.D, where)) return rc;
    if (g_adv.steps < 1) {
        set_err(where, "steps must be >= 1 (call ZQ_Cuda_Prepare_Advection first)");
        return ZQ3D_ERR_ARG;
    }
    return ZQ3D_OK;
}
}  // namespace

extern "C" float Advect_Velocity(float* u, float* v, float* w, const bool* occupy, bool is_open,
                                 enum ZQ_CUDA_Advection3D::AdvectVelocityType type) {
    const char* where = "Advect_Velocity";
    if (adv_ready(where)) return -1.0f;
    if (!u || !v || !w || !occupy) { set_err(where, "null pointer"); return -1.0f; }
    const int mode = (int)type;
    if (mode < 0 || mode > 5) {
        set_err(where, "unknown advection mode");
        return -1.0f;
    }
    const bool bfecc = mode == 1 || mode == 3 || mode == 5;
    const Dims d{(int)g_adv.W, (int)g_adv.H, (int)g_adv.D};
    const Sizes z(d);
    Workspace& W = ws();
    std::lock_guard<std::mutex> lk(W.mu);
    Timer t;
    Errs E;
    if (W.ensure(S_U, z.Nu * 4) || W.ensure(S_V, z.Nv * 4) || W.ensure(S_W, z.Nw * 4) || W.ensure(S_OCC, z.N) ||
        W.ensure(S_VEL4, z.N * 16) || W.ensure(S_SCR, (z.Nu + z.Nv + z.Nw) * 4) ||
        (bfecc && W.ensure(S_BFECC, bfecc_scratch_values(d) * 4))) {
        set_err(where, "device allocation failed");
        return -1.0f;
    }
    float* du = (float*)W.ptr[S_U];
    float* dv = (float*)W.ptr[S_V];
    float* dw = (float*)W.ptr[S_W];
    uint8_t* docc = (uint8_t*)W.ptr[S_OCC];
    E(hipMemcpy(du, u, z.Nu * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(dv, v, z.Nv * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(dw, w, z.Nw * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(docc, occupy, z.N, hipMemcpyHostToDevice));
    const AdvParams ap{g_adv.h, g_adv.dtt, (int)g_adv.steps};
    launch_advect_velocity(d, du, dv, dw, docc, is_open ? 1 : 0, mode, ap, (float4*)W.ptr[S_VEL4],
                           (float*)W.ptr[S_SCR], 0, bfecc ? (float*)W.ptr[S_BFECC] : nullptr);
    E(hipMemcpy(u, du, z.Nu * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(v, dv, z.Nv * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(w, dw, z.Nw * 4, hipMemcpyDeviceToHost));
    E(hipGetLastError());
    const hipError_t e = E.e;
    const float ms = t.stop();
    if (e != hipSuccess) { set_err(where, hipGetErrorString(e)); return -1.0f; }
    return ms;
}

extern "C" float Advect_Scalar(const float* u, const float* v, const float* w, const bool* occupy,
                               const float* input_temperature, const float* input_density, float* output_temperature,
                               float* output_density, enum ZQ_CUDA_Advection3D::AdvectScalarType type) {
    const char* where = "Advect_Scalar";
    if (adv_ready(where)) return -1.0f;
    if (!u || !v || !w || !occupy || !input_temperature || !input_density || !output_temperature || !output_density) {
        set_err(where, "null pointer");
        return -1.0f;
    }
    const int mode = (int)type;
    if (mode < 0 || mode > 3) { set_err(where, "unknown scalar advection mode"); return -1.0f; }
    const Dims d{(int)g_adv.W, (int)g_adv.H, (int)g_adv.D};
    const Sizes z(d);
    if (mode == 1 || mode == 3) {
        // The reference's switch has no BFECC case and returns its zero-initialised outputs.
        Timer t;
        std::memset(output_temperature, 0, z.N * 4);
        std::memset(output_density, 0, z.N * 4);
        return t.stop();
    }
    Workspace& W = ws();
    std::lock_guard<std::mutex> lk(W.mu);
    Timer t;
    Errs E;
    if (W.ensure(S_U, z.Nu * 4) || W.ensure(S_V, z.Nv * 4) || W.ensure(S_W, z.Nw * 4) || W.ensure(S_OCC, z.N) ||
        W.ensure(S_T, z.N * 4) || W.ensure(S_R, z.N * 4) || W.ensure(S_T2, z.N * 4) || W.ensure(S_R2, z.N * 4) ||
        W.ensure(S_VEL4, z.N * 16)) {
        set_err(where, "device allocation failed");
        return -1.0f;
    }
    E(hipMemcpy(W.ptr[S_U], u, z.Nu * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_V], v, z.Nv * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_W], w, z.Nw * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_OCC], occupy, z.N, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_T], input_temperature, z.N * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_R], input_density, z.N * 4, hipMemcpyHostToDevice));
    const AdvParams ap{g_adv.h, g_adv.dtt, (int)g_adv.steps};
    launch_advect_scalar(d, (float*)W.ptr[S_U], (float*)W.ptr[S_V], (float*)W.ptr[S_W], (uint8_t*)W.ptr[S_OCC],
                         (float*)W.ptr[S_T], (float*)W.ptr[S_R], (float*)W.ptr[S_T2], (float*)W.ptr[S_R2],
                         mode == 0 ? 1 : 0, ap, (float4*)W.ptr[S_VEL4], 0);
    E(hipMemcpy(output_temperature, W.ptr[S_T2], z.N * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(output_density, W.ptr[S_R2], z.N * 4, hipMemcpyDeviceToHost));
    E(hipGetLastError());
    const hipError_t e = E.e;
    const float ms = t.stop();
    if (e != hipSuccess) { set_err(where, hipGetErrorString(e)); return -1.0f; }
    return ms;
}

extern "C" float AddForce3D(float* mac_u, float* mac_v, float* mac_w, const bool* occupy, const float* temperature,
                            const float deltah, const float deltat, const float buoyCoeff, const float confineCoeff,
                            const float Tamb, const int width, const int height, const int depth,
                            enum ZQ_CUDA_AddForce3D::AddForceType type) {
    const char* where = "AddForce3D";
    (void)type;
    if (check_dims(width, height, depth, where)) return -1.0f;
    if (!mac_u || !mac_v || !mac_w || !occupy || !temperature) { set_err(where, "null pointer"); return -1.0f; }
    const Dims d{width, height, depth};
    const Sizes z(d);
    Workspace& W = ws();
    std::lock_guard<std::mutex> lk(W.mu);
    Timer t;
    Errs E;
    if (W.ensure(S_U, z.Nu * 4) || W.ensure(S_V, z.Nv * 4) || W.ensure(S_W, z.Nw * 4) || W.ensure(S_OCC, z.N) ||
        W.ensure(S_T, z.N * 4) || W.ensure(S_VEL4, z.N * 16) || W.ensure(S_SCR, (z.Nu + z.Nv + z.Nw) * 4)) {
        set_err(where, "device allocation failed");
        return -1.0f;
    }
    E(hipMemcpy(W.ptr[S_U], mac_u, z.Nu * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_V], mac_v, z.Nv * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_W], mac_w, z.Nw * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_OCC], occupy, z.N, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_T], temperature, z.N * 4, hipMemcpyHostToDevice));
    const ForceParams fp{deltah, deltat, buoyCoeff, 0.0f, confineCoeff, Tamb};
    launch_force_atten<float>(d, (float*)W.ptr[S_U], (float*)W.ptr[S_V], (float*)W.ptr[S_W], (float*)W.ptr[S_T],
                              nullptr, (uint8_t*)W.ptr[S_OCC], nullptr, fp, 0, 1.0f, 1.0f, 1.0f,
                              (float4*)W.ptr[S_VEL4], (float*)W.ptr[S_SCR], 0);
    E(hipMemcpy(mac_u, W.ptr[S_U], z.Nu * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_v, W.ptr[S_V], z.Nv * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_w, W.ptr[S_W], z.Nw * 4, hipMemcpyDeviceToHost));
    E(hipGetLastError());
    const hipError_t e = E.e;
    const float ms = t.stop();
    if (e != hipSuccess) { set_err(where, hipGetErrorString(e)); return -1.0f; }
    return ms;
}

extern "C" float ZQ3D_AddForce3D_ex(float* mac_u, float* mac_v, float* mac_w, const bool* occupy,
                                    const float* temperature, const float* density, const float* extra_force,
                                    const float deltah, const float deltat, const float buoyCoeff,
                                    const float gravityCoeff, const float confineCoeff, const float Tamb,
                                    const int width, const int height, const int depth) {
    const char* where = "ZQ3D_AddForce3D_ex";
    if (check_dims(width, height, depth, where)) return -1.0f;
    if (!mac_u || !mac_v || !mac_w || !occupy || !temperature || (gravityCoeff != 0.0f && !density)) {
        set_err(where, "null pointer");
        return -1.0f;
    }
    const Dims d{width, height, depth};
    const Sizes z(d);
    Workspace& W = ws();
    std::lock_guard<std::mutex> lk(W.mu);
    Timer t;
    Errs E;
    if (W.ensure(S_U, z.Nu * 4) || W.ensure(S_V, z.Nv * 4) || W.ensure(S_W, z.Nw * 4) || W.ensure(S_OCC, z.N) ||
        W.ensure(S_T, z.N * 4) || W.ensure(S_R, z.N * 4) || W.ensure(S_T2, z.N * 12) || W.ensure(S_VEL4, z.N * 16) ||
        W.ensure(S_SCR, (z.Nu + z.Nv + z.Nw) * 4)) {
        set_err(where, "device allocation failed");
        return -1.0f;
    }
    E(hipMemcpy(W.ptr[S_U], mac_u, z.Nu * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_V], mac_v, z.Nv * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_W], mac_w, z.Nw * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_OCC], occupy, z.N, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_T], temperature, z.N * 4, hipMemcpyHostToDevice));
    if (density) E(hipMemcpy(W.ptr[S_R], density, z.N * 4, hipMemcpyHostToDevice));
    if (extra_force) E(hipMemcpy(W.ptr[S_T2], extra_force, z.N * 12, hipMemcpyHostToDevice));
    const ForceParams fp{deltah, deltat, buoyCoeff, gravityCoeff, confineCoeff, Tamb};
    launch_force_atten<float>(d, (float*)W.ptr[S_U], (float*)W.ptr[S_V], (float*)W.ptr[S_W], (float*)W.ptr[S_T],
                       density ? (float*)W.ptr[S_R] : nullptr, (uint8_t*)W.ptr[S_OCC],
                       extra_force ? (float*)W.ptr[S_T2] : nullptr, fp, 0, 1.0f, 1.0f, 1.0f, (float4*)W.ptr[S_VEL4],
                       (float*)W.ptr[S_SCR], 0);
    E(hipMemcpy(mac_u, W.ptr[S_U], z.Nu * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_v, W.ptr[S_V], z.Nv * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_w, W.ptr[S_W], z.Nw * 4, hipMemcpyDeviceToHost));
    E(hipGetLastError());
    const float ms = t.stop();
    if (E.e != hipSuccess) { set_err(where, hipGetErrorString(E.e)); return -1.0f; }
    return ms;
}

extern "C" float Attenuation3D(float* mac_u, float* mac_v, float* mac_w, float* temperature, float* density,
                               const bool* occupy, const float velAtten, const float tempAtten,
                               const float densityAtten, const int width, const int height, const int depth) {
    const char* where = "Attenuation3D";
    if (check_dims(width, height, depth, where)) return -1.0f;
    if (!mac_u || !mac_v || !mac_w || !occupy || !temperature || !density) {
        set_err(where, "null pointer");
        return -1.0f;
    }
    const Dims d{width, height, depth};
    const Sizes z(d);
    Workspace& W = ws();
    std::lock_guard<std::mutex> lk(W.mu);
    Timer t;
    Errs E;
    if (W.ensure(S_U, z.Nu * 4) || W.ensure(S_V, z.Nv * 4) || W.ensure(S_W, z.Nw * 4) || W.ensure(S_OCC, z.N) ||
        W.ensure(S_T, z.N * 4) || W.ensure(S_R, z.N * 4)) {
        set_err(where, "device allocation failed");
        return -1.0f;
    }
    E(hipMemcpy(W.ptr[S_U], mac_u, z.Nu * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_V], mac_v, z.Nv * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_W], mac_w, z.Nw * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_OCC], occupy, z.N, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_T], temperature, z.N * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_R], density, z.N * 4, hipMemcpyHostToDevice));
    launch_attenuation(d, (float*)W.ptr[S_U], (float*)W.ptr[S_V], (float*)W.ptr[S_W], (float*)W.ptr[S_T],
                       (float*)W.ptr[S_R], (uint8_t*)W.ptr[S_OCC], velAtten, tempAtten, densityAtten, 0);
    E(hipMemcpy(mac_u, W.ptr[S_U], z.Nu * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_v, W.ptr[S_V], z.Nv * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_w, W.ptr[S_W], z.Nw * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(temperature, W.ptr[S_T], z.N * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(density, W.ptr[S_R], z.N * 4, hipMemcpyDeviceToHost));
    E(hipGetLastError());
    const hipError_t e = E.e;
    const float ms = t.stop();
    if (e != hipSuccess) { set_err(where, hipGetErrorString(e)); return -1.0f; }
    return ms;
}

// =============================== Part 2 ===============================

namespace {

float solve_host(float* u, float* v, float* w, const bool* occupy, int W_, int H_, int D_, const zq3d_sor_opts& o,
                 float given_mean, zq3d_proj_stats* st, const char* where) {
    if (check_dims(W_, H_, D_, where)) return -1.0f;
    if (check_sor_opts(o, where)) return -1.0f;
    if (!u || !v || !w || !occupy) { set_err(where, "null pointer"); return -1.0f; }
    const Dims d{W_, H_, D_};
    const Sizes z(d);
    Workspace& W = ws();
    std::lock_guard<std::mutex> lk(W.mu);
    Timer t;
    Errs E;
    if (W.ensure(S_U, z.Nu * 4) || W.ensure(S_V, z.Nv * 4) || W.ensure(S_W, z.Nw * 4) || W.ensure(S_OCC, z.N) ||
        W.ensure(S_NC, z.N) || W.ensure(S_B, z.N * 4) || W.ensure(S_P, z.Np * 4) || W.ensure(S_P2, z.Np * 4) ||
        W.ensure(S_PART, (size_t)reduction_partials(d) * 8) || W.ensure(S_SCAL, sizeof(Scalars)) ||
        (o.mixed_precision && (W.ensure(S_B64, z.N * 8) || W.ensure(S_P64, z.Np * 8)))) {
        set_err(where, "device allocation failed");
        return -1.0f;
    }
    float* du = (float*)W.ptr[S_U];
    float* dv = (float*)W.ptr[S_V];
    float* dw = (float*)W.ptr[S_W];
    uint8_t* docc = (uint8_t*)W.ptr[S_OCC];
    E(hipMemcpy(du, u, z.Nu * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(dv, v, z.Nv * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(dw, w, z.Nw * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(docc, occupy, z.N, hipMemcpyHostToDevice));
    // gated passes run on colour-split storage: pressure pairs, right-hand sides, counts
    SplitBufs split{};
    if (o.mixed_precision && o.pass_rule == 2) {
        const size_t sp = split_padded_size(d), sc = split_cell_size(d);
        if (W.ensure(S_SPLIT, 4 * sp * 4 + 2 * sc * 4 + 2 * sc)) {
            set_err(where, "device allocation failed");
            return -1.0f;
        }
        float* f = (float*)W.ptr[S_SPLIT];
        uint8_t* n8 = (uint8_t*)(f + 4 * sp + 2 * sc);
        split = SplitBufs{{f, f + sp}, {f + 4 * sp, f + 4 * sp + sc}, {n8, n8 + sc}, {f + 2 * sp, f + 3 * sp}};
        launch_fill_f32(f, 4 * sp, 0.0f, 0);  // padding slots hold 0
    }
    const ProjBufs B{(uint8_t*)W.ptr[S_NC],  (float*)W.ptr[S_B],      (float*)W.ptr[S_P],
                     (float*)W.ptr[S_P2],    (double*)W.ptr[S_B64],   (double*)W.ptr[S_P64],
                     (double*)W.ptr[S_PART], (Scalars*)W.ptr[S_SCAL], SorShape{0, 0}, split};
    launch_fill_f32(B.p_alt, z.Np, 0.0f, 0);  // zero ghost layer of the second sweep buffer
    int rc;
    if (o.mixed_precision) {
        launch_fill_f64(B.p64, z.Np, 0.0, 0);
        rc = project_mixed(d, du, dv, dw, docc, o, B, st, 0);
    } else {
        rc = project_plain(d, du, dv, dw, docc, o, given_mean, o.closed_mean ? 1 : 0, B, st, 0);
    }
    if (rc) return -1.0f;
    E(hipMemcpy(u, du, z.Nu * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(v, dv, z.Nv * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(w, dw, z.Nw * 4, hipMemcpyDeviceToHost));
    E(hipGetLastError());
    const hipError_t e = E.e;
    const float ms = t.stop();
    if (e != hipSuccess) { set_err(where, hipGetErrorString(e)); return -1.0f; }
    return ms;
}

}  // namespace

extern "C" float SolveOpenPoissonRedBlackwithOccupy3D_MAC(float* mac_u, float* mac_v, float* mac_w,
                                                          const bool* occupy, const int width, const int height,
                                                          const int depth, const int maxIter) {
    zq3d_sor_opts o{};
    o.open_mask = ZQ3D_OPEN_ALL;
    o.closed_mean = 0;
    o.omega = 1.0f;
    o.iters = maxIter;
    return solve_host(mac_u, mac_v, mac_w, occupy, width, height, depth, o, 0.0f, nullptr,
                      "SolveOpenPoissonRedBlackwithOccupy3D_MAC");
}

extern "C" float SolveClosedPoissonRedBlackwithOccupy3D_MAC(float* mac_u, float* mac_v, float* mac_w,
                                                            const bool* occupy, const float div_per_volume,
                                                            const int width, const int height, const int depth,
                                                            const int maxIter) {
    zq3d_sor_opts o{};
    o.open_mask = 0;
    o.closed_mean = 0;  // the caller's div_per_volume is used as given
    o.omega = 1.0f;
    o.iters = maxIter;
    return solve_host(mac_u, mac_v, mac_w, occupy, width, height, depth, o, div_per_volume, nullptr,
                      "SolveClosedPoissonRedBlackwithOccupy3D_MAC");
}

extern "C" float ZQ3D_SolvePoissonRedBlack_ex(float* mac_u, float* mac_v, float* mac_w, const bool* occupy,
                                              const int width, const int height, const int depth,
                                              const zq3d_sor_opts* opts, zq3d_proj_stats* stats) {
    if (!opts) { set_err("ZQ3D_SolvePoissonRedBlack_ex", "null options"); return -1.0f; }
    return solve_host(mac_u, mac_v, mac_w, occupy, width, height, depth, *opts, 0.0f, stats,
                      "ZQ3D_SolvePoissonRedBlack_ex");
}

extern "C" float ZQ3D_MACtoRegular4(const float* mac_u, const float* mac_v, const float* mac_w, float* vel4,
                                    const int width, const int height, const int depth) {
    const char* where = "ZQ3D_MACtoRegular4";
    if (check_dims(width, height, depth, where)) return -1.0f;
    if (!mac_u || !mac_v || !mac_w || !vel4) { set_err(where, "null pointer"); return -1.0f; }
    const Dims d{width, height, depth};
    const Sizes z(d);
    Workspace& W = ws();
    std::lock_guard<std::mutex> lk(W.mu);
    Timer t;
    Errs E;
    if (W.ensure(S_U, z.Nu * 4) || W.ensure(S_V, z.Nv * 4) || W.ensure(S_W, z.Nw * 4) || W.ensure(S_VEL4, z.N * 16)) {
        set_err(where, "device allocation failed");
        return -1.0f;
    }
    E(hipMemcpy(W.ptr[S_U], mac_u, z.Nu * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_V], mac_v, z.Nv * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(W.ptr[S_W], mac_w, z.Nw * 4, hipMemcpyHostToDevice));
    launch_mac_to_cell4(d, (float*)W.ptr[S_U], (float*)W.ptr[S_V], (float*)W.ptr[S_W], (float4*)W.ptr[S_VEL4], 0);
    E(hipMemcpy(vel4, W.ptr[S_VEL4], z.N * 16, hipMemcpyDeviceToHost));
    E(hipGetLastError());
    const hipError_t e = E.e;
    const float ms = t.stop();
    if (e != hipSuccess) { set_err(where, hipGetErrorString(e)); return -1.0f; }
    return ms;
}

// =============================== Part 3 ===============================

struct zq3d_sim {
    zq3d_config cfg;
    Dims d;
    hipStream_t stream = nullptr;
    float *u = nullptr, *v = nullptr, *w = nullptr;
    float *T = nullptr, *R = nullptr, *T2 = nullptr, *R2 = nullptr;
    uint8_t* occ = nullptr;
    float4* vel4 = nullptr;       // also the vorticity scratch of the force stage
    float* scratch = nullptr;     // Nu+Nv+Nw floats: face scratch, cell force, advected cells
    float* bfecc = nullptr;       // BFECC velocity modes: bfecc_scratch_values(d) floats
    uint8_t* nc = nullptr;
    float* b = nullptr;           // FP32 divergence / FP32 residual of the mixed solve
    float* p = nullptr;           // padded FP32 pressure / correction e of the mixed solve
    float* p_alt = nullptr;       // second sweep buffer (ping-pong), zero ghost layer
    double* part = nullptr;       // reduction partials
    float* split_p = nullptr;     // colour-split pressures (2 pairs x 2 colours, padded), rhs, counts
    float* split_b = nullptr;
    uint8_t* split_n = nullptr;
    double* b64 = nullptr;
    double* p64 = nullptr;
    Scalars* sc = nullptr;
    std::vector<float> boxes, spheres;
    bool p64_valid = false;
    double* p64_prev = nullptr;    // warm_start 2: the pressure of the step before p64's
    bool p64_prev_valid = false;
    bool prof = false;      // stage timing (an event pair per stage, waited for)
    bool sor_prof = false;  // sweep timing (pooled event pairs, not waited for until read)
    hipEvent_t ev[2] = {nullptr, nullptr};
    SorProf sorprof;
    zq3d_timing tm{};
    SorShape shape{0, 0};  // fused sweep tile shape chosen by zq3d_autotune_sor / zq3d_set_sor_shape
    Readback rb;           // pinned read-back slots of the projection's reductions
    hipEvent_t gate_ev[2] = {nullptr, nullptr};  // progress checks of gated passes
    CoarseState coarse;    // sub_size > 0: the coarse grid of the coarse + sub-grid projection
    OctreeState octree;    // octree != 0: the leaves and level arrays of the octree projection
};

namespace {

template <class T>
hipError_t dalloc(T** p, size_t bytes) {
    return hipMalloc((void**)p, bytes);
}

int field_info(zq3d_sim* s, int field, void** ptr, size_t* bytes) {
    const Sizes z(s->d);
    switch (field) {
    case ZQ3D_FIELD_U: *ptr = s->u; *bytes = z.Nu * 4; return 0;
    case ZQ3D_FIELD_V: *ptr = s->v; *bytes = z.Nv * 4; return 0;
    case ZQ3D_FIELD_W: *ptr = s->w; *bytes = z.Nw * 4; return 0;
    case ZQ3D_FIELD_T: *ptr = s->T; *bytes = z.N * 4; return 0;
    case ZQ3D_FIELD_RHO: *ptr = s->R; *bytes = z.N * 4; return 0;
    case ZQ3D_FIELD_OCC: *ptr = s->occ; *bytes = z.N; return 0;
    case ZQ3D_FIELD_P64: *ptr = s->p64; *bytes = z.Np * 8; return 0;
    case ZQ3D_FIELD_VEL4: *ptr = s->vel4; *bytes = z.N * 16; return 0;
    default: return -1;
    }
}

ProjBufs proj_bufs(zq3d_sim* s) {
    const size_t sp = split_padded_size(s->d), sc = split_cell_size(s->d);
    const SplitBufs split{{s->split_p, s->split_p + sp},
                          {s->split_b, s->split_b + sc},
                          {s->split_n, s->split_n + sc},
                          {s->split_p + 2 * sp, s->split_p + 3 * sp}};
    return ProjBufs{s->nc, s->b, s->p, s->p_alt, s->b64, s->p64, s->part, s->sc, s->shape, split, s->rb,
                    {s->gate_ev[0], s->gate_ev[1]}};
}

// Times one stage of a simulation with an event pair when profiling is on.
struct StageTimer {
    zq3d_sim* s;
    double* acc;
    StageTimer(zq3d_sim* s_, double* acc_) : s(s_), acc(acc_) {
        if (s->prof) (void)hipEventRecord(s->ev[0], s->stream);
    }
    int done(const char* where) {
        if (!s->prof) return ZQ3D_OK;
        ZQ_HIP(hipEventRecord(s->ev[1], s->stream), where);
        ZQ_HIP(hipEventSynchronize(s->ev[1]), where);
        float ms = 0.0f;
        ZQ_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]), where);
        *acc += ms;
        return ZQ3D_OK;
    }
};

int check_config(const zq3d_config& c, const char* where) {
    if (c.struct_size != (int)sizeof(zq3d_config)) {
        set_err(where, "zq3d_config.struct_size mismatch (header/library version)");
        return ZQ3D_ERR_ARG;
    }
    if (int rc = check_dims(c.width, c.height, c.depth, where)) return rc;
    if (c.steps < 1) { set_err(where, "steps must be >= 1"); return ZQ3D_ERR_ARG; }
    if (!(c.voxel_size > 0.0f) || !(c.step_time >= 0.0f)) {
        set_err(where, "voxel_size must be > 0 and step_time >= 0");
        return ZQ3D_ERR_ARG;
    }
    if (c.adv_vel_type < 0 || c.adv_vel_type > 5) { set_err(where, "unknown adv_vel_type"); return ZQ3D_ERR_ARG; }
    if (c.adv_sca_type < 0 || c.adv_sca_type > 3) { set_err(where, "unknown adv_sca_type"); return ZQ3D_ERR_ARG; }
    if (c.warm_start < 0 || c.warm_start > 2) { set_err(where, "warm_start must be 0, 1 or 2"); return ZQ3D_ERR_ARG; }
    if (c.warm_start == 2 && !(c.warm_extrapolate >= 0.0f && c.warm_extrapolate <= 2.0f)) {
        set_err(where, "warm_extrapolate must be in [0, 2]");
        return ZQ3D_ERR_ARG;
    }
    if (c.sub_size < 0 || c.sub_size > COARSE_MAX_S ||
        (c.sub_size && (c.width % c.sub_size || c.height % c.sub_size || c.depth % c.sub_size))) {
        set_err(where, "sub_size must be 0 or in [1, 8] dividing width, height and depth");
        return ZQ3D_ERR_ARG;
    }
    if (c.coarse_flux && !c.sub_size) {
        set_err(where, "coarse_flux selects the flux form of the coarse solve: it needs sub_size > 0");
        return ZQ3D_ERR_ARG;
    }
    if (c.octree) {
        if (c.width % 8 || c.height % 8 || c.depth % 8) {
            set_err(where, "octree: width, height and depth must be multiples of 8");
            return ZQ3D_ERR_ARG;
        }
        if ((long long)c.width * c.height * c.depth >= (1LL << 30)) {
            set_err(where, "octree: at most 2^30 cells (neighbour entries hold 30-bit indices)");
            return ZQ3D_ERR_ARG;
        }
        if (c.sub_size || c.sor.mixed_precision) {
            set_err(where, "octree runs its own FP32 solve: sub_size and sor.mixed_precision must be 0");
            return ZQ3D_ERR_ARG;
        }
        if (c.octree_thresh[0] < 0 || c.octree_thresh[1] < 0 || c.octree_thresh[2] < 0) {
            set_err(where, "octree_thresh must be >= 0");
            return ZQ3D_ERR_ARG;
        }
    }
    if (c.sub_size && c.sor.mixed_precision) {
        set_err(where, "sub_size > 0 runs the coarse FP32 SOR: sor.mixed_precision must be 0");
        return ZQ3D_ERR_ARG;
    }
    return check_sor_opts(c.sor, where);
}

}  // namespace

extern "C" int zq3d_config_default(zq3d_config* c, int width, int height, int depth) {
    if (!c) { set_err("zq3d_config_default", "null config"); return ZQ3D_ERR_ARG; }
    std::memset(c, 0, sizeof(*c));
    c->struct_size = (int)sizeof(zq3d_config);
    c->width = width;
    c->height = height;
    c->depth = depth;
    c->voxel_size = (float)(1.0 / width);
    c->step_time = 0.01f;
    c->steps = 20;
    c->Tamb = 10.0f;
    c->buoyancy = 1.0f;
    c->gravity = 0.0f;
    c->confine = 0.0f;
    c->vel_atten = c->temp_atten = c->dens_atten = 1.0f;
    c->adv_vel_type = 0;  // ADV_VEL_INREG_OUTREG
    c->adv_sca_type = 2;  // ADV_SCA_REG
    c->sor.open_mask = ZQ3D_OPEN_ALL;
    c->sor.closed_mean = 0;
    c->sor.omega = 1.0f;
    c->sor.iters = 100;
    c->sor.mixed_precision = 0;
    c->sor.max_outer = 1000;  // maxOuter's default in the config reader
    c->sor.tol = 1e-6;
    c->warm_start = 1;
    c->warm_extrapolate = 0.5f;
    c->octree_thresh[0] = 2;  // octree_thresh1..3 (ZQ_SmokeSimulationUtils3D.cpp:52-54)
    c->octree_thresh[1] = 4;
    c->octree_thresh[2] = 16;
    return ZQ3D_OK;
}

extern "C" int zq3d_create(const zq3d_config* cfg, zq3d_sim** out) {
    const char* where = "zq3d_create";
    if (!cfg || !out) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    *out = nullptr;
    if (int rc = check_config(*cfg, where)) return rc;
    zq3d_sim* s = new zq3d_sim();
    s->cfg = *cfg;
    s->d = Dims{cfg->width, cfg->height, cfg->depth};
    const Sizes z(s->d);
    hipError_t e = hipSuccess;
    auto A = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    A(dalloc(&s->u, z.Nu * 4));
    A(dalloc(&s->v, z.Nv * 4));
    A(dalloc(&s->w, z.Nw * 4));
    A(dalloc(&s->T, z.N * 4));
    A(dalloc(&s->R, z.N * 4));
    A(dalloc(&s->T2, z.N * 4));
    A(dalloc(&s->R2, z.N * 4));
    A(dalloc(&s->occ, z.N));
    A(dalloc(&s->vel4, z.N * 16));
    A(dalloc(&s->scratch, (z.Nu + z.Nv + z.Nw) * 4));
    A(dalloc(&s->nc, z.N));
    A(dalloc(&s->b, z.N * 4));
    A(dalloc(&s->p, z.Np * 4));
    A(dalloc(&s->p_alt, z.Np * 4));
    A(dalloc(&s->part, (size_t)reduction_partials(s->d) * 8));
    A(dalloc(&s->split_p, 4 * split_padded_size(s->d) * 4));
    A(dalloc(&s->split_b, 2 * split_cell_size(s->d) * 4));
    A(dalloc(&s->split_n, 2 * split_cell_size(s->d)));
    A(dalloc(&s->sc, sizeof(Scalars)));
    A(s->rb.alloc());
    A(hipEventCreateWithFlags(&s->gate_ev[0], hipEventDisableTiming));
    A(hipEventCreateWithFlags(&s->gate_ev[1], hipEventDisableTiming));
    if (cfg->sor.mixed_precision) {
        A(dalloc(&s->b64, z.N * 8));
        A(dalloc(&s->p64, z.Np * 8));
        if (cfg->warm_start == 2) A(dalloc(&s->p64_prev, z.Np * 8));
    }
    if (cfg->adv_vel_type == 1 || cfg->adv_vel_type == 3 || cfg->adv_vel_type == 5)
        A(dalloc(&s->bfecc, bfecc_scratch_values(s->d) * 4));
    if (e != hipSuccess) {
        set_err(where, hipGetErrorString(e));
        zq3d_destroy(s);
        return ZQ3D_ERR_HIP;
    }
    if (cfg->sub_size > 0) {
        if (int rc = coarse_create(s->d, cfg->sub_size, &s->coarse, where)) {
            zq3d_destroy(s);
            return rc;
        }
        s->coarse.flux = cfg->coarse_flux ? 1 : 0;
    }
    if (cfg->octree) {
        if (int rc = octree_create(s->d, cfg->octree_thresh, &s->octree, where)) {
            zq3d_destroy(s);
            return rc;
        }
    }
    hipStream_t st = nullptr;
    launch_fill_f32(s->u, z.Nu, 0.0f, st);
    launch_fill_f32(s->v, z.Nv, 0.0f, st);
    launch_fill_f32(s->w, z.Nw, 0.0f, st);
    launch_fill_f32(s->T, z.N, 0.0f, st);
    launch_fill_f32(s->R, z.N, 0.0f, st);
    launch_fill_u8(s->occ, z.N, 0, st);
    launch_fill_f32(s->p, z.Np, 0.0f, st);
    launch_fill_f32(s->p_alt, z.Np, 0.0f, st);
    launch_fill_f32(s->split_p, 4 * split_padded_size(s->d), 0.0f, st);
    if (s->p64) launch_fill_f64(s->p64, z.Np, 0.0, st);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        set_err(where, hipGetErrorString(e));
        zq3d_destroy(s);
        return ZQ3D_ERR_HIP;
    }
    *out = s;
    return ZQ3D_OK;
}

extern "C" void zq3d_destroy(zq3d_sim* s) {
    if (!s) return;
    void* ptrs[] = {s->u, s->v, s->w, s->T, s->R, s->T2, s->R2, s->occ, s->vel4, s->scratch, s->bfecc, s->nc, s->b, s->p, s->p_alt, s->part,
                    s->split_p, s->split_b, s->split_n,
                    s->b64, s->p64, s->p64_prev, s->sc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : {s->ev[0], s->ev[1], s->gate_ev[0], s->gate_ev[1]})
        if (e) (void)hipEventDestroy(e);
    s->sorprof.release();
    s->rb.release();
    coarse_destroy(&s->coarse);
    octree_destroy(&s->octree);
    delete s;
}

extern "C" int zq3d_set_stream(zq3d_sim* s, void* stream) {
    if (!s) { set_err("zq3d_set_stream", "null sim"); return ZQ3D_ERR_ARG; }
    s->stream = (hipStream_t)stream;
    return ZQ3D_OK;
}

extern "C" int zq3d_synchronize(zq3d_sim* s) {
    if (!s) { set_err("zq3d_synchronize", "null sim"); return ZQ3D_ERR_ARG; }
    ZQ_HIP(hipStreamSynchronize(s->stream), "zq3d_synchronize");
    return ZQ3D_OK;
}

extern "C" int zq3d_upload(zq3d_sim* s, int field, const void* host, size_t bytes) {
    const char* where = "zq3d_upload";
    if (!s || !host) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    void* p;
    size_t n;
    if (field_info(s, field, &p, &n) || !p) { set_err(where, "unknown or unallocated field"); return ZQ3D_ERR_ARG; }
    if (bytes != n) { set_err(where, "size mismatch"); return ZQ3D_ERR_ARG; }
    ZQ_HIP(hipMemcpyAsync(p, host, n, hipMemcpyHostToDevice, s->stream), where);
    ZQ_HIP(hipStreamSynchronize(s->stream), where);
    if (field == ZQ3D_FIELD_P64) {
        s->p64_valid = true;
        s->p64_prev_valid = false;  // the extrapolation starts again from the uploaded pressure
    }
    return ZQ3D_OK;
}

extern "C" int zq3d_download(zq3d_sim* s, int field, void* host, size_t bytes) {
    const char* where = "zq3d_download";
    if (!s || !host) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    void* p;
    size_t n;
    if (field_info(s, field, &p, &n) || !p) { set_err(where, "unknown or unallocated field"); return ZQ3D_ERR_ARG; }
    if (bytes != n) { set_err(where, "size mismatch"); return ZQ3D_ERR_ARG; }
    ZQ_HIP(hipMemcpyAsync(host, p, n, hipMemcpyDeviceToHost, s->stream), where);
    ZQ_HIP(hipStreamSynchronize(s->stream), where);
    return ZQ3D_OK;
}

extern "C" int zq3d_device_ptr(zq3d_sim* s, int field, void** dev, size_t* bytes) {
    const char* where = "zq3d_device_ptr";
    if (!s || !dev) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    size_t n;
    if (field_info(s, field, dev, &n) || !*dev) { set_err(where, "unknown or unallocated field"); return ZQ3D_ERR_ARG; }
    if (bytes) *bytes = n;
    return ZQ3D_OK;
}

extern "C" int zq3d_set_box_sources(zq3d_sim* s, int count, const float* boxes8) {
    if (!s || count < 0 || (count && !boxes8)) { set_err("zq3d_set_box_sources", "bad argument"); return ZQ3D_ERR_ARG; }
    s->boxes.assign(boxes8, boxes8 + 8 * count);
    return ZQ3D_OK;
}

extern "C" int zq3d_set_sphere_sources(zq3d_sim* s, int count, const float* sph6) {
    if (!s || count < 0 || (count && !sph6)) { set_err("zq3d_set_sphere_sources", "bad argument"); return ZQ3D_ERR_ARG; }
    s->spheres.assign(sph6, sph6 + 6 * count);
    return ZQ3D_OK;
}

extern "C" int zq3d_stamp_sphere_obstacle(zq3d_sim* s, float cx, float cy, float cz, float radius, float su, float sv,
                                          float sw, int shell_mask) {
    if (!s || (shell_mask & ~63)) { set_err("zq3d_stamp_sphere_obstacle", "bad argument"); return ZQ3D_ERR_ARG; }
    launch_stamp_sphere(s->d, cx, cy, cz, radius, su, sv, sw, shell_mask, s->occ, s->u, s->v, s->w, s->stream);
    ZQ_HIP(hipGetLastError(), "zq3d_stamp_sphere_obstacle");
    return ZQ3D_OK;
}

extern "C" int zq3d_add_force_attenuate(zq3d_sim* s, const float* extra) {
    if (!s) { set_err("zq3d_add_force_attenuate", "null sim"); return ZQ3D_ERR_ARG; }
    const zq3d_config& c = s->cfg;
    const ForceParams fp{c.voxel_size, c.step_time, c.buoyancy, c.gravity, c.confine, c.Tamb};
    StageTimer t(s, &s->tm.ms_force);
    // vel4 is scratch here (the advection stage fills it again), so the vorticity goes in planes
    launch_force_atten(s->d, s->u, s->v, s->w, s->T, s->R, s->occ, extra, fp, 1, c.vel_atten, c.temp_atten,
                       c.dens_atten, s->vel4, s->scratch, s->stream, true);
    ZQ_HIP(hipGetLastError(), "zq3d_add_force_attenuate");
    return t.done("zq3d_add_force_attenuate");
}

extern "C" int zq3d_advect_velocity(zq3d_sim* s) {
    if (!s) { set_err("zq3d_advect_velocity", "null sim"); return ZQ3D_ERR_ARG; }
    const zq3d_config& c = s->cfg;
    const AdvParams ap{c.voxel_size, c.step_time / c.steps, c.steps};
    // the reference's stage always advects with is_open = true (ZQ_SmokeSimulationUtils3D.cpp:2163)
    StageTimer t(s, &s->tm.ms_advect_velocity);
    launch_advect_velocity(s->d, s->u, s->v, s->w, s->occ, 1, c.adv_vel_type, ap, s->vel4, s->scratch, s->stream,
                           s->bfecc);
    ZQ_HIP(hipGetLastError(), "zq3d_advect_velocity");
    return t.done("zq3d_advect_velocity");
}

extern "C" int zq3d_project(zq3d_sim* s, zq3d_proj_stats* st) {
    if (!s) { set_err("zq3d_project", "null sim"); return ZQ3D_ERR_ARG; }
    const zq3d_config& c = s->cfg;
    StageTimer t(s, &s->tm.ms_project);
    SorProf* pf = s->sor_prof ? &s->sorprof : nullptr;
    int rc;
    if (c.sub_size > 0) {
        rc = project_coarse(s->coarse, s->u, s->v, s->w, s->occ, c.sor, proj_bufs(s), st, s->stream, pf);
    } else if (c.octree) {
        rc = project_octree(s->octree, s->u, s->v, s->w, s->occ, c.sor, proj_bufs(s), st, s->stream, pf);
    } else if (c.sor.mixed_precision) {
        if (!c.warm_start || !s->p64_valid) {
            launch_fill_f64(s->p64, Sizes(s->d).Np, 0.0, s->stream);
            s->p64_prev_valid = false;
        } else if (c.warm_start == 2) {
            launch_extrapolate_p64(s->p64, s->p64_prev, Sizes(s->d).Np, s->p64_prev_valid ? 1 : 0, c.warm_extrapolate,
                                   s->stream);
            s->p64_prev_valid = true;
        }
        rc = project_mixed(s->d, s->u, s->v, s->w, s->occ, c.sor, proj_bufs(s), st, s->stream, pf);
        s->p64_valid = (rc == ZQ3D_OK);
        if (!s->p64_valid) s->p64_prev_valid = false;
    } else {
        rc = project_plain(s->d, s->u, s->v, s->w, s->occ, c.sor, 0.0f, c.sor.closed_mean ? 1 : 0, proj_bufs(s), st,
                           s->stream, pf);
    }
    if (rc) return rc;
    ZQ_HIP(hipGetLastError(), "zq3d_project");
    return t.done("zq3d_project");
}

extern "C" int zq3d_set_coarse_sweep(zq3d_sim* s, int k, int zc) {
    const char* where = "zq3d_set_coarse_sweep";
    if (!s) { set_err(where, "null sim"); return ZQ3D_ERR_ARG; }
    if (s->cfg.sub_size <= 0) { set_err(where, "the simulation has no coarse grid (sub_size == 0)"); return ZQ3D_ERR_STATE; }
    const CoarseSweep sw{k, zc};
    if (k < 0 || (k > 0 && !coarse_sweep_valid(s->coarse.cd, sw))) {
        set_err(where, "k must be 0, or k > 0 and zc > 0 with the slab's planes within 160 KiB of LDS");
        return ZQ3D_ERR_ARG;
    }
    s->coarse.sweep = k == 0 ? CoarseSweep{0, 0} : sw;
    return ZQ3D_OK;
}

extern "C" int zq3d_set_guide(zq3d_sim* s, float coeff, const float* gu, const float* gv, const float* gw) {
    const char* where = "zq3d_set_guide";
    if (!s) { set_err(where, "null sim"); return ZQ3D_ERR_ARG; }
    if (s->cfg.octree) {
        // the octree's guiding is the special-region projection step; the guide values are not read
        s->octree.guiding = gu != nullptr;
        return ZQ3D_OK;
    }
    if (s->cfg.sub_size <= 0) {
        set_err(where, "guiding needs the coarse grid (sub_size > 0) or the octree projection");
        return ZQ3D_ERR_STATE;
    }
    return coarse_set_guide(&s->coarse, coeff, gu, gv, gw, s->stream, where);
}

extern "C" int zq3d_get_coarse_sweep(zq3d_sim* s, int* k, int* zc) {
    const char* where = "zq3d_get_coarse_sweep";
    if (!s || !k || !zc) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    if (s->cfg.sub_size <= 0) { set_err(where, "the simulation has no coarse grid (sub_size == 0)"); return ZQ3D_ERR_STATE; }
    *k = s->coarse.sweep.k;
    *zc = s->coarse.sweep.zc;
    return ZQ3D_OK;
}

extern "C" int zq3d_advect_scalar(zq3d_sim* s) {
    if (!s) { set_err("zq3d_advect_scalar", "null sim"); return ZQ3D_ERR_ARG; }
    const zq3d_config& c = s->cfg;
    const Sizes z(s->d);
    StageTimer t(s, &s->tm.ms_advect_scalar);
    if (c.adv_sca_type == 1 || c.adv_sca_type == 3) {
        // reference behaviour: the BFECC scalar modes return zero fields (ZQ_CUDA_Advection3D.cu:1782-1794)
        launch_fill_f32(s->T, z.N, 0.0f, s->stream);
        launch_fill_f32(s->R, z.N, 0.0f, s->stream);
    } else {
        const AdvParams ap{c.voxel_size, c.step_time / c.steps, c.steps};
        launch_advect_scalar(s->d, s->u, s->v, s->w, s->occ, s->T, s->R, s->T2, s->R2, c.adv_sca_type == 0 ? 1 : 0,
                             ap, s->vel4, s->stream);
        std::swap(s->T, s->T2);
        std::swap(s->R, s->R2);
    }
    ZQ_HIP(hipGetLastError(), "zq3d_advect_scalar");
    return t.done("zq3d_advect_scalar");
}

extern "C" int zq3d_reinject(zq3d_sim* s) {
    if (!s) { set_err("zq3d_reinject", "null sim"); return ZQ3D_ERR_ARG; }
    StageTimer t(s, &s->tm.ms_reinject);
    for (size_t i = 0; i + 8 <= s->boxes.size(); i += 8)
        launch_reinject_box(s->d, s->boxes.data() + i, s->T, s->R, s->stream);
    for (size_t i = 0; i + 6 <= s->spheres.size(); i += 6)
        launch_reinject_sphere(s->d, s->spheres.data() + i, s->T, s->R, s->stream);
    ZQ_HIP(hipGetLastError(), "zq3d_reinject");
    return t.done("zq3d_reinject");
}

extern "C" int zq3d_step(zq3d_sim* s, const float* extra, zq3d_proj_stats* st) {
    int rc;
    if ((rc = zq3d_add_force_attenuate(s, extra))) return rc;
    if ((rc = zq3d_advect_velocity(s))) return rc;
    if ((rc = zq3d_project(s, st))) return rc;
    if ((rc = zq3d_advect_scalar(s))) return rc;
    if ((rc = zq3d_reinject(s))) return rc;
    s->tm.steps++;
    return ZQ3D_OK;
}

extern "C" int zq3d_set_profiling(zq3d_sim* s, int on) {
    const char* where = "zq3d_set_profiling";
    if (!s) { set_err(where, "null sim"); return ZQ3D_ERR_ARG; }
    if (on < 0 || on > 2) { set_err(where, "on must be 0, 1 or 2"); return ZQ3D_ERR_ARG; }
    if (on == 1) {
        for (hipEvent_t* e : {&s->ev[0], &s->ev[1]})
            if (!*e) ZQ_HIP(hipEventCreate(e), where);
    }
    ZQ_HIP(s->sorprof.resolve(), where);  // pairs still pending belong to the previous period
    s->prof = on == 1;
    s->sor_prof = on != 0;
    s->tm = zq3d_timing{};
    s->sorprof.ms = 0.0;
    s->sorprof.launches = 0;
    return ZQ3D_OK;
}

extern "C" int zq3d_set_sor_shape(zq3d_sim* s, int ty, int zc) {
    if (!s || zc < 0) { set_err("zq3d_set_sor_shape", "bad argument"); return ZQ3D_ERR_ARG; }
    SorShape cand[32];
    const int n = sor_shape_candidates(s->d, cand, 32);
    bool ok = (ty == 0 && zc == 0);
    for (int i = 0; i < n && !ok; i++) ok = cand[i].ty == ty && cand[i].zc == zc;
    if (!ok) { set_err("zq3d_set_sor_shape", "shape not supported for this grid"); return ZQ3D_ERR_ARG; }
    s->shape = SorShape{ty, zc};
    return ZQ3D_OK;
}

// Times every fused-sweep tile shape valid for the grid on the simulation's own occupancy (the
// sweep's cost depends on it only through skipped solid cells) and keeps the fastest.
extern "C" int zq3d_autotune_sor(zq3d_sim* s, int iters, zq3d_tune_report* rep) {
    const char* where = "zq3d_autotune_sor";
    if (!s || iters < 1) { set_err(where, "bad argument"); return ZQ3D_ERR_ARG; }
    const Sizes z(s->d);
    SorShape cand[ZQ3D_TUNE_MAX];
    const int n = sor_shape_candidates(s->d, cand, ZQ3D_TUNE_MAX);
    if (rep) std::memset(rep, 0, sizeof(*rep));
    if (n == 0) {
        s->shape = SorShape{0, 0};
        if (rep) rep->best = -1;
        return ZQ3D_OK;
    }
    hipEvent_t a = nullptr, b = nullptr;
    ZQ_HIP(hipEventCreate(&a), where);
    ZQ_HIP(hipEventCreate(&b), where);
    launch_fill_f32(&s->sc->mean32, 1, 0.0f, s->stream);
    launch_proj_setup(s->d, s->u, s->v, s->w, s->occ, s->cfg.sor.open_mask, s->nc, s->b, nullptr, s->p, nullptr,
                      nullptr, s->stream);
    int best = -1;
    float best_us = 0.0f;
    Errs E;
    const ProjBufs B = proj_bufs(s);
    for (int i = 0; i < n; i++) {
        launch_fill_f32(s->p, z.Np, 0.0f, s->stream);
        launch_sor(s->d, s->nc, s->b, &s->sc->mean32, s->p, s->p_alt, 1.8f, 2, s->stream, cand[i], &B.split);
        // the events bracket the sweeps only (not the split-layout conversions)
        launch_sor(s->d, s->nc, s->b, &s->sc->mean32, s->p, s->p_alt, 1.8f, iters, s->stream, cand[i], &B.split, a,
                   b);
        E(hipGetLastError());
        E(hipEventSynchronize(b));
        float ms = 0.0f;
        E(hipEventElapsedTime(&ms, a, b));
        const float us = 1e3f * ms / iters;
        if (rep) {
            rep->ty[i] = cand[i].ty;
            rep->zc[i] = cand[i].zc;
            rep->us_per_iter[i] = us;
        }
        if (best < 0 || us < best_us) {
            best = i;
            best_us = us;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    launch_fill_f32(s->p, z.Np, 0.0f, s->stream);
    if (E.e != hipSuccess) { set_err(where, hipGetErrorString(E.e)); return ZQ3D_ERR_HIP; }
    s->shape = cand[best];
    if (rep) {
        rep->count = n;
        rep->best = best;
    }
    return ZQ3D_OK;
}

extern "C" const char* zq3d_sor_kernel(zq3d_sim* s) {
    if (!s) return "";
    return sor_kernel_name(s->d, s->shape);
}

extern "C" int zq3d_get_timing(zq3d_sim* s, zq3d_timing* out) {
    if (!s || !out) { set_err("zq3d_get_timing", "null argument"); return ZQ3D_ERR_ARG; }
    ZQ_HIP(s->sorprof.resolve(), "zq3d_get_timing");
    *out = s->tm;
    out->ms_sor = s->sorprof.ms;
    out->sor_launches = s->sorprof.launches;
    return ZQ3D_OK;
}

extern "C" int zq3d_mac_to_regular(zq3d_sim* s) {
    if (!s) { set_err("zq3d_mac_to_regular", "null sim"); return ZQ3D_ERR_ARG; }
    launch_mac_to_cell4(s->d, s->u, s->v, s->w, s->vel4, s->stream);
    ZQ_HIP(hipGetLastError(), "zq3d_mac_to_regular");
    return ZQ3D_OK;
}

extern "C" const char* zq3d_last_error(void) { return g_err.c_str(); }

extern "C" void zq3d_release_workspace(void) {
    Workspace& W = ws();
    std::lock_guard<std::mutex> lk(W.mu);
    W.release();
}
