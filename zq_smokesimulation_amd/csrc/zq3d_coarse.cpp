// zq3d_coarse.cpp -- host side of the coarse + sub-grid projection (COARSE_OPEN_POISSON_SOR,
// ZQ_SmokeSimulationUtils3D.cpp:3688-3795): the device-resident projection of a zq3d_sim with
// zq3d_config.sub_size > 0, and the two ZQlibCuda entry points that path calls in the reference
// (the face-fraction red-black solve and the matrix product behind SubProjection3DCuda).
#include <cstring>
#include <mutex>

#include "zq3d_internal.h"

namespace zq3d {
namespace detail {

namespace {

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// One allocation holds every coarse array; returns its size and, with base, fills B.
size_t coarse_layout(const CoarseDims& cd, char* base, CoarseBufs* B) {
    size_t off = 0;
    auto f = [&](float** p, size_t n) {
        if (base) *p = (float*)(base + off);
        off += align256(n * sizeof(float));
    };
    auto b = [&](uint8_t** p, size_t n) {
        if (base) *p = (uint8_t*)(base + off);
        off += align256(n);
    };
    CoarseBufs t{};
    CoarseBufs* o = B ? B : &t;
    f(&o->rU, cd.Nu());
    f(&o->rV, cd.Nv());
    f(&o->rW, cd.Nw());
    f(&o->cu, cd.Nu());
    f(&o->cv, cd.Nv());
    f(&o->cw, cd.Nw());
    f(&o->inv, cd.Nc());
    f(&o->div, cd.Nc());
    f(&o->p, cd.Np());
    f(&o->p2, cd.Np());
    f(&o->vol, cd.Nc());
    if (base) o->mean = (double*)(base + off);
    off += 256;
    f(&o->gU, cd.Nu());
    f(&o->gV, cd.Nv());
    f(&o->gW, cd.Nw());
    b(&o->cocc, cd.Nc());
    b(&o->regular, cd.Nc());
    return off;
}

int check_block(const Dims& d, int s, const char* where) {
    if (s < 1 || s > COARSE_MAX_S) {
        set_err(where, "sub_size must be in [1, 8]");
        return ZQ3D_ERR_ARG;
    }
    if (d.W % s || d.H % s || d.D % s) {
        set_err(where, "sub_size must divide width, height and depth");
        return ZQ3D_ERR_ARG;
    }
    return ZQ3D_OK;
}

}  // namespace

// Times every valid sweep schedule (iters iterations each, HIP events on the null stream) on the
// coarse arrays as they stand and keeps the fastest; all schedules give the same bits.
int tune_coarse_sweep(CoarseState* c, int iters, const char* where) {
    CoarseSweep cand[16];
    const int n = coarse_sweep_candidates(c->cd, cand, 16);
    hipEvent_t a = nullptr, b = nullptr;
    ZQ_HIP(hipEventCreate(&a), where);
    ZQ_HIP(hipEventCreate(&b), where);
    float best = 0.0f;
    int bi = 0;
    for (int i = 0; i < n; i++) {
        (void)launch_coarse_sor(c->cd, c->B, 1.0f, 2, cand[i], nullptr);  // first-launch costs
        (void)hipEventRecord(a, nullptr);
        (void)launch_coarse_sor(c->cd, c->B, 1.0f, iters, cand[i], nullptr);
        (void)hipEventRecord(b, nullptr);
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (i == 0 || ms < best) {
            best = ms;
            bi = i;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ZQ_HIP(hipGetLastError(), where);
    c->sweep = cand[bi];
    return ZQ3D_OK;
}

int coarse_create(const Dims& d, int s, CoarseState* out, const char* where) {
    if (int rc = check_block(d, s, where)) return rc;
    out->cd = coarse_dims(d, s);
    const size_t bytes = coarse_layout(out->cd, nullptr, nullptr);
    ZQ_HIP(hipMalloc(&out->mem, bytes), where);
    coarse_layout(out->cd, (char*)out->mem, &out->B);
    // a fluid coarse grid with unit fractions to time the schedules on
    launch_fill_f32(out->B.rU, out->cd.Nu(), 1.0f, nullptr);
    launch_fill_f32(out->B.rV, out->cd.Nv(), 1.0f, nullptr);
    launch_fill_f32(out->B.rW, out->cd.Nw(), 1.0f, nullptr);
    launch_fill_f32(out->B.inv, out->cd.Nc(), 1.0f / 6.0f, nullptr);
    launch_fill_f32(out->B.div, out->cd.Nc(), 0.0f, nullptr);
    launch_fill_f32(out->B.p, out->cd.Np(), 0.0f, nullptr);
    launch_fill_f32(out->B.p2, out->cd.Np(), 0.0f, nullptr);
    ZQ_HIP(hipStreamSynchronize(nullptr), where);
    return tune_coarse_sweep(out, 20, where);
}

int coarse_set_guide(CoarseState* c, float coeff, const float* gu, const float* gv, const float* gw, hipStream_t s,
                     const char* where) {
    if (!gu) {
        c->guide = 0;
        return ZQ3D_OK;
    }
    if (!gv || !gw) {
        set_err(where, "gu, gv and gw must all be given (or gu == NULL to turn guiding off)");
        return ZQ3D_ERR_ARG;
    }
    ZQ_HIP(hipMemcpyAsync(c->B.gU, gu, c->cd.Nu() * 4, hipMemcpyHostToDevice, s), where);
    ZQ_HIP(hipMemcpyAsync(c->B.gV, gv, c->cd.Nv() * 4, hipMemcpyHostToDevice, s), where);
    ZQ_HIP(hipMemcpyAsync(c->B.gW, gw, c->cd.Nw() * 4, hipMemcpyHostToDevice, s), where);
    ZQ_HIP(hipStreamSynchronize(s), where);
    c->guide = 1;
    c->guide_coeff = coeff;
    return ZQ3D_OK;
}

void coarse_destroy(CoarseState* c) {
    if (c && c->mem) (void)hipFree(c->mem);
    if (c) c->mem = nullptr;
}

int project_coarse(const CoarseState& c, float* u, float* v, float* w, const uint8_t* occ, const zq3d_sor_opts& o,
                   const ProjBufs& B, zq3d_proj_stats* st, hipStream_t s, SorProf* pf) {
    const char* where = "project_coarse";
    // the fractions follow the occupancy, which callers may change between projections
    launch_coarse_setup(c.cd, occ, o.open_mask, c.B, s, c.flux);
    launch_coarse_restrict(c.cd, u, v, w, c.B, s);
    if (o.closed_mean) launch_coarse_remove_mean(c.cd, c.B, s);
    hipEvent_t ea = nullptr, eb = nullptr;
    if (pf) {
        ZQ_HIP(pf->next(&ea, &eb), where);
        ZQ_HIP(hipEventRecord(ea, s), where);
    }
    const float* p = launch_coarse_sor(c.cd, c.B, o.omega, o.iters, c.sweep, s);
    if (pf) {
        ZQ_HIP(hipEventRecord(eb, s), where);
        pf->launches += c.sweep.k > 0 ? (o.iters + c.sweep.k - 1) / c.sweep.k : 2LL * o.iters;
    }
    launch_coarse_prolong(c.cd, c.B, p, o.open_mask, occ, u, v, w, s, c.guide, c.guide_coeff);
    if (c.cd.s > 1) launch_sub_project(c.cd, c.B, occ, u, v, w, s);
    ZQ_HIP(hipGetLastError(), where);
    if (st) {
        Scalars* sc = B.sc;
        launch_fill_f64(&sc->mean64, 1, 0.0, s);
        launch_proj_setup(c.cd.f, u, v, w, occ, o.open_mask, B.nc, B.b, nullptr, B.p, nullptr, nullptr, s);
        launch_max_div64(c.cd.f, u, v, w, B.nc, &sc->mean64, B.part, B.target(1, &sc->dmax), s);
        double md = 0.0;
        if (int rc = read_result(B, 1, &sc->dmax, &md, s, where)) return rc;
        st->outer = 0;
        st->sweeps = o.iters;
        st->max_residual = -1.0;
        st->max_div = md;
        st->converged = -1;
        st->mean = 0.0;
    }
    return ZQ3D_OK;
}

}  // namespace detail
}  // namespace zq3d

using namespace zq3d;
using namespace zq3d::detail;

namespace {

// grow-only device arena of the host-pointer entry points below
struct Arena {
    void* ptr = nullptr;
    size_t cap = 0;
    std::mutex mu;
    hipError_t ensure(size_t bytes) {
        if (cap >= bytes) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&ptr, bytes);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
};

Arena& arena() {
    static Arena a;
    return a;
}

struct EvTimer {
    hipEvent_t a = nullptr, b = nullptr;
    EvTimer() {
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
    ~EvTimer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

}  // namespace

extern "C" float SolveOpenPoissonRedBlackwithFaceRatio3D_MAC(float* mac_u, float* mac_v, float* mac_w,
                                                            const bool* occupy, const float* unoccupyU,
                                                            const float* unoccupyV, const float* unoccupyW,
                                                            const int width, const int height, const int depth,
                                                            const int maxIter) {
    const char* where = "SolveOpenPoissonRedBlackwithFaceRatio3D_MAC";
    if (check_dims(width, height, depth, where)) return -1.0f;
    if (!mac_u || !mac_v || !mac_w || !occupy || !unoccupyU || !unoccupyV || !unoccupyW) {
        set_err(where, "null pointer");
        return -1.0f;
    }
    if (maxIter < 0) { set_err(where, "maxIter must be >= 0"); return -1.0f; }
    const CoarseDims cd = coarse_dims(Dims{width, height, depth}, 1);
    Arena& A = arena();
    std::lock_guard<std::mutex> lk(A.mu);
    const size_t bytes = coarse_layout(cd, nullptr, nullptr);
    if (A.ensure(bytes) != hipSuccess) { set_err(where, "device allocation failed"); return -1.0f; }
    CoarseBufs B{};
    coarse_layout(cd, (char*)A.ptr, &B);
    EvTimer t;
    Errs E;
    static_assert(sizeof(bool) == 1, "occupancy is one byte per cell");
    E(hipMemcpy(B.cu, mac_u, cd.Nu() * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(B.cv, mac_v, cd.Nv() * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(B.cw, mac_w, cd.Nw() * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(B.rU, unoccupyU, cd.Nu() * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(B.rV, unoccupyV, cd.Nv() * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(B.rW, unoccupyW, cd.Nw() * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(B.cocc, occupy, cd.Nc(), hipMemcpyHostToDevice));
    launch_coarse_prepare(cd, B.cu, B.cv, B.cw, ZQ3D_OPEN_ALL, B, 0);
    const CoarseSweep sw = coarse_sweep_valid(cd, CoarseSweep{2, 1}) ? CoarseSweep{2, 1} : CoarseSweep{0, 0};
    const float* p = launch_coarse_sor(cd, B, 1.0f, maxIter, sw, 0);
    launch_coarse_correct(cd, B, p, ZQ3D_OPEN_ALL, B.cu, B.cv, B.cw, 0);
    E(hipGetLastError());
    E(hipMemcpy(mac_u, B.cu, cd.Nu() * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_v, B.cv, cd.Nv() * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_w, B.cw, cd.Nw() * 4, hipMemcpyDeviceToHost));
    const float ms = t.stop();
    if (E.e != hipSuccess) { set_err(where, hipGetErrorString(E.e)); return -1.0f; }
    return ms;
}

extern "C" float MatrixMul(const float* A_, const float* B_, const int h1, const int w1, const int w2, float* C_) {
    const char* where = "MatrixMul";
    if (!A_ || !B_ || !C_) { set_err(where, "null pointer"); return -1.0f; }
    if (h1 <= 0 || w1 <= 0 || w2 <= 0 || (long long)h1 * w1 >= (1LL << 31) || (long long)w1 * w2 >= (1LL << 31) ||
        (long long)h1 * w2 >= (1LL << 31)) {
        set_err(where, "sizes must be positive and each matrix below 2^31 entries");
        return -1.0f;
    }
    const size_t na = (size_t)h1 * w1, nb = (size_t)w1 * w2, nc = (size_t)h1 * w2;
    Arena& A = arena();
    std::lock_guard<std::mutex> lk(A.mu);
    const size_t oa = 0, ob = align256(na * 4), oc = ob + align256(nb * 4);
    if (A.ensure(oc + nc * 4) != hipSuccess) { set_err(where, "device allocation failed"); return -1.0f; }
    float* dA = (float*)((char*)A.ptr + oa);
    float* dB = (float*)((char*)A.ptr + ob);
    float* dC = (float*)((char*)A.ptr + oc);
    EvTimer t;
    Errs E;
    E(hipMemcpy(dA, A_, na * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(dB, B_, nb * 4, hipMemcpyHostToDevice));
    launch_batched_matvec(h1, w1, w2, dA, dB, dC, 0);
    E(hipGetLastError());
    E(hipMemcpy(C_, dC, nc * 4, hipMemcpyDeviceToHost));
    const float ms = t.stop();
    if (E.e != hipSuccess) { set_err(where, hipGetErrorString(E.e)); return -1.0f; }
    return ms;
}
