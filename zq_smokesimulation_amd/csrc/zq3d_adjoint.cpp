// zq3d_adjoint.cpp -- Part 5 of the C ABI: checkpointed rollout and hand-written adjoint.
//
// Forward step s (the order of ZQ_SmokeSimulation3D::UpdateOneFrame, ZQ_SmokeSimulation3D.cpp:156-243):
//   force+atten(u_s, T_s, rho_s; c_s) -> (u_f, T_f, rho_f); advect u_f -> u_a; project u_a -> u_p;
//   advect (T_f, rho_f) through u_p; reinject.
// Adjoint step s, in reverse: reinjected cells forget their adjoint; scalar advection (field and
// velocity adjoints); projection; velocity advection; force (control, T, rho and u adjoints).
// The projection's adjoint: with free faces F (between two fluid cells, open domain faces) and fixed
// faces X, the forward is u_F' = u_F - D_F^T L^+ Pi (D_F u_F + D_X u_X), u_X' = u_X (Pi removes
// the closed-domain mean).  Its transpose needs the same Poisson solve on the free-face divergence
// of the adjoint, the free constant of a closed-domain pressure removed, and the correction applied
// to every face (fixed faces feed the divergence).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "zq3d_adjoint.h"
#include "zq3d_internal.h"
#include "zq3d_kernels.h"
#include "zq_smoke3d.h"

using namespace zq3d;
using namespace zq3d::detail;

namespace {

// Mixed-precision projection on faces of type F.  Forward: tolerance o.tol on max|r|.  adjoint:
// right-hand side from the free faces, tolerance rtol * max|rhs|, closed-domain constant removed,
// correction on every face.  *bmax (nullable) receives max|rhs| over counted cells.
template <class F>
int project_mp(const Dims& d, F* u, F* v, F* w, const uint8_t* occ, const zq3d_sor_opts& o, const ProjBufs& B,
               bool adjoint, double rtol, zq3d_proj_stats* st, hipStream_t s, double* bmax_out) {
    const char* where = adjoint ? "adjoint projection" : "projection";
    const Sizes z(d);
    Scalars* sc = B.sc;
    launch_fill_f32(&sc->zero32, 1, 0.0f, s);
    if (adjoint) {
        if constexpr (sizeof(F) == 4)
            launch_proj_setup_free(d, u, v, w, occ, o.open_mask, B.nc, B.b64, B.p64, B.part, s);
        else
            launch_proj_setup64(d, u, v, w, occ, o.open_mask, 1, B.nc, B.b64, B.p64, B.part, s);
    } else {
        if constexpr (sizeof(F) == 4)
            launch_proj_setup(d, u, v, w, occ, o.open_mask, B.nc, nullptr, B.b64, nullptr, B.p64, B.part, s);
        else
            launch_proj_setup64(d, u, v, w, occ, o.open_mask, 0, B.nc, B.b64, B.p64, B.part, s);
    }
    // the FP32-face forward set-up marches z (its own partial layout); the others run per plane
    if (sizeof(F) == 4 && !adjoint)
        launch_mean_finalize_mp(d, B.part, &sc->mean64, o.closed_mean, s);
    else
        launch_mean_finalize(d, B.part, nullptr, &sc->mean64, o.closed_mean, s);
    double tol = o.tol;
    if (adjoint || bmax_out) {
        double bmax = 0.0;
        launch_max_abs64(d, B.nc, B.b64, B.part, B.target(1, &sc->dmax), s);
        if (int rc = read_result(B, 1, &sc->dmax, &bmax, s, where)) return rc;
        if (adjoint) tol = rtol * bmax;
        if (bmax_out) *bmax_out = bmax;
    }
    // the forward checks the divergence its stored faces will have; the adjoint's faces are not
    // projected fields (its correction goes on every face)
    const DivCheck chk{u, v, w, occ, o.open_mask, sizeof(F) == 8};
    RefineOut ro;
    if (int rc = refine_mixed(d, o, tol, B, s, nullptr, where, adjoint ? nullptr : &chk, &ro)) return rc;
    double md = ro.dmax;
    if (adjoint) {
        if (o.closed_mean) launch_remove_mean_p64(d, occ, B.p64, B.part, &sc->aux64, s);
        launch_correct_all64<F>(d, B.p64, u, v, w, s);
        md = -1.0;
    } else {
        launch_correct64<F>(d, occ, o.open_mask, B.p64, u, v, w, s);
        if (md < 0.0) {  // the passes ran out before the residual met tol: measure the stored faces
            launch_max_div64<F>(d, u, v, w, B.nc, &sc->mean64, B.part, B.target(1, &sc->dmax), s);
            if (int rc = read_result(B, 1, &sc->dmax, &md, s, where)) return rc;
        }
    }
    ZQ_HIP(hipGetLastError(), where);
    if (st) {
        st->outer = ro.outer;
        st->sweeps = (int)ro.sweeps;
        st->max_residual = ro.rmax;
        st->max_div = md;
        st->mean = 0.0;
        st->converged = adjoint ? -1 : (md <= o.tol ? 1 : 0);
    }
    return ZQ3D_OK;
}

}  // namespace

struct zq3d_adj {
    zq3d_config cfg;
    Dims d;
    int prec = 4;  // bytes per value
    int nsteps = 0, K = 1, ncp = 0;
    hipStream_t stream = nullptr;
    std::vector<float> boxes, spheres, boxes0, spheres0;  // *0: source values zeroed (adjoint)
    double rtol = 0.0;          // adjoint projection tolerance relative to max|rhs|; 0: match the forward
    SorShape shape{SOR_SPLIT, 0};  // sweep algorithm of every projection (forward and adjoint)
    std::vector<double> fwd_rel;  // per step: forward tolerance / max|rhs| of its projection
    std::vector<void*> allocs;
    // state: u, v, w, T, rho, T2, rho2 (value type), vel4 (4-wide), scratch faces (Nu+Nv+Nw)
    void *u, *v, *w, *T, *R, *T2, *R2, *vel4, *scratch;
    void *u0, *v0, *w0, *T0, *R0;  // initial state
    double* p640;                  // initial warm start of the forward projection
    uint8_t* occ;
    // projection
    uint8_t* nc;
    float *b, *p, *p_alt, *split_p, *split_b;
    uint8_t* split_n;
    double *b64, *p64, *p64_adj, *part;
    Scalars* sc;
    Readback rb;  // pinned read-back slots of the projections' reductions
    hipEvent_t gate_ev[2] = {nullptr, nullptr};  // progress checks of gated passes
    // checkpoints: ncp x (u, v, w, T, rho values + p64)
    char* cp;
    size_t cp_bytes;
    // optional store of every forward step's final FP64 pressure (nsteps x Np): a recomputed step
    // then applies the forward's own pressure instead of solving again (same correction, same bits)
    double* pstore = nullptr;
    bool pstore_on = false, pstore_valid = false;
    // tape of one segment: K x (us, uf, up faces; Tf, Rf cells)
    char* tape;
    size_t tape_bytes;
    int tape_seg = -1;
    // adjoint buffers
    void *ub, *vb, *wb, *Tb, *Rb, *Tb2, *Rb2, *cellb, *advb, *gb, *wbar, *cb, *outb, *vort;
    bool forward_done = false, backward_done = false;
    long long generation = 0;  // forwards started; a backward differentiates the latest one
    bool state_final = false;  // the working state holds the last forward's final state
    double last_max_div = 0.0;
};

namespace {

int adj_alloc(zq3d_adj* a, void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return 1;
    }
    a->allocs.push_back(*p);
    return 0;
}

ProjBufs proj_bufs(zq3d_adj* a, double* p64) {
    const size_t sp = split_padded_size(a->d), sc = split_cell_size(a->d);
    const SplitBufs split{{a->split_p, a->split_p + sp},
                          {a->split_b, a->split_b + sc},
                          {a->split_n, a->split_n + sc},
                          {a->split_p + 2 * sp, a->split_p + 3 * sp}};
    return ProjBufs{a->nc, a->b, a->p, a->p_alt, a->b64, p64, a->part, a->sc, a->shape, split, a->rb,
                    {a->gate_ev[0], a->gate_ev[1]}};
}

// pointers into one tape slot
template <class R>
struct Slot {
    R *us, *vs, *ws, *uf, *vf, *wf, *up, *vp, *wp, *Tf, *Rf;
};

template <class R>
Slot<R> tape_slot(zq3d_adj* a, int t) {
    const Sizes z(a->d);
    const size_t faces = z.Nu + z.Nv + z.Nw;
    R* base = (R*)(a->tape + (size_t)t * (3 * faces + 2 * z.N) * sizeof(R));
    Slot<R> sl;
    sl.us = base;
    sl.vs = sl.us + z.Nu;
    sl.ws = sl.vs + z.Nv;
    sl.uf = base + faces;
    sl.vf = sl.uf + z.Nu;
    sl.wf = sl.vf + z.Nv;
    sl.up = base + 2 * faces;
    sl.vp = sl.up + z.Nu;
    sl.wp = sl.vp + z.Nv;
    sl.Tf = base + 3 * faces;
    sl.Rf = sl.Tf + z.N;
    return sl;
}

template <class R>
hipError_t copy_faces(zq3d_adj* a, R* du, R* dv, R* dw, const R* su, const R* sv, const R* sw) {
    const Sizes z(a->d);
    Errs E;
    E(hipMemcpyAsync(du, su, z.Nu * sizeof(R), hipMemcpyDeviceToDevice, a->stream));
    E(hipMemcpyAsync(dv, sv, z.Nv * sizeof(R), hipMemcpyDeviceToDevice, a->stream));
    E(hipMemcpyAsync(dw, sw, z.Nw * sizeof(R), hipMemcpyDeviceToDevice, a->stream));
    return E.e;
}

// checkpoint c: u, v, w, T, rho, then p64
template <class R>
hipError_t checkpoint(zq3d_adj* a, int c, bool save) {
    const Sizes z(a->d);
    char* base = a->cp + (size_t)c * a->cp_bytes;
    const size_t sz[5] = {z.Nu, z.Nv, z.Nw, z.N, z.N};
    void* fld[5] = {a->u, a->v, a->w, a->T, a->R};
    Errs E;
    size_t off = 0;
    for (int f = 0; f < 5; f++) {
        const size_t n = sz[f] * sizeof(R);
        if (save) E(hipMemcpyAsync(base + off, fld[f], n, hipMemcpyDeviceToDevice, a->stream));
        else E(hipMemcpyAsync(fld[f], base + off, n, hipMemcpyDeviceToDevice, a->stream));
        off += n;
    }
    if (save) E(hipMemcpyAsync(base + off, a->p64, z.Np * 8, hipMemcpyDeviceToDevice, a->stream));
    else E(hipMemcpyAsync(a->p64, base + off, z.Np * 8, hipMemcpyDeviceToDevice, a->stream));
    return E.e;
}

AdvParams adv_params(const zq3d_config& c) { return AdvParams{c.voxel_size, c.step_time / c.steps, c.steps}; }

// forward: the rollout's forward (stores the pressure when the store is on); solve: a recomputed
// step solved again; replay: a recomputed step that applies the stored pressure; evaluate: a step of
// a side evaluation (solved, recording nothing the backward reads)
enum class StepMode { forward, solve, replay, evaluate };

template <class R>
int forward_step(zq3d_adj* a, int s, const R* controls, Slot<R>* sl, zq3d_adj_stats* st, StepMode mode) {
    const char* where = "zq3d_adj_forward";
    const zq3d_config& c = a->cfg;
    const Dims& d = a->d;
    const Sizes z(d);
    R *u = (R*)a->u, *v = (R*)a->v, *w = (R*)a->w, *T = (R*)a->T, *Rr = (R*)a->R;
    if (sl) ZQ_HIP(copy_faces<R>(a, sl->us, sl->vs, sl->ws, u, v, w), where);
    const ForceParams fp{c.voxel_size, c.step_time, c.buoyancy, c.gravity, c.confine, c.Tamb};
    const R* ctrl = controls ? controls + (size_t)s * 3 * z.N : nullptr;
    launch_force_atten<R>(d, u, v, w, T, Rr, a->occ, ctrl, fp, 1, c.vel_atten, c.temp_atten, c.dens_atten,
                          (vec4_t<R>*)a->vel4, (R*)a->scratch, a->stream);
    if (sl) {
        ZQ_HIP(copy_faces<R>(a, sl->uf, sl->vf, sl->wf, u, v, w), where);
        ZQ_HIP(hipMemcpyAsync(sl->Tf, T, z.N * sizeof(R), hipMemcpyDeviceToDevice, a->stream), where);
        ZQ_HIP(hipMemcpyAsync(sl->Rf, Rr, z.N * sizeof(R), hipMemcpyDeviceToDevice, a->stream), where);
    }
    const AdvParams ap = adv_params(c);
    launch_advect_velocity<R>(d, u, v, w, a->occ, 1, c.adv_vel_type, ap, (vec4_t<R>*)a->vel4, (R*)a->scratch,
                              a->stream);
    zq3d_proj_stats ps{};
    if (mode == StepMode::replay) {
        // the correction the forward applied, from the forward's own final pressure of this step
        launch_correct64<R>(d, a->occ, c.sor.open_mask, a->pstore + (size_t)s * z.Np, u, v, w, a->stream);
    } else {
        double bmax = 0.0;
        if (int rc = project_mp<R>(d, u, v, w, a->occ, c.sor, proj_bufs(a, a->p64), false, 0.0, &ps, a->stream,
                                   &bmax))
            return rc;
        if (mode != StepMode::evaluate) a->fwd_rel[s] = bmax > 0.0 ? c.sor.tol / bmax : 1.0;
        if (mode == StepMode::forward && a->pstore_on)
            ZQ_HIP(hipMemcpyAsync(a->pstore + (size_t)s * z.Np, a->p64, z.Np * 8, hipMemcpyDeviceToDevice, a->stream),
                   where);
    }
    if (sl) ZQ_HIP(copy_faces<R>(a, sl->up, sl->vp, sl->wp, u, v, w), where);
    launch_advect_scalar<R>(d, u, v, w, a->occ, T, Rr, (R*)a->T2, (R*)a->R2, c.adv_sca_type == 0 ? 1 : 0, ap,
                            (vec4_t<R>*)a->vel4, a->stream);
    std::swap(a->T, a->T2);
    std::swap(a->R, a->R2);
    for (size_t i = 0; i + 8 <= a->boxes.size(); i += 8)
        launch_reinject_box<R>(d, a->boxes.data() + i, (R*)a->T, (R*)a->R, a->stream);
    for (size_t i = 0; i + 6 <= a->spheres.size(); i += 6)
        launch_reinject_sphere<R>(d, a->spheres.data() + i, (R*)a->T, (R*)a->R, a->stream);
    ZQ_HIP(hipGetLastError(), where);
    if (st) {
        st->forward_sweeps += ps.sweeps;
        st->max_div = std::max(st->max_div, ps.max_div);
    }
    return ZQ3D_OK;
}

template <class R>
int adjoint_step(zq3d_adj* a, int s, const Slot<R>& sl, R* controls_bar, zq3d_adj_stats* st) {
    const char* where = "zq3d_adj_backward";
    const zq3d_config& c = a->cfg;
    const Dims& d = a->d;
    const Sizes z(d);
    R *ub = (R*)a->ub, *vb = (R*)a->vb, *wb = (R*)a->wb;
    const AdvParams ap = adv_params(c);
    vec4_t<R>* vel4 = (vec4_t<R>*)a->vel4;
    // reinjection: the source cells were overwritten
    for (size_t i = 0; i + 8 <= a->boxes0.size(); i += 8)
        launch_reinject_box<R>(d, a->boxes0.data() + i, (R*)a->Tb, (R*)a->Rb, a->stream);
    for (size_t i = 0; i + 6 <= a->spheres0.size(); i += 6)
        launch_reinject_sphere<R>(d, a->spheres0.data() + i, (R*)a->Tb, (R*)a->Rb, a->stream);
    // scalar advection through u_p
    const int mac = c.adv_sca_type == 0 ? 1 : 0;
    if (!mac) launch_mac_to_cell4<R>(d, sl.up, sl.vp, sl.wp, vel4, a->stream);
    launch_adj_advect_scalar<R>(d, sl.up, sl.vp, sl.wp, ub, vb, wb, a->occ, sl.Tf, sl.Rf, (R*)a->Tb, (R*)a->Rb,
                                (R*)a->Tb2, (R*)a->Rb2, mac, ap, vel4, (R*)a->cellb, a->stream);
    std::swap(a->Tb, a->Tb2);
    std::swap(a->Rb, a->Rb2);
    // projection
    zq3d_proj_stats ps{};
    // the same relative accuracy as the forward projection of this step, unless set explicitly
    const double rtol = a->rtol > 0.0 ? a->rtol : std::min(1.0, a->fwd_rel[s]);
    if (int rc = project_mp<R>(d, ub, vb, wb, a->occ, c.sor, proj_bufs(a, a->p64_adj), true, rtol, &ps, a->stream,
                               nullptr))
        return rc;
    // velocity advection of u_f
    if (c.adv_vel_type != 2) launch_mac_to_cell4<R>(d, sl.uf, sl.vf, sl.wf, vel4, a->stream);
    launch_adj_advect_velocity<R>(d, sl.uf, sl.vf, sl.wf, ub, vb, wb, a->occ, 1, c.adv_vel_type, ap, vel4,
                                  (R*)a->cellb, (R*)a->advb, (R*)a->outb, a->stream);
    // force + attenuation of u_s
    const ForceParams fp{c.voxel_size, c.step_time, c.buoyancy, c.gravity, c.confine, c.Tamb};
    launch_vorticity<R>(d, sl.us, sl.vs, sl.ws, (vec4_t<R>*)a->vort, c.voxel_size, a->stream);
    launch_adj_force<R>(d, ub, vb, wb, (R*)a->Tb, (R*)a->Rb, controls_bar + (size_t)s * 3 * z.N, a->occ,
                        (const vec4_t<R>*)a->vort, fp, c.vel_atten, c.temp_atten, c.dens_atten, (R*)a->gb,
                        (R*)a->wbar, (R*)a->cb, a->stream);
    ZQ_HIP(hipGetLastError(), where);
    if (st) st->adjoint_sweeps += ps.sweeps;
    return ZQ3D_OK;
}

template <class R>
int run_forward(zq3d_adj* a, const R* controls, R* rho_out, zq3d_adj_stats* st) {
    const char* where = "zq3d_adj_forward";
    const Sizes z(a->d);
    ZQ_HIP(copy_faces<R>(a, (R*)a->u, (R*)a->v, (R*)a->w, (const R*)a->u0, (const R*)a->v0, (const R*)a->w0), where);
    ZQ_HIP(hipMemcpyAsync(a->T, a->T0, z.N * sizeof(R), hipMemcpyDeviceToDevice, a->stream), where);
    ZQ_HIP(hipMemcpyAsync(a->R, a->R0, z.N * sizeof(R), hipMemcpyDeviceToDevice, a->stream), where);
    ZQ_HIP(hipMemcpyAsync(a->p64, a->p640, z.Np * 8, hipMemcpyDeviceToDevice, a->stream), where);
    const int last_seg = a->ncp - 1;
    a->pstore_valid = false;
    for (int s = 0; s < a->nsteps; s++) {
        if (s % a->K == 0) ZQ_HIP(checkpoint<R>(a, s / a->K, true), where);
        // the last segment's tape is recorded now, so the backward need not recompute it
        Slot<R> sl;
        const bool rec = s / a->K == last_seg;
        if (rec) sl = tape_slot<R>(a, s - last_seg * a->K);
        if (int rc = forward_step<R>(a, s, controls, rec ? &sl : nullptr, st, StepMode::forward)) return rc;
    }
    a->tape_seg = last_seg;
    a->pstore_valid = a->pstore_on;
    if (rho_out) ZQ_HIP(hipMemcpyAsync(rho_out, a->R, z.N * sizeof(R), hipMemcpyDeviceToDevice, a->stream), where);
    return ZQ3D_OK;
}

template <class R>
int run_backward(zq3d_adj* a, const R* controls, const R* rho_bar, R* controls_bar, zq3d_adj_stats* st) {
    const char* where = "zq3d_adj_backward";
    const Sizes z(a->d);
    launch_fill_adj<R>((R*)a->ub, z.Nu, a->stream);
    launch_fill_adj<R>((R*)a->vb, z.Nv, a->stream);
    launch_fill_adj<R>((R*)a->wb, z.Nw, a->stream);
    launch_fill_adj<R>((R*)a->Tb, z.N, a->stream);
    ZQ_HIP(hipMemcpyAsync(a->Rb, rho_bar, z.N * sizeof(R), hipMemcpyDeviceToDevice, a->stream), where);
    launch_fill_f64(a->p64_adj, z.Np, 0.0, a->stream);
    for (int seg = a->ncp - 1; seg >= 0; seg--) {
        const int s0 = seg * a->K, s1 = std::min(a->nsteps, s0 + a->K);
        if (a->tape_seg != seg) {
            ZQ_HIP(checkpoint<R>(a, seg, false), where);
            const StepMode mode = a->pstore_on && a->pstore_valid ? StepMode::replay : StepMode::solve;
            for (int s = s0; s < s1; s++) {
                Slot<R> sl = tape_slot<R>(a, s - s0);
                zq3d_adj_stats tmp{};
                if (int rc = forward_step<R>(a, s, controls, &sl, &tmp, mode)) return rc;
                if (st) {
                    st->forward_sweeps += tmp.forward_sweeps;
                    st->recomputed_steps++;
                }
            }
            a->tape_seg = seg;
        }
        for (int s = s1 - 1; s >= s0; s--)
            if (int rc = adjoint_step<R>(a, s, tape_slot<R>(a, s - s0), controls_bar, st)) return rc;
    }
    return ZQ3D_OK;
}

// The final density for `controls` from the latest forward's checkpoint at or before step `from`
// (the steps before it are that forward's); the checkpoints, tape and pressure store are kept.
template <class R>
int run_evaluate(zq3d_adj* a, int from, const R* controls, R* rho_out, zq3d_adj_stats* st) {
    const char* where = "zq3d_adj_evaluate";
    const Sizes z(a->d);
    const int c = from / a->K;
    ZQ_HIP(checkpoint<R>(a, c, false), where);
    for (int s = c * a->K; s < a->nsteps; s++)
        if (int rc = forward_step<R>(a, s, controls, nullptr, st, StepMode::evaluate)) return rc;
    if (rho_out) ZQ_HIP(hipMemcpyAsync(rho_out, a->R, z.N * sizeof(R), hipMemcpyDeviceToDevice, a->stream), where);
    return ZQ3D_OK;
}

struct CallTimer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s;
    explicit CallTimer(hipStream_t s_) : s(s_) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, s);
    }
    double stop() {
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return ms;
    }
    ~CallTimer() {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
};

}  // namespace

extern "C" int zq3d_adj_create(const zq3d_config* cfg, int precision, int nsteps, int checkpoint_every,
                               zq3d_adj** out) {
    const char* where = "zq3d_adj_create";
    if (!cfg || !out) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    *out = nullptr;
    if (cfg->struct_size != (int)sizeof(zq3d_config)) { set_err(where, "zq3d_config.struct_size mismatch"); return ZQ3D_ERR_ARG; }
    if (int rc = check_dims(cfg->width, cfg->height, cfg->depth, where)) return rc;
    if (int rc = check_sor_opts(cfg->sor, where)) return rc;
    if (precision != 4 && precision != 8) { set_err(where, "precision must be 4 or 8"); return ZQ3D_ERR_ARG; }
    if (nsteps < 1 || checkpoint_every < 1) { set_err(where, "nsteps and checkpoint_every must be >= 1"); return ZQ3D_ERR_ARG; }
    if (!cfg->sor.mixed_precision) { set_err(where, "the adjoint rollout needs sor.mixed_precision"); return ZQ3D_ERR_ARG; }
    if (cfg->steps < 1 || cfg->steps > ZQ3D_ADJ_MAX_STEPS) { set_err(where, "steps must be in [1, 32]"); return ZQ3D_ERR_ARG; }
    if (cfg->adv_vel_type != 0 && cfg->adv_vel_type != 2 && cfg->adv_vel_type != 4) {
        set_err(where, "adv_vel_type must be a non-BFECC mode");
        return ZQ3D_ERR_ARG;
    }
    if (cfg->adv_sca_type != 0 && cfg->adv_sca_type != 2) { set_err(where, "adv_sca_type must be MAC or Reg"); return ZQ3D_ERR_ARG; }
    if (!(cfg->voxel_size > 0.0f)) { set_err(where, "voxel_size must be > 0"); return ZQ3D_ERR_ARG; }
    zq3d_adj* a = new zq3d_adj();
    a->cfg = *cfg;
    a->d = Dims{cfg->width, cfg->height, cfg->depth};
    a->prec = precision;
    a->nsteps = nsteps;
    a->K = std::min(checkpoint_every, nsteps);
    a->fwd_rel.assign(nsteps, 1.0);
    a->ncp = (nsteps + a->K - 1) / a->K;
    const Sizes z(a->d);
    const size_t e = (size_t)precision;
    const size_t faces = z.Nu + z.Nv + z.Nw;
    int bad = 0;
    auto A = [&](void** p, size_t bytes) { bad |= adj_alloc(a, p, bytes); };
    A(&a->u, z.Nu * e); A(&a->v, z.Nv * e); A(&a->w, z.Nw * e);
    A(&a->T, z.N * e); A(&a->R, z.N * e); A(&a->T2, z.N * e); A(&a->R2, z.N * e);
    A(&a->vel4, z.N * 4 * e); A(&a->scratch, faces * e);
    A(&a->u0, z.Nu * e); A(&a->v0, z.Nv * e); A(&a->w0, z.Nw * e); A(&a->T0, z.N * e); A(&a->R0, z.N * e);
    A((void**)&a->occ, z.N); A((void**)&a->nc, z.N);
    A((void**)&a->b, z.N * 4); A((void**)&a->p, z.Np * 4); A((void**)&a->p_alt, z.Np * 4);
    A((void**)&a->split_p, 4 * split_padded_size(a->d) * 4);
    A((void**)&a->split_b, 2 * split_cell_size(a->d) * 4);
    A((void**)&a->split_n, 2 * split_cell_size(a->d));
    A((void**)&a->b64, z.N * 8); A((void**)&a->p64, z.Np * 8); A((void**)&a->p64_adj, z.Np * 8);
    A((void**)&a->p640, z.Np * 8);
    A((void**)&a->part, (size_t)reduction_partials(a->d) * 8);
    A((void**)&a->sc, sizeof(Scalars));
    bad |= a->rb.alloc() != hipSuccess;
    bad |= hipEventCreateWithFlags(&a->gate_ev[0], hipEventDisableTiming) != hipSuccess;
    bad |= hipEventCreateWithFlags(&a->gate_ev[1], hipEventDisableTiming) != hipSuccess;
    a->cp_bytes = (faces + 2 * z.N) * e + z.Np * 8;
    a->cp_bytes = (a->cp_bytes + 255) & ~(size_t)255;
    A((void**)&a->cp, a->cp_bytes * a->ncp);
    a->tape_bytes = (3 * faces + 2 * z.N) * e * a->K;
    A((void**)&a->tape, a->tape_bytes);
    A(&a->ub, z.Nu * e); A(&a->vb, z.Nv * e); A(&a->wb, z.Nw * e);
    A(&a->Tb, z.N * e); A(&a->Rb, z.N * e); A(&a->Tb2, z.N * e); A(&a->Rb2, z.N * e);
    A(&a->cellb, 3 * z.N * e); A(&a->advb, 3 * z.N * e); A(&a->gb, 3 * z.N * e); A(&a->wbar, 3 * z.N * e);
    A(&a->cb, 3 * z.N * e); A(&a->outb, faces * e); A(&a->vort, 4 * z.N * e);
    if (bad) {
        set_err(where, "device allocation failed");
        zq3d_adj_destroy(a);
        return ZQ3D_ERR_HIP;
    }
    hipStream_t s = nullptr;
    (void)hipMemsetAsync(a->u0, 0, z.Nu * e, s);
    (void)hipMemsetAsync(a->v0, 0, z.Nv * e, s);
    (void)hipMemsetAsync(a->w0, 0, z.Nw * e, s);
    (void)hipMemsetAsync(a->T0, 0, z.N * e, s);
    (void)hipMemsetAsync(a->R0, 0, z.N * e, s);
    (void)hipMemsetAsync(a->occ, 0, z.N, s);
    (void)hipMemsetAsync(a->p, 0, z.Np * 4, s);
    (void)hipMemsetAsync(a->p_alt, 0, z.Np * 4, s);
    (void)hipMemsetAsync(a->split_p, 0, 4 * split_padded_size(a->d) * 4, s);
    (void)hipMemsetAsync(a->p64, 0, z.Np * 8, s);
    (void)hipMemsetAsync(a->p64_adj, 0, z.Np * 8, s);
    (void)hipMemsetAsync(a->p640, 0, z.Np * 8, s);
    if (hipStreamSynchronize(s) != hipSuccess) {
        set_err(where, "initialisation failed");
        zq3d_adj_destroy(a);
        return ZQ3D_ERR_HIP;
    }
    *out = a;
    return ZQ3D_OK;
}

extern "C" void zq3d_adj_destroy(zq3d_adj* a) {
    if (!a) return;
    for (void* p : a->allocs) (void)hipFree(p);
    a->rb.release();
    for (hipEvent_t e : {a->gate_ev[0], a->gate_ev[1]})
        if (e) (void)hipEventDestroy(e);
    delete a;
}

extern "C" int zq3d_adj_set_stream(zq3d_adj* a, void* stream) {
    if (!a) { set_err("zq3d_adj_set_stream", "null rollout"); return ZQ3D_ERR_ARG; }
    a->stream = (hipStream_t)stream;
    return ZQ3D_OK;
}

extern "C" int zq3d_adj_set_sources(zq3d_adj* a, int n_box, const float* boxes8, int n_sphere, const float* sph6) {
    if (!a || n_box < 0 || n_sphere < 0 || (n_box && !boxes8) || (n_sphere && !sph6)) {
        set_err("zq3d_adj_set_sources", "bad argument");
        return ZQ3D_ERR_ARG;
    }
    a->boxes.assign(boxes8, boxes8 + 8 * n_box);
    a->pstore_valid = false;
    a->spheres.assign(sph6, sph6 + 6 * n_sphere);
    a->boxes0 = a->boxes;
    a->spheres0 = a->spheres;
    for (int i = 0; i < n_box; i++) a->boxes0[8 * i + 6] = a->boxes0[8 * i + 7] = 0.0f;
    for (int i = 0; i < n_sphere; i++) a->spheres0[6 * i + 4] = a->spheres0[6 * i + 5] = 0.0f;
    return ZQ3D_OK;
}

extern "C" int zq3d_adj_set_adjoint_tol(zq3d_adj* a, double rtol) {
    if (!a || !(rtol >= 0.0)) { set_err("zq3d_adj_set_adjoint_tol", "bad argument"); return ZQ3D_ERR_ARG; }
    a->rtol = rtol;
    return ZQ3D_OK;
}

extern "C" int zq3d_adj_set_pressure_store(zq3d_adj* a, int on) {
    const char* where = "zq3d_adj_set_pressure_store";
    if (!a) { set_err(where, "null handle"); return ZQ3D_ERR_ARG; }
    if (on && !a->pstore) {
        const Sizes z(a->d);
        if (adj_alloc(a, (void**)&a->pstore, (size_t)a->nsteps * z.Np * 8)) {
            set_err(where, "device allocation failed (nsteps x the padded FP64 pressure)");
            return ZQ3D_ERR_HIP;
        }
    }
    a->pstore_on = on != 0;
    a->pstore_valid = false;  // filled by the next forward
    return ZQ3D_OK;
}

extern "C" int zq3d_adj_set_sor_shape(zq3d_adj* a, int ty, int zc) {
    const char* where = "zq3d_adj_set_sor_shape";
    if (!a) { set_err(where, "null handle"); return ZQ3D_ERR_ARG; }
    SorShape cand[ZQ3D_TUNE_MAX];
    const int n = sor_shape_candidates(a->d, cand, ZQ3D_TUNE_MAX);
    bool ok = false;
    for (int i = 0; i < n; i++) ok = ok || (cand[i].ty == ty && cand[i].zc == zc);
    if (!ok) { set_err(where, "shape not among the sweep candidates for this grid"); return ZQ3D_ERR_ARG; }
    a->shape = SorShape{ty, zc};
    return ZQ3D_OK;
}

namespace {

int field_ptr(zq3d_adj* a, int field, bool initial, bool grad, void** p, size_t* bytes) {
    const Sizes z(a->d);
    const size_t e = (size_t)a->prec;
    switch (field) {
    case ZQ3D_FIELD_U: *p = grad ? a->ub : (initial ? a->u0 : a->u); *bytes = z.Nu * e; return 0;
    case ZQ3D_FIELD_V: *p = grad ? a->vb : (initial ? a->v0 : a->v); *bytes = z.Nv * e; return 0;
    case ZQ3D_FIELD_W: *p = grad ? a->wb : (initial ? a->w0 : a->w); *bytes = z.Nw * e; return 0;
    case ZQ3D_FIELD_T: *p = grad ? a->Tb : (initial ? a->T0 : a->T); *bytes = z.N * e; return 0;
    case ZQ3D_FIELD_RHO: *p = grad ? a->Rb : (initial ? a->R0 : a->R); *bytes = z.N * e; return 0;
    case ZQ3D_FIELD_OCC: if (grad) return 1; *p = a->occ; *bytes = z.N; return 0;
    case ZQ3D_FIELD_P64: if (grad || !initial) return 1; *p = a->p640; *bytes = z.Np * 8; return 0;
    default: return 1;
    }
}

}  // namespace

extern "C" int zq3d_adj_set_initial(zq3d_adj* a, int field, const void* dev) {
    const char* where = "zq3d_adj_set_initial";
    if (!a || !dev) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    void* p;
    size_t n;
    if (field_ptr(a, field, true, false, &p, &n)) { set_err(where, "unknown field"); return ZQ3D_ERR_ARG; }
    ZQ_HIP(hipMemcpyAsync(p, dev, n, hipMemcpyDeviceToDevice, a->stream), where);
    a->forward_done = a->backward_done = false;
    a->pstore_valid = false;
    a->tape_seg = -1;
    return ZQ3D_OK;
}

extern "C" int zq3d_adj_forward(zq3d_adj* a, const void* controls, zq3d_adj_stats* st, void* rho_out) {
    const char* where = "zq3d_adj_forward";
    if (!a) { set_err(where, "null rollout"); return ZQ3D_ERR_ARG; }
    zq3d_adj_stats local{};
    a->generation++;
    CallTimer t(a->stream);
    int rc = a->prec == 4 ? run_forward<float>(a, (const float*)controls, (float*)rho_out, &local)
                          : run_forward<double>(a, (const double*)controls, (double*)rho_out, &local);
    local.ms = t.stop();
    a->forward_done = rc == ZQ3D_OK;
    a->state_final = a->forward_done;
    a->backward_done = false;
    if (st) *st = local;
    return rc;
}

extern "C" int zq3d_adj_backward(zq3d_adj* a, const void* controls, const void* rho_bar, void* controls_bar,
                                 zq3d_adj_stats* st) {
    const char* where = "zq3d_adj_backward";
    if (!a || !rho_bar || !controls_bar) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    if (!a->forward_done) { set_err(where, "no forward to differentiate"); return ZQ3D_ERR_STATE; }
    zq3d_adj_stats local{};
    CallTimer t(a->stream);
    int rc = a->prec == 4
                 ? run_backward<float>(a, (const float*)controls, (const float*)rho_bar, (float*)controls_bar, &local)
                 : run_backward<double>(a, (const double*)controls, (const double*)rho_bar, (double*)controls_bar,
                                        &local);
    local.ms = t.stop();
    a->backward_done = rc == ZQ3D_OK;
    if (local.recomputed_steps) a->state_final = false;
    if (st) *st = local;
    return rc;
}

extern "C" int zq3d_adj_evaluate(zq3d_adj* a, int from_step, const void* controls, zq3d_adj_stats* st,
                                 void* rho_out) {
    const char* where = "zq3d_adj_evaluate";
    if (!a || !rho_out) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    if (!a->forward_done) { set_err(where, "no forward whose checkpoints to start from"); return ZQ3D_ERR_STATE; }
    if (from_step < 0 || from_step >= a->nsteps) { set_err(where, "from_step out of range"); return ZQ3D_ERR_ARG; }
    zq3d_adj_stats local{};
    CallTimer t(a->stream);
    int rc = a->prec == 4 ? run_evaluate<float>(a, from_step, (const float*)controls, (float*)rho_out, &local)
                          : run_evaluate<double>(a, from_step, (const double*)controls, (double*)rho_out, &local);
    local.ms = t.stop();
    a->state_final = false;  // the working state is this evaluation's
    if (st) *st = local;
    return rc;
}

extern "C" long long zq3d_adj_generation(zq3d_adj* a) { return a ? a->generation : -1; }

extern "C" int zq3d_adj_get_tape(zq3d_adj* a, int step, int field, void* dev) {
    const char* where = "zq3d_adj_get_tape";
    if (!a || !dev) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    if (a->tape_seg < 0) { set_err(where, "nothing taped yet"); return ZQ3D_ERR_STATE; }
    const int s0 = a->tape_seg * a->K, s1 = std::min(s0 + a->K, a->nsteps);
    if (step < s0 || step >= s1) { set_err(where, "step is not in the taped segment"); return ZQ3D_ERR_ARG; }
    const Sizes z(a->d);
    const size_t e = (size_t)a->prec;
    const size_t faces = z.Nu + z.Nv + z.Nw;
    const char* base = a->tape + (size_t)(step - s0) * (3 * faces + 2 * z.N) * e;
    size_t off, n;
    switch (field) {
    case ZQ3D_FIELD_U: off = 2 * faces; n = z.Nu; break;
    case ZQ3D_FIELD_V: off = 2 * faces + z.Nu; n = z.Nv; break;
    case ZQ3D_FIELD_W: off = 2 * faces + z.Nu + z.Nv; n = z.Nw; break;
    case ZQ3D_FIELD_T: off = 3 * faces; n = z.N; break;
    case ZQ3D_FIELD_RHO: off = 3 * faces + z.N; n = z.N; break;
    default: set_err(where, "unknown field"); return ZQ3D_ERR_ARG;
    }
    ZQ_HIP(hipMemcpyAsync(dev, base + off * e, n * e, hipMemcpyDeviceToDevice, a->stream), where);
    return ZQ3D_OK;
}

extern "C" int zq3d_adj_get(zq3d_adj* a, int field, int grad, void* dev) {
    const char* where = "zq3d_adj_get";
    if (!a || !dev) { set_err(where, "null argument"); return ZQ3D_ERR_ARG; }
    if (grad ? !a->backward_done : !a->state_final) {
        set_err(where, grad ? "no backward yet" : "no final state (none computed, or overwritten by a backward)");
        return ZQ3D_ERR_STATE;
    }
    void* p;
    size_t n;
    if (field_ptr(a, field, false, grad != 0, &p, &n)) { set_err(where, "unknown field"); return ZQ3D_ERR_ARG; }
    ZQ_HIP(hipMemcpyAsync(dev, p, n, hipMemcpyDeviceToDevice, a->stream), where);
    ZQ_HIP(hipStreamSynchronize(a->stream), where);
    return ZQ3D_OK;
}
