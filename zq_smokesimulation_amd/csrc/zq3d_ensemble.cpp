// zq3d_ensemble.cpp -- Part 6 of the C ABI: an ensemble of independent smoke grids of uneven sizes
// stepped together, one persistent launch per stage (kernels_ensemble.hip), the whole step
// captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "zq3d_ensemble.h"
#include "zq3d_internal.h"
#include "zq3d_kernels.h"
#include "zq_smoke3d.h"

using namespace zq3d;
using namespace zq3d::detail;

struct zq3d_ens {
    int count = 0;
    std::vector<zq3d_config> cfg;
    std::vector<Dims> dims;
    std::vector<EnsMember> host;  // descriptors as uploaded
    EnsMember* dev_members = nullptr;
    int* dev_tables = nullptr;
    std::vector<void*> allocs;
    hipStream_t stream = nullptr;
    EnsLaunch L{};
    // pool offsets per member (elements)
    std::vector<size_t> off_u, off_v, off_w, off_c, off_p, off_part, off_sc, off_sp;
    float *u = nullptr, *v = nullptr, *w = nullptr, *T = nullptr, *R = nullptr, *T2 = nullptr, *R2 = nullptr;
    uint8_t *occ = nullptr, *nc = nullptr;
    float4* vel4 = nullptr;
    float *scratch = nullptr, *b = nullptr, *p = nullptr, *mean32 = nullptr;
    double* part = nullptr;
    uint8_t* split_n = nullptr;  // colour-split counts (2 per member), right-hand sides, pressures (4)
    float *split_b = nullptr, *split_p = nullptr;
    size_t Nu = 0, Nv = 0, Nw = 0, Nc = 0, Np = 0, Nparts = 0, Nsc = 0, Nsp = 0;
    long long cells = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_stale = true;
};

namespace {

int cdiv(int a, int b) { return (a + b - 1) / b; }

// whether the one-launch split sweeps can serve every member (split_buffers_fit)
bool rb_fits(const zq3d_ens* e) {
    for (const Dims& d : e->dims)
        if (!split_buffers_fit(d)) return false;
    return true;
}

void drop_graph(zq3d_ens* e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    e->graph_stale = true;
}

int check_member(const zq3d_config& c, const char* where) {
    if (c.struct_size != (int)sizeof(zq3d_config)) { set_err(where, "zq3d_config.struct_size mismatch"); return ZQ3D_ERR_ARG; }
    if (int rc = check_dims(c.width, c.height, c.depth, where)) return rc;
    if (int rc = check_sor_opts(c.sor, where)) return rc;
    if (c.steps < 1 || !(c.voxel_size > 0.0f)) { set_err(where, "steps >= 1 and voxel_size > 0 required"); return ZQ3D_ERR_ARG; }
    if (c.adv_vel_type != 0 || c.adv_sca_type != 2) {
        set_err(where, "ensemble members use ADV_VEL_INREG_OUTREG and ADV_SCA_REG");
        return ZQ3D_ERR_ARG;
    }
    if (c.sor.mixed_precision) {
        set_err(where, "ensemble members use the plain FP32 projection (a fixed sweep count, so a step can be captured)");
        return ZQ3D_ERR_ARG;
    }
    return ZQ3D_OK;
}

template <class T>
int ens_alloc(zq3d_ens* e, T** p, size_t bytes) {
    void* q = nullptr;
    if (hipMalloc(&q, bytes ? bytes : 1) != hipSuccess) return 1;
    e->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

// tables: cells, u/v/w faces, and the half-row grid per sweep shape; each as start[count+1], gx, gy
void build_tables(zq3d_ens* e, std::vector<int>& flat, std::vector<size_t>& at) {
    const int B = e->count;
    auto add = [&](auto grid_of) {
        at.push_back(flat.size());
        std::vector<int> start(B + 1), gx(B), gy(B);
        int total = 0;
        for (int s = 0; s < B; s++) {
            int x, y, z;
            grid_of(e->dims[s], x, y, z);
            start[s] = total;
            gx[s] = x;
            gy[s] = y;
            total += x * y * z;
        }
        start[B] = total;
        flat.insert(flat.end(), start.begin(), start.end());
        flat.insert(flat.end(), gx.begin(), gx.end());
        flat.insert(flat.end(), gy.begin(), gy.end());
    };
    add([](const Dims& d, int& x, int& y, int& z) { x = cdiv(d.W, 64), y = cdiv(d.H, 4), z = d.D; });
    add([](const Dims& d, int& x, int& y, int& z) { x = cdiv(d.W + 1, 64), y = cdiv(d.H, 4), z = d.D; });
    add([](const Dims& d, int& x, int& y, int& z) { x = cdiv(d.W, 64), y = cdiv(d.H + 1, 4), z = d.D; });
    add([](const Dims& d, int& x, int& y, int& z) { x = cdiv(d.W, 64), y = cdiv(d.H, 4), z = d.D + 1; });
    for (int sh = 0; sh < kEnsSweepShapeCount; sh++) {
        const int tbx = kEnsSweepShapes[sh][0], tby = kEnsSweepShapes[sh][1];
        add([tbx, tby](const Dims& d, int& x, int& y, int& z) { x = cdiv((d.W + 1) / 2, tbx), y = cdiv(d.H, tby), z = d.D; });
    }
    for (int r = 0; r < kEnsRbShapeCount; r++) {
        const int tx = kEnsRbShapes[r][0], ty = kEnsRbShapes[r][1], zc = kEnsRbShapes[r][2];
        add([tx, ty, zc](const Dims& d, int& x, int& y, int& z) {
            x = cdiv((d.W + 1) / 2, tx), y = cdiv(d.H, ty), z = cdiv(d.D, zc);
        });
    }
}

EnsTable table_at(const zq3d_ens* e, const std::vector<int>& flat, size_t at) {
    const int B = e->count;
    return EnsTable{e->dev_tables + at, e->dev_tables + at + B + 1, e->dev_tables + at + 2 * B + 1, flat[at + B]};
}

int field_span(zq3d_ens* e, int s, int field, void** p, size_t* bytes) {
    const Sizes z(e->dims[s]);
    switch (field) {
    case ZQ3D_FIELD_U: *p = e->u + e->off_u[s]; *bytes = z.Nu * 4; return 0;
    case ZQ3D_FIELD_V: *p = e->v + e->off_v[s]; *bytes = z.Nv * 4; return 0;
    case ZQ3D_FIELD_W: *p = e->w + e->off_w[s]; *bytes = z.Nw * 4; return 0;
    case ZQ3D_FIELD_T: *p = e->T + e->off_c[s]; *bytes = z.N * 4; return 0;
    case ZQ3D_FIELD_RHO: *p = e->R + e->off_c[s]; *bytes = z.N * 4; return 0;
    case ZQ3D_FIELD_OCC: *p = e->occ + e->off_c[s]; *bytes = z.N; return 0;
    default: return 1;
    }
}

int upload_members(zq3d_ens* e, const char* where) {
    ZQ_HIP(hipMemcpyAsync(e->dev_members, e->host.data(), sizeof(EnsMember) * e->count, hipMemcpyHostToDevice,
                          e->stream),
           where);
    ZQ_HIP(hipStreamSynchronize(e->stream), where);
    return ZQ3D_OK;
}

}  // namespace

extern "C" int zq3d_ens_create(int count, const zq3d_config* cfgs, zq3d_ens** out) {
    const char* where = "zq3d_ens_create";
    if (!out || !cfgs || count < 1) { set_err(where, "bad argument"); return ZQ3D_ERR_ARG; }
    *out = nullptr;
    for (int s = 0; s < count; s++)
        if (int rc = check_member(cfgs[s], where)) return rc;
    zq3d_ens* e = new zq3d_ens();
    e->count = count;
    e->cfg.assign(cfgs, cfgs + count);
    for (int s = 0; s < count; s++) {
        const Dims d{cfgs[s].width, cfgs[s].height, cfgs[s].depth};
        const Sizes z(d);
        e->dims.push_back(d);
        e->off_u.push_back(e->Nu);
        e->off_v.push_back(e->Nv);
        e->off_w.push_back(e->Nw);
        e->off_c.push_back(e->Nc);
        e->off_p.push_back(e->Np);
        e->off_part.push_back(e->Nparts);
        e->off_sc.push_back(e->Nsc);
        e->off_sp.push_back(e->Nsp);
        e->Nsc += 2 * split_cell_size(d);
        e->Nsp += 4 * split_padded_size(d);
        e->Nu += z.Nu;
        e->Nv += z.Nv;
        e->Nw += z.Nw;
        e->Nc += z.N;
        e->Np += z.Np;
        e->Nparts += 2 * (size_t)cdiv(d.W, 64) * cdiv(d.H, 4) * d.D;
        e->cells += (long long)z.N;
    }
    std::vector<int> flat;
    std::vector<size_t> at;
    build_tables(e, flat, at);
    int bad = 0;
    bad |= ens_alloc(e, &e->u, e->Nu * 4) | ens_alloc(e, &e->v, e->Nv * 4) | ens_alloc(e, &e->w, e->Nw * 4);
    bad |= ens_alloc(e, &e->T, e->Nc * 4) | ens_alloc(e, &e->R, e->Nc * 4) | ens_alloc(e, &e->T2, e->Nc * 4) |
           ens_alloc(e, &e->R2, e->Nc * 4);
    bad |= ens_alloc(e, &e->occ, e->Nc) | ens_alloc(e, &e->nc, e->Nc) | ens_alloc(e, &e->vel4, e->Nc * 16);
    bad |= ens_alloc(e, &e->scratch, (e->Nu + e->Nv + e->Nw) * 4) | ens_alloc(e, &e->b, e->Nc * 4) |
           ens_alloc(e, &e->p, e->Np * 4);
    bad |= ens_alloc(e, &e->part, e->Nparts * 8) | ens_alloc(e, &e->mean32, (size_t)count * 4);
    bad |= ens_alloc(e, &e->split_n, e->Nsc) | ens_alloc(e, &e->split_b, e->Nsc * 4) |
           ens_alloc(e, &e->split_p, e->Nsp * 4);
    bad |= ens_alloc(e, &e->dev_members, sizeof(EnsMember) * count) | ens_alloc(e, &e->dev_tables, flat.size() * 4);
    if (bad || hipStreamCreate(&e->stream) != hipSuccess) {
        set_err(where, "device allocation failed");
        zq3d_ens_destroy(e);
        return ZQ3D_ERR_HIP;
    }
    size_t off_scr = 0;
    for (int s = 0; s < count; s++) {
        const zq3d_config& c = cfgs[s];
        const Dims d = e->dims[s];
        const Sizes z(d);
        EnsMember m;
        std::memset(&m, 0, sizeof(m));
        m.d = d;
        m.u = e->u + e->off_u[s];
        m.v = e->v + e->off_v[s];
        m.w = e->w + e->off_w[s];
        m.T = e->T + e->off_c[s];
        m.R = e->R + e->off_c[s];
        m.T2 = e->T2 + e->off_c[s];
        m.R2 = e->R2 + e->off_c[s];
        m.occ = e->occ + e->off_c[s];
        m.nc = e->nc + e->off_c[s];
        m.vel4 = e->vel4 + e->off_c[s];
        m.scratch = e->scratch + off_scr;
        off_scr += z.Nu + z.Nv + z.Nw;
        m.b = e->b + e->off_c[s];
        m.p = e->p + e->off_p[s];
        m.part = e->part + e->off_part[s];
        m.mean32 = e->mean32 + s;
        const size_t sc = split_cell_size(d), sp = split_padded_size(d);
        for (int c = 0; c < 2; c++) {
            m.sn[c] = e->split_n + e->off_sc[s] + c * sc;
            m.sb[c] = e->split_b + e->off_sc[s] + c * sc;
        }
        for (int q = 0; q < 4; q++) m.sp[q] = e->split_p + e->off_sp[s] + q * sp;
        // the single-grid launchers' host arithmetic (launch_force_atten, launch_advect_*, launch_sor)
        m.h = c.voxel_size;
        m.inv_h = step_scale<float>(c.voxel_size).inv;
        m.dt = c.step_time;
        m.dth = c.step_time * 0.5f;
        m.two_h = 2.0f * c.voxel_size;
        m.ch = c.confine * c.voxel_size;
        m.dtt = c.step_time / c.steps;
        m.steps = c.steps;
        m.beta = c.buoyancy;
        m.alpha = c.gravity;
        m.Tamb = c.Tamb;
        m.aV = c.vel_atten;
        m.aT = c.temp_atten;
        m.aR = c.dens_atten;
        m.open_mask = c.sor.open_mask;
        m.closed_mean = c.sor.closed_mean ? 1 : 0;
        m.iters = c.sor.iters;
        m.omega = c.sor.omega;
        m.omw = 1.0f - c.sor.omega;
        e->host.push_back(m);
        e->L.max_iters = std::max(e->L.max_iters, c.sor.iters);
    }
    Errs E;
    E(hipMemcpy(e->dev_tables, flat.data(), flat.size() * 4, hipMemcpyHostToDevice));
    E(hipMemsetAsync(e->u, 0, e->Nu * 4, e->stream));
    E(hipMemsetAsync(e->v, 0, e->Nv * 4, e->stream));
    E(hipMemsetAsync(e->w, 0, e->Nw * 4, e->stream));
    E(hipMemsetAsync(e->T, 0, e->Nc * 4, e->stream));
    E(hipMemsetAsync(e->R, 0, e->Nc * 4, e->stream));
    E(hipMemsetAsync(e->occ, 0, e->Nc, e->stream));
    E(hipMemsetAsync(e->p, 0, e->Np * 4, e->stream));
    E(hipMemsetAsync(e->nc, 0, e->Nc, e->stream));
    E(hipMemsetAsync(e->b, 0, e->Nc * 4, e->stream));
    // slots past a row's end and the padding of the split arrays stay 0 from here on
    E(hipMemsetAsync(e->split_n, 0, e->Nsc, e->stream));
    E(hipMemsetAsync(e->split_b, 0, e->Nsc * 4, e->stream));
    E(hipMemsetAsync(e->split_p, 0, e->Nsp * 4, e->stream));
    if (E.e != hipSuccess) {
        set_err(where, hipGetErrorString(E.e));
        zq3d_ens_destroy(e);
        return ZQ3D_ERR_HIP;
    }
    if (int rc = upload_members(e, where)) {
        zq3d_ens_destroy(e);
        return rc;
    }
    EnsLaunch& L = e->L;
    L.members = e->dev_members;
    L.count = count;
    L.cells = table_at(e, flat, at[0]);
    L.ufaces = table_at(e, flat, at[1]);
    L.vfaces = table_at(e, flat, at[2]);
    L.wfaces = table_at(e, flat, at[3]);
    for (int sh = 0; sh < kEnsSweepShapeCount; sh++) L.half[sh] = table_at(e, flat, at[4 + sh]);
    for (int r = 0; r < kEnsRbShapeCount; r++) L.rb[r] = table_at(e, flat, at[4 + kEnsSweepShapeCount + r]);
    L.sweep_shape = 0;
    // one-launch iterations, 64 slots x 16 rows x 16 planes per workgroup, when every member's split
    // arrays are small enough for them; else the half sweeps
    L.sweep_rb = rb_fits(e) ? 4 : 0;
    L.blocks = 0;
    L.sweep_blocks = 0;
    L.p_pool = e->p;
    L.p_pool_size = e->Np;
    L.T_pool = e->T;
    L.T2_pool = e->T2;
    L.R_pool = e->R;
    L.R2_pool = e->R2;
    L.cell_pool_size = e->Nc;
    *out = e;
    return ZQ3D_OK;
}

extern "C" void zq3d_ens_destroy(zq3d_ens* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    for (void* p : e->allocs) (void)hipFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

extern "C" int zq3d_ens_set_sources(zq3d_ens* e, int member, int n_box, const float* boxes8, int n_sphere,
                                    const float* spheres6) {
    const char* where = "zq3d_ens_set_sources";
    if (!e || member < 0 || member >= e->count || n_box < 0 || n_sphere < 0 || n_box > kEnsMaxSources ||
        n_sphere > kEnsMaxSources || (n_box && !boxes8) || (n_sphere && !spheres6)) {
        set_err(where, "bad argument (at most 4 box and 4 sphere sources per member)");
        return ZQ3D_ERR_ARG;
    }
    EnsMember& m = e->host[member];
    const Dims d = e->dims[member];
    m.n_box = n_box;
    m.n_sph = n_sphere;
    for (int q = 0; q < n_box; q++) {
        const float* b = boxes8 + 8 * q;
        // launch_reinject_box: centre and half extents scaled by the grid on the host
        const float v[8] = {b[0] * d.W, b[1] * d.H, b[2] * d.D, b[3] * d.W, b[4] * d.H, b[5] * d.D, b[6], b[7]};
        std::memcpy(m.box[q], v, sizeof(v));
    }
    for (int q = 0; q < n_sphere; q++) {
        const float* sp = spheres6 + 6 * q;
        const float rr = sp[3] * d.W;  // launch_reinject_sphere
        const float v[6] = {sp[0] * d.W, sp[1] * d.H, sp[2] * d.D, rr * rr, sp[4], sp[5]};
        std::memcpy(m.sph[q], v, sizeof(v));
    }
    return upload_members(e, where);
}

extern "C" int zq3d_ens_upload(zq3d_ens* e, int member, int field, const void* host, size_t bytes) {
    const char* where = "zq3d_ens_upload";
    if (!e || !host || member < 0 || member >= e->count) { set_err(where, "bad argument"); return ZQ3D_ERR_ARG; }
    void* p;
    size_t n;
    if (field_span(e, member, field, &p, &n)) { set_err(where, "unknown field"); return ZQ3D_ERR_ARG; }
    if (n != bytes) { set_err(where, "size mismatch"); return ZQ3D_ERR_ARG; }
    ZQ_HIP(hipMemcpyAsync(p, host, n, hipMemcpyHostToDevice, e->stream), where);
    ZQ_HIP(hipStreamSynchronize(e->stream), where);
    return ZQ3D_OK;
}

extern "C" int zq3d_ens_download(zq3d_ens* e, int member, int field, void* host, size_t bytes) {
    const char* where = "zq3d_ens_download";
    if (!e || !host || member < 0 || member >= e->count) { set_err(where, "bad argument"); return ZQ3D_ERR_ARG; }
    void* p;
    size_t n;
    if (field_span(e, member, field, &p, &n)) { set_err(where, "unknown field"); return ZQ3D_ERR_ARG; }
    if (n != bytes) { set_err(where, "size mismatch"); return ZQ3D_ERR_ARG; }
    ZQ_HIP(hipMemcpyAsync(host, p, n, hipMemcpyDeviceToHost, e->stream), where);
    ZQ_HIP(hipStreamSynchronize(e->stream), where);
    return ZQ3D_OK;
}

extern "C" int zq3d_ens_step(zq3d_ens* e, int steps, int use_graph) {
    const char* where = "zq3d_ens_step";
    if (!e || steps < 0) { set_err(where, "bad argument"); return ZQ3D_ERR_ARG; }
    if (!use_graph) {
        for (int n = 0; n < steps; n++) ens_launch_step(e->L, e->stream);
        ZQ_HIP(hipGetLastError(), where);
        return ZQ3D_OK;
    }
    if (e->graph_stale) {
        drop_graph(e);
        ZQ_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), where);
        ens_launch_step(e->L, e->stream);
        const hipError_t cap = hipGetLastError();
        const hipError_t end = hipStreamEndCapture(e->stream, &e->graph);
        ZQ_HIP(cap, where);
        ZQ_HIP(end, where);
        ZQ_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0), where);
        e->graph_stale = false;
    }
    for (int n = 0; n < steps; n++) ZQ_HIP(hipGraphLaunch(e->exec, e->stream), where);
    return ZQ3D_OK;
}

extern "C" int zq3d_ens_synchronize(zq3d_ens* e) {
    if (!e) { set_err("zq3d_ens_synchronize", "null ensemble"); return ZQ3D_ERR_ARG; }
    ZQ_HIP(hipStreamSynchronize(e->stream), "zq3d_ens_synchronize");
    return ZQ3D_OK;
}

extern "C" long long zq3d_ens_cells(zq3d_ens* e) { return e ? e->cells : 0; }

extern "C" int zq3d_ens_stream(zq3d_ens* e, void** stream) {
    if (!e || !stream) { set_err("zq3d_ens_stream", "bad argument"); return ZQ3D_ERR_ARG; }
    *stream = e->stream;
    return ZQ3D_OK;
}

extern "C" int zq3d_ens_set_sweep(zq3d_ens* e, int kind, int tbx, int tby, int blocks) {
    const char* where = "zq3d_ens_set_sweep";
    if (!e) { set_err(where, "null ensemble"); return ZQ3D_ERR_ARG; }
    EnsLaunch L = e->L;
    bool ok = false;
    if (kind == 0) {
        for (int sh = 0; sh < kEnsSweepShapeCount; sh++)
            if (kEnsSweepShapes[sh][0] == tbx && kEnsSweepShapes[sh][1] == tby && blocks >= 0) {
                L.sweep_rb = 0;
                L.sweep_shape = sh;
                L.sweep_blocks = blocks;
                ok = true;
            }
    } else if (kind == 1 && rb_fits(e)) {
        for (int r = 0; r < kEnsRbShapeCount; r++)
            if (kEnsRbShapes[r][0] == tbx && kEnsRbShapes[r][1] == tby && kEnsRbShapes[r][2] == blocks) {
                L.sweep_rb = r + 1;
                ok = true;
            }
    }
    if (!ok) { set_err(where, "not a sweep candidate (see zq3d_ens_autotune)"); return ZQ3D_ERR_ARG; }
    e->L = L;
    e->graph_stale = true;
    return ZQ3D_OK;
}

// Times the sweep kernel for every tile shape and persistent-launch size, and the cell kernels for
// every launch size, on the ensemble's own grids; keeps the fastest of each.  The fields are left
// as they were (only projection and advection scratch is written).
extern "C" int zq3d_ens_autotune(zq3d_ens* e, int iters, zq3d_ens_tune* rep) {
    const char* where = "zq3d_ens_autotune";
    if (!e || iters < 1) { set_err(where, "bad argument"); return ZQ3D_ERR_ARG; }
    static const int blocks[] = {0, 256, 1024, 4096};
    const int nb = (int)(sizeof(blocks) / sizeof(blocks[0]));
    hipEvent_t a = nullptr, b = nullptr;
    ZQ_HIP(hipEventCreate(&a), where);
    ZQ_HIP(hipEventCreate(&b), where);
    if (rep) std::memset(rep, 0, sizeof(*rep));
    EnsLaunch L = e->L;
    Errs E;
    auto time_it = [&](auto&& launch) {
        launch();  // warm
        E(hipEventRecord(a, e->stream));
        launch();
        E(hipEventRecord(b, e->stream));
        E(hipEventSynchronize(b));
        float ms = 0.0f;
        E(hipEventElapsedTime(&ms, a, b));
        return ms;
    };
    // the sweeps run on the live counts and right-hand sides (interleaved and colour-split)
    ens_launch_setup(e->L, e->stream);
    ens_launch_split_in(e->L, e->stream);
    int best_sh = 0, best_sb = 0, best_rb = 0;
    float best = -1.0f;
    int n = 0;
    auto consider = [&](int tbx, int tby, int blk, float us, int sh, int sb, int rb) {
        if (rep && n < ZQ3D_ENS_TUNE_MAX) {
            rep->kind[n] = rb ? 1 : 0;
            rep->tbx[n] = tbx;
            rep->tby[n] = tby;
            rep->blocks[n] = blk;
            rep->us_per_iter[n] = us;
            n++;
        }
        if (best < 0.0f || us < best) {
            best = us;
            best_sh = sh;
            best_sb = sb;
            best_rb = rb;
        }
    };
    L.sweep_rb = 0;
    for (int sh = 0; sh < kEnsSweepShapeCount; sh++)
        for (int bi = 0; bi < nb; bi++) {
            L.sweep_shape = sh;
            L.sweep_blocks = blocks[bi];
            const float ms = time_it([&] { ens_launch_sweeps(L, 0, iters, e->stream); });
            consider(kEnsSweepShapes[sh][0], kEnsSweepShapes[sh][1], blocks[bi], 1e3f * ms / iters, sh, blocks[bi], 0);
        }
    for (int r = 0; r < kEnsRbShapeCount && rb_fits(e); r++) {
        L.sweep_rb = r + 1;
        const float ms = time_it([&] { ens_launch_sweeps(L, 0, iters, e->stream); });
        consider(kEnsRbShapes[r][0], kEnsRbShapes[r][1], kEnsRbShapes[r][2], 1e3f * ms / iters, 0, 0, r + 1);
    }
    L.sweep_shape = best_sh;
    L.sweep_blocks = best_sb;
    L.sweep_rb = best_rb;
    int best_b = 0;
    float best_cell = -1.0f;
    for (int bi = 0; bi < nb; bi++) {
        L.blocks = blocks[bi];
        const float ms = time_it([&] { ens_launch_probe(L, e->stream); });
        if (best_cell < 0.0f || ms < best_cell) {
            best_cell = ms;
            best_b = blocks[bi];
        }
    }
    L.blocks = best_b;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (E.e != hipSuccess) { set_err(where, hipGetErrorString(E.e)); return ZQ3D_ERR_HIP; }
    e->L = L;
    e->graph_stale = true;
    if (rep) {
        rep->count = n;
        rep->best_kind = best_rb ? 1 : 0;
        rep->best_tbx = best_rb ? kEnsRbShapes[best_rb - 1][0] : kEnsSweepShapes[best_sh][0];
        rep->best_tby = best_rb ? kEnsRbShapes[best_rb - 1][1] : kEnsSweepShapes[best_sh][1];
        rep->best_sweep_blocks = best_rb ? kEnsRbShapes[best_rb - 1][2] : best_sb;
        rep->best_us_per_iter = best;
        rep->best_blocks = best_b;
        rep->probe_ms = best_cell;
    }
    return ZQ3D_OK;
}
