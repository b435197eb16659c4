// zq3d_octree.cpp -- the octree projection on the device-resident simulation, and the two
// ZQlibCuda octree solver calls of ZQ_SmokeOctreeGrid3D (host arrays).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "zq3d_internal.h"

namespace zq3d {
namespace detail {

int octree_create(const Dims& d, const int thresh[3], OctreeState* out, const char* where) {
    out->d = d;
    for (int i = 0; i < 3; i++) out->thresh[i] = thresh[i];
    ZQ_HIP(hipMalloc(&out->mem, octree_bytes(d)), where);
    octree_layout(d, (char*)out->mem, &out->B);
    return ZQ3D_OK;
}

void octree_destroy(OctreeState* o) {
    if (!o) return;
    if (o->mem) (void)hipFree(o->mem);
    o->mem = nullptr;
    octree_free_lists(&o->lists);
}

namespace {

// The special-region projections of octree guiding (see project_octree's declaration).
int octree_special_regions(OctreeState& t, float* u, float* v, float* w, int iters, hipStream_t s,
                           const char* where) {
    const Dims& d = t.d;
    const Sizes z(d);
    const int W = d.W, H = d.H, D = d.D;
    std::vector<uint8_t> lev(z.N);
    std::vector<float> hu(z.Nu), hv(z.Nv), hw(z.Nw);
    ZQ_HIP(hipMemcpyAsync(lev.data(), t.B.lev, z.N, hipMemcpyDeviceToHost, s), where);
    ZQ_HIP(hipMemcpyAsync(hu.data(), u, z.Nu * 4, hipMemcpyDeviceToHost, s), where);
    ZQ_HIP(hipMemcpyAsync(hv.data(), v, z.Nv * 4, hipMemcpyDeviceToHost, s), where);
    ZQ_HIP(hipMemcpyAsync(hw.data(), w, z.Nw * 4, hipMemcpyDeviceToHost, s), where);
    ZQ_HIP(hipStreamSynchronize(s), where);
    // regions: the labelling of _find_special_label (ZQ_SmokeOctreeGrid3D.cpp:5169-5320) over the
    // cells in level-0 and level-1 leaves.  A pass seeds the first unlabelled cell of the lowest plane
    // that has one and, in each plane above, the first unlabelled cell of row 0 only, all under the
    // pass's label, then fills from the last seed: z neighbours are queued unlabelled, y and x
    // neighbours labelled as they are queued (order z+1, z-1, y+1, y-1, x+1, x-1).  A cell reached
    // only along z seeds a later pass; two such cells above each other queue each other without end,
    // which overflows the reference's W*H*D queue: an error here.
    const size_t plane = (size_t)W * H;
    std::vector<int32_t> lab(z.N, 0);
    for (size_t o = 0; o < z.N; o++)
        if (lev[o] == 0 || lev[o] == 1) lab[o] = -1;
    int regions = 0;
    std::vector<size_t> queue;
    for (;;) {
        size_t seed = z.N;
        bool seeded = false;
        for (int k = 0; k < D; k++) {
            const int rows = seeded ? 1 : H;
            bool here = false;
            for (int j = 0; j < rows && !here; j++)
                for (int i = 0; i < W; i++) {
                    const size_t c = (size_t)k * plane + (size_t)j * W + i;
                    if (lab[c] == -1) {
                        lab[c] = regions + 1;
                        seed = c;
                        here = true;
                        break;
                    }
                }
            seeded = seeded || here;
        }
        if (!seeded) break;
        const int r = ++regions;
        queue.assign(1, seed);
        for (size_t h = 0; h < queue.size(); h++) {
            const size_t c = queue[h];
            const int x = (int)(c % W), y = (int)((c / W) % H), k = (int)(c / plane);
            auto visit = [&](size_t n, bool label) {
                if (lab[n] != -1) return;
                if (label) lab[n] = r;
                queue.push_back(n);
            };
            if (k + 1 < D) visit(c + plane, false);
            if (k >= 1) visit(c - plane, false);
            if (y + 1 < H) visit(c + W, true);
            if (y >= 1) visit(c - W, true);
            if (x + 1 < W) visit(c + 1, true);
            if (x >= 1) visit(c - 1, true);
            if (queue.size() > z.N) {
                set_err(where, "octree guiding: the reference's special-region labelling does not end on these "
                               "leaves (two cells above each other reached only along z)");
                return ZQ3D_ERR_STATE;
            }
        }
    }
    const size_t Wu = (size_t)W + 1, Hv = (size_t)H + 1;
    // every region's cell count, float divergence sum and bounding box in one scan-order pass, where
    // the reference scans the grid once per region (ZQ_SmokeOctreeGrid3D.cpp:561-590): each sum still
    // takes its region's cells in scan order, and a region's solve changes only faces between two of
    // its own cells, so the faces a later region sums are the ones before any solve
    struct Box {
        int x0, y0, z0, x1, y1, z1, count;
        float sum;
    };
    std::vector<Box> boxes(regions + 1, Box{W, H, D, -1, -1, -1, 0, 0.0f});
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                const int r = lab[((size_t)k * H + j) * W + i];
                if (r <= 0) continue;
                Box& b = boxes[r];
                b.count++;
                b.sum += hu[((size_t)k * H + j) * Wu + i + 1] - hu[((size_t)k * H + j) * Wu + i] +
                         hv[((size_t)k * Hv + j + 1) * W + i] - hv[((size_t)k * Hv + j) * W + i] +
                         hw[((size_t)(k + 1) * H + j) * W + i] - hw[((size_t)k * H + j) * W + i];
                b.x0 = std::min(b.x0, i), b.y0 = std::min(b.y0, j), b.z0 = std::min(b.z0, k);
                b.x1 = std::max(b.x1, i), b.y1 = std::max(b.y1, j), b.z1 = std::max(b.z1, k);
            }
    for (int r = 1; r <= regions; r++) {
        const Box& b = boxes[r];
        const int x0 = b.x0, y0 = b.y0, z0 = b.z0, count = b.count;
        const float mean = b.sum / count;
        const int sx = b.x1 - x0 + 1, sy = b.y1 - y0 + 1, sz = b.z1 - z0 + 1;
        std::vector<uint8_t> bo((size_t)sx * sy * sz);
        std::vector<float> bu((size_t)(sx + 1) * sy * sz), bv((size_t)sx * (sy + 1) * sz), bw((size_t)sx * sy * (sz + 1));
        for (int k = 0; k < sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++)
                    bo[((size_t)k * sy + j) * sx + i] = lab[((size_t)(z0 + k) * H + y0 + j) * W + x0 + i] != r;
        for (int k = 0; k < sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i <= sx; i++)
                    bu[((size_t)k * sy + j) * (sx + 1) + i] = hu[((size_t)(z0 + k) * H + y0 + j) * Wu + x0 + i];
        for (int k = 0; k < sz; k++)
            for (int j = 0; j <= sy; j++)
                for (int i = 0; i < sx; i++)
                    bv[((size_t)k * (sy + 1) + j) * sx + i] = hv[((size_t)(z0 + k) * Hv + y0 + j) * W + x0 + i];
        for (int k = 0; k <= sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++)
                    bw[((size_t)k * sy + j) * sx + i] = hw[((size_t)(z0 + k) * H + y0 + j) * W + x0 + i];
        if (ZQ_CUDA_PoissonSolver3D::SolveClosedPoissonRedBlackwithOccupy3D_MAC(
                bu.data(), bv.data(), bw.data(), reinterpret_cast<const bool*>(bo.data()), mean, sx, sy, sz,
                std::min(iters, count)) < 0.0f)
            return ZQ3D_ERR_HIP;
        // written back as the reference does: the box's u faces except its high x side (:700-712)
        for (int k = 0; k < sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++)
                    hu[((size_t)(z0 + k) * H + y0 + j) * Wu + x0 + i] = bu[((size_t)k * sy + j) * (sx + 1) + i];
        for (int k = 0; k < sz; k++)
            for (int j = 0; j <= sy; j++)
                for (int i = 0; i < sx; i++)
                    hv[((size_t)(z0 + k) * Hv + y0 + j) * W + x0 + i] = bv[((size_t)k * (sy + 1) + j) * sx + i];
        for (int k = 0; k <= sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++)
                    hw[((size_t)(z0 + k) * H + y0 + j) * W + x0 + i] = bw[((size_t)k * sy + j) * sx + i];
    }
    ZQ_HIP(hipMemcpyAsync(u, hu.data(), z.Nu * 4, hipMemcpyHostToDevice, s), where);
    ZQ_HIP(hipMemcpyAsync(v, hv.data(), z.Nv * 4, hipMemcpyHostToDevice, s), where);
    ZQ_HIP(hipMemcpyAsync(w, hw.data(), z.Nw * 4, hipMemcpyHostToDevice, s), where);
    ZQ_HIP(hipStreamSynchronize(s), where);  // the host vectors go out of scope
    return ZQ3D_OK;
}

}  // namespace

int project_octree(OctreeState& t, float* u, float* v, float* w, const uint8_t* occ, const zq3d_sor_opts& o,
                   const ProjBufs& B, zq3d_proj_stats* st, hipStream_t s, SorProf* pf) {
    const char* where = "project_octree";
    // the leaves follow the occupancy, which callers may change between projections
    // (ReBuildFromGlobalGrid each frame with moving objects, ZQ_SmokeSimulation3D.cpp:482-491)
    launch_octree_build(t.d, occ, t.thresh[0], t.thresh[1], t.thresh[2], t.B, s);
    const float* mean = nullptr;
    if (o.closed_mean) {
        // the mean divergence over the fluid cells (_solveGlobalClosedPoissonSOR, ZQ_SmokeOctreeGrid3D.cpp:2935-2964)
        launch_proj_setup(t.d, u, v, w, occ, 0, B.nc, B.b, nullptr, B.p, nullptr, B.part, s);
        launch_mean_finalize(t.d, B.part, &B.sc->mean32, &B.sc->mean64, 1, s);
        mean = &B.sc->mean32;
    }
    ZQ_HIP(octree_build_lists(t.d, t.B, &t.lists, s), where);
    launch_octree_prepare(t.d, u, v, w, t.B, s);
    hipEvent_t ea = nullptr, eb = nullptr;
    if (pf) {
        ZQ_HIP(pf->next(&ea, &eb), where);
        ZQ_HIP(hipEventRecord(ea, s), where);
    }
    const int launches = launch_octree_sweeps(t.d, t.B, t.lists, o.iters, mean, s);
    if (pf) {
        ZQ_HIP(hipEventRecord(eb, s), where);
        pf->launches += launches;
    }
    launch_octree_adjust(t.d, t.B, u, v, w, s);
    if (t.guiding)
        if (int rc = octree_special_regions(t, u, v, w, o.iters, s, where)) return rc;
    launch_octree_sub_project(t.d, t.B, u, v, w, s);
    ZQ_HIP(hipGetLastError(), where);
    if (st) {
        Scalars* sc = B.sc;
        launch_fill_f64(&sc->mean64, 1, 0.0, s);
        launch_proj_setup(t.d, u, v, w, occ, ZQ3D_OPEN_ALL, B.nc, B.b, nullptr, B.p, nullptr, nullptr, s);
        launch_max_div64(t.d, u, v, w, B.nc, &sc->mean64, B.part, B.target(1, &sc->dmax), s);
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

size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

// grow-only device arena of the two host-pointer entry points below
struct OctArena {
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

OctArena& oct_arena() {
    static OctArena a;
    return a;
}

OctreeLists& oct_arena_lists() {  // the entry buffer of the host-pointer calls (guarded by the arena's mutex)
    static OctreeLists l;
    return l;
}

// One octree solve on host arrays: faces and leaf flags in, corrected faces out.  The index and
// neighbour lists the reference passes are a function of the leaf flags alone (its
// _selectInfos_for_CUDAOpenPoissonRedBlack3 builds them from nothing else); the kernels derive the
// same neighbours, in the same order, from the flags, so the lists are accepted and not read.
float octree_call(const char* where, float* mac_u, float* mac_v, float* mac_w, const bool* leaf0, const bool* leaf1,
                  const bool* leaf2, const bool* leaf3, const float* div_per_volume, int width, int height, int depth,
                  int maxIter) {
    if (check_dims(width, height, depth, where)) return -1.0f;
    if (width % 8 || height % 8 || depth % 8) {
        set_err(where, "width, height and depth must be multiples of 8 (the octree's 8^3 blocks)");
        return -1.0f;
    }
    if ((long long)width * height * depth >= (1LL << 30)) {
        set_err(where, "at most 2^30 cells (neighbour entries hold 30-bit indices)");
        return -1.0f;
    }
    if (!mac_u || !mac_v || !mac_w || !leaf0 || !leaf1 || !leaf2 || !leaf3) {
        set_err(where, "null pointer");
        return -1.0f;
    }
    if (maxIter < 0) { set_err(where, "maxIter must be >= 0"); return -1.0f; }
    static_assert(sizeof(bool) == 1, "leaf flags are one byte per cell");
    const Dims d{width, height, depth};
    const Sizes z(d);
    const size_t n1 = z.N / 8, n2 = z.N / 64, n3 = z.N / 512;
    const size_t o_u = 0, o_v = o_u + up256(z.Nu * 4), o_w = o_v + up256(z.Nv * 4), o_l0 = o_w + up256(z.Nw * 4);
    const size_t o_l1 = o_l0 + up256(z.N), o_l2 = o_l1 + up256(n1), o_l3 = o_l2 + up256(n2), o_m = o_l3 + up256(n3);
    const size_t o_t = o_m + 256, total = o_t + octree_bytes(d);
    OctArena& A = oct_arena();
    std::lock_guard<std::mutex> lk(A.mu);
    if (A.ensure(total) != hipSuccess) { set_err(where, "device allocation failed"); return -1.0f; }
    char* base = (char*)A.ptr;
    float *du = (float*)(base + o_u), *dv = (float*)(base + o_v), *dw = (float*)(base + o_w);
    uint8_t *l0 = (uint8_t*)(base + o_l0), *l1 = (uint8_t*)(base + o_l1), *l2 = (uint8_t*)(base + o_l2);
    uint8_t* l3 = (uint8_t*)(base + o_l3);
    float* mean = (float*)(base + o_m);
    OctreeBufs B{};
    octree_layout(d, base + o_t, &B);
    hipEvent_t a = nullptr, b = nullptr;
    Errs E;
    E(hipEventCreate(&a));
    E(hipEventCreate(&b));
    E(hipEventRecord(a, 0));
    E(hipMemcpy(du, mac_u, z.Nu * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(dv, mac_v, z.Nv * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(dw, mac_w, z.Nw * 4, hipMemcpyHostToDevice));
    E(hipMemcpy(l0, leaf0, z.N, hipMemcpyHostToDevice));
    E(hipMemcpy(l1, leaf1, n1, hipMemcpyHostToDevice));
    E(hipMemcpy(l2, leaf2, n2, hipMemcpyHostToDevice));
    E(hipMemcpy(l3, leaf3, n3, hipMemcpyHostToDevice));
    if (div_per_volume) E(hipMemcpy(mean, div_per_volume, 4, hipMemcpyHostToDevice));
    if (E.e == hipSuccess) {
        launch_octree_from_leaves(d, l0, l1, l2, l3, B, 0);
        OctreeLists& Ls = oct_arena_lists();
        E(octree_build_lists(d, B, &Ls, 0));
        launch_octree_prepare(d, du, dv, dw, B, 0);
        if (E.e == hipSuccess) launch_octree_sweeps(d, B, Ls, maxIter, div_per_volume ? mean : nullptr, 0);
        launch_octree_adjust(d, B, du, dv, dw, 0);
        E(hipGetLastError());
    }
    E(hipMemcpy(mac_u, du, z.Nu * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_v, dv, z.Nv * 4, hipMemcpyDeviceToHost));
    E(hipMemcpy(mac_w, dw, z.Nw * 4, hipMemcpyDeviceToHost));
    E(hipEventRecord(b, 0));
    E(hipEventSynchronize(b));
    float ms = 0.0f;
    E(hipEventElapsedTime(&ms, a, b));
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (E.e != hipSuccess) { set_err(where, hipGetErrorString(E.e)); return -1.0f; }
    return ms;
}

}  // namespace

#define ZQ3D_OCT_LISTS                                                                                          \
    const int, const int*, const int, const int*, const int, const int*, const int, const int*, const int,     \
        const int*, const int, const int*, const int, const int*, const int, const int*, const int, const int*, \
        const int, const int*, const int, const int*, const int, const int*, const int, const int*, const int,  \
        const int*, const int, const int*, const int, const int*

extern "C" float SolveOpenOctreePoissonRedBlack3_3D_MAC(float* mac_u, float* mac_v, float* mac_w, const bool* leaf0,
                                                       const bool* leaf1, const bool* leaf2, const bool* leaf3,
                                                       const int width, const int height, const int depth,
                                                       const int maxIter, ZQ3D_OCT_LISTS) {
    return octree_call("SolveOpenOctreePoissonRedBlack3_3D_MAC", mac_u, mac_v, mac_w, leaf0, leaf1, leaf2, leaf3,
                       nullptr, width, height, depth, maxIter);
}

extern "C" float SolveClosedOctreePoissonRedBlack3_3D_MAC(float* mac_u, float* mac_v, float* mac_w, const bool* leaf0,
                                                         const bool* leaf1, const bool* leaf2, const bool* leaf3,
                                                         const float div_per_volume, const int width,
                                                         const int height, const int depth, const int maxIter,
                                                         ZQ3D_OCT_LISTS) {
    return octree_call("SolveClosedOctreePoissonRedBlack3_3D_MAC", mac_u, mac_v, mac_w, leaf0, leaf1, leaf2, leaf3,
                       &div_per_volume, width, height, depth, maxIter);
}
