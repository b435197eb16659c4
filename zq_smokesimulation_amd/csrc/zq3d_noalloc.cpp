// zq3d_noalloc.cpp -- ZQ3D_SolvePoissonRedBlack_noalloc: the projection of ZQ3D_SolvePoissonRedBlack_ex
// on caller device pointers, with every scratch buffer (neighbour counts, right-hand side, the two
// sweep pressures, reduction partials, scalars, and the FP64 pressure and right-hand side of the
// mixed mode) carved out of one file-scope device array, so the call allocates nothing.
//
// Each GPU holds its own copy of `scratch` from the moment the library's code object is loaded on
// it until the process ends; its address on a device comes from hipGetSymbolAddress, looked up
// once per device (with that device current) at the first call there.  Host code never passes the
// array's name to a kernel: in host code the name is a stand-in, not a device address.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "zq3d_internal.h"
#include "zq3d_kernels.h"
#include "zq_smoke3d.h"

using namespace zq3d;
using namespace zq3d::detail;

namespace {

static __device__ __align__(256) unsigned char scratch[ZQ3D_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
unsigned char* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// The current device's address of `scratch` (looked up once per device).
int scratch_base(unsigned char** out, const char* where) {
    int dev = -1;
    ZQ_HIP(hipGetDevice(&dev), where);
    if (dev < 0 || dev >= kMaxDevices) {
        set_err(where, "device index out of range");
        return ZQ3D_ERR_STATE;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        ZQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), where);
        if (!p) {
            set_err(where, "hipGetSymbolAddress returned no address for the scratch array");
            return ZQ3D_ERR_HIP;
        }
        g_scratch[dev] = (unsigned char*)p;
    }
    *out = g_scratch[dev];
    return ZQ3D_OK;
}

// Byte offsets of the projection buffers inside the scratch array (each 256-byte aligned).
struct Layout {
    size_t nc, b, p, p_alt, part, sc, b64, p64, total;
};

Layout layout(const Dims& d, bool mixed) {
    const Sizes z(d);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    Layout L{};
    L.nc = take(z.N);
    L.b = take(z.N * sizeof(float));
    L.p = take(z.Np * sizeof(float));
    L.p_alt = take(z.Np * sizeof(float));
    L.part = take((size_t)reduction_partials(d) * sizeof(double));
    L.sc = take(sizeof(Scalars));
    L.b64 = mixed ? take(z.N * sizeof(double)) : 0;
    L.p64 = mixed ? take(z.Np * sizeof(double)) : 0;
    L.total = off;
    return L;
}

}  // namespace

extern "C" int ZQ3D_SolvePoissonRedBlack_noalloc(float* mac_u, float* mac_v, float* mac_w, const bool* occupy,
                                                 const int width, const int height, const int depth,
                                                 const zq3d_sor_opts* opts, zq3d_proj_stats* stats, void* stream) {
    const char* where = "ZQ3D_SolvePoissonRedBlack_noalloc";
    if (!opts) { set_err(where, "null options"); return ZQ3D_ERR_ARG; }
    if (int rc = check_dims(width, height, depth, where)) return rc;
    if (int rc = check_sor_opts(*opts, where)) return rc;
    if (!mac_u || !mac_v || !mac_w || !occupy) { set_err(where, "null pointer"); return ZQ3D_ERR_ARG; }
    const Dims d{width, height, depth};
    const bool mixed = opts->mixed_precision != 0;
    const Layout L = layout(d, mixed);
    if (L.total > (size_t)ZQ3D_NOALLOC_SCRATCH_BYTES) {
        char msg[256];
        std::snprintf(msg, sizeof(msg),
                      "a %dx%dx%d %s projection needs %zu bytes of scratch; this entry point holds at most %zu "
                      "(use ZQ3D_SolvePoissonRedBlack_ex or a zq3d_sim for larger grids)",
                      width, height, depth, mixed ? "mixed-precision" : "FP32", L.total,
                      (size_t)ZQ3D_NOALLOC_SCRATCH_BYTES);
        set_err(where, msg);
        return ZQ3D_ERR_ARG;
    }
    unsigned char* base = nullptr;
    if (int rc = scratch_base(&base, where)) return rc;
    const hipStream_t s = (hipStream_t)stream;
    const Sizes z(d);
    const ProjBufs B{(uint8_t*)(base + L.nc),
                     (float*)(base + L.b),
                     (float*)(base + L.p),
                     (float*)(base + L.p_alt),
                     mixed ? (double*)(base + L.b64) : nullptr,
                     mixed ? (double*)(base + L.p64) : nullptr,
                     (double*)(base + L.part),
                     (Scalars*)(base + L.sc),
                     SorShape{0, 0},
                     SplitBufs{}};
    // The array keeps the previous call's contents: the second sweep buffer's ghost layer and the
    // FP64 pressure (cold start) must be zero; every other buffer is written before it is read.
    launch_fill_f32(B.p_alt, z.Np, 0.0f, s);
    int rc;
    if (mixed) {
        launch_fill_f64(B.p64, z.Np, 0.0, s);
        rc = project_mixed(d, mac_u, mac_v, mac_w, (const uint8_t*)occupy, *opts, B, stats, s);
    } else {
        rc = project_plain(d, mac_u, mac_v, mac_w, (const uint8_t*)occupy, *opts, 0.0f, opts->closed_mean ? 1 : 0, B,
                           stats, s);
    }
    if (rc) return rc;
    ZQ_HIP(hipGetLastError(), where);
    return ZQ3D_OK;
}
