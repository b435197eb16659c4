// zq3d_device.h -- device-side helpers shared by the gfx950 kernels.
//
// Layout is the reference's (ZQ_SmokeGrid3D.cpp:5-18), x fastest:
//   cell (i,j,k)   : (k*H + j)*W + i
//   u face (i,j,k) : (k*H + j)*(W+1) + i      i in [0,W]
//   v face (i,j,k) : (k*(H+1) + j)*W + i      j in [0,H]
//   w face (i,j,k) : (k*H + j)*W + i          k in [0,D]
//   padded pressure: ((k+1)*(H+2) + (j+1))*(W+2) + (i+1), ghost layer held at 0.
// Every kernel maps x to the 64 lanes of a wavefront, so each global access of a
// row is one coalesced 256-byte transaction per wave.
//
// Arithmetic is written operation by operation in a fixed order and the library
// is compiled with -ffp-contract=off, so results are reproducible bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>
#include <cmath>
#include <type_traits>

namespace zq3d {

struct Dims {
    int W, H, D;
};

__device__ __forceinline__ int cidx(const Dims& d, int i, int j, int k) { return (k * d.H + j) * d.W + i; }
__device__ __forceinline__ int uidx(const Dims& d, int i, int j, int k) { return (k * d.H + j) * (d.W + 1) + i; }
__device__ __forceinline__ int vidx(const Dims& d, int i, int j, int k) { return (k * (d.H + 1) + j) * d.W + i; }
__device__ __forceinline__ int widx(const Dims& d, int i, int j, int k) { return (k * d.H + j) * d.W + i; }
__device__ __forceinline__ int pidx(const Dims& d, int i, int j, int k) {
    return ((k + 1) * (d.H + 2) + (j + 1)) * (d.W + 2) + (i + 1);
}

// Real-type plumbing: R = float (the product path) or double (the FP64 check mode of the adjoint).
template <class R>
struct V4;
template <>
struct V4<float> {
    using type = float4;
};
template <>
struct V4<double> {
    using type = double4;
};
template <class R>
using vec4_t = typename V4<R>::type;

template <class R>
struct R3 {
    R x, y, z;
};

__device__ __forceinline__ float zfloor(float x) { return floorf(x); }
__device__ __forceinline__ double zfloor(double x) { return floor(x); }
__device__ __forceinline__ float zsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double zsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float zmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double zmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float zmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double zmax(double a, double b) { return fmax(a, b); }

template <class R>
__device__ __forceinline__ R lerp(R a, R b, R t) {
    return a + t * (b - a);
}
__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// One axis of a clamp-to-edge linear filter in texel space: texels i0, i1 and weight a.
// The coordinate is first limited to [-1, n], which leaves the result unchanged and keeps the
// float->int conversion defined for any input (also NaN, which fmin/fmax map to a bound).
template <class R>
__device__ __forceinline__ void axis_setup(R x, int n, int& i0, int& i1, R& a) {
    x = zmin(zmax(x, R(-1)), (R)n);
    const R fx = zfloor(x);
    const int k = (int)fx;
    a = x - fx;
    i1 = clampi(k + 1, 0, n - 1);
    i0 = clampi(k, 0, n - 1);
}

// Trilinear sample of a scalar grid nx*ny*nz at texel coordinates (x,y,z).
template <class R>
__device__ __forceinline__ R sample3(const R* __restrict__ g, int nx, int ny, int nz, R x, R y, R z) {
    int i0, i1, j0, j1, k0, k1;
    R ax, ay, az;
    axis_setup(x, nx, i0, i1, ax);
    axis_setup(y, ny, j0, j1, ay);
    axis_setup(z, nz, k0, k1, az);
    const int r00 = (k0 * ny + j0) * nx, r10 = (k0 * ny + j1) * nx;
    const int r01 = (k1 * ny + j0) * nx, r11 = (k1 * ny + j1) * nx;
    const R c00 = lerp(g[r00 + i0], g[r00 + i1], ax);
    const R c10 = lerp(g[r10 + i0], g[r10 + i1], ax);
    const R c01 = lerp(g[r01 + i0], g[r01 + i1], ax);
    const R c11 = lerp(g[r11 + i0], g[r11 + i1], ax);
    const R c0 = lerp(c00, c10, ay);
    const R c1 = lerp(c01, c11, ay);
    return lerp(c0, c1, az);
}

// Trilinear sample of the first three components of a 4-wide cell grid (one 16-byte load per
// tap in FP32).
template <class R>
__device__ __forceinline__ R3<R> sample4(const vec4_t<R>* __restrict__ g, int nx, int ny, int nz, R x, R y, R z) {
    int i0, i1, j0, j1, k0, k1;
    R ax, ay, az;
    axis_setup(x, nx, i0, i1, ax);
    axis_setup(y, ny, j0, j1, ay);
    axis_setup(z, nz, k0, k1, az);
    const int r00 = (k0 * ny + j0) * nx, r10 = (k0 * ny + j1) * nx;
    const int r01 = (k1 * ny + j0) * nx, r11 = (k1 * ny + j1) * nx;
    const vec4_t<R> t000 = g[r00 + i0], t100 = g[r00 + i1], t010 = g[r10 + i0], t110 = g[r10 + i1];
    const vec4_t<R> t001 = g[r01 + i0], t101 = g[r01 + i1], t011 = g[r11 + i0], t111 = g[r11 + i1];
    R3<R> out;
    {
        const R c00 = lerp(t000.x, t100.x, ax), c10 = lerp(t010.x, t110.x, ax);
        const R c01 = lerp(t001.x, t101.x, ax), c11 = lerp(t011.x, t111.x, ax);
        out.x = lerp(lerp(c00, c10, ay), lerp(c01, c11, ay), az);
    }
    {
        const R c00 = lerp(t000.y, t100.y, ax), c10 = lerp(t010.y, t110.y, ax);
        const R c01 = lerp(t001.y, t101.y, ax), c11 = lerp(t011.y, t111.y, ax);
        out.y = lerp(lerp(c00, c10, ay), lerp(c01, c11, ay), az);
    }
    {
        const R c00 = lerp(t000.z, t100.z, ax), c10 = lerp(t010.z, t110.z, ax);
        const R c01 = lerp(t001.z, t101.z, ax), c11 = lerp(t011.z, t111.z, ax);
        out.z = lerp(lerp(c00, c10, ay), lerp(c01, c11, ay), az);
    }
    return out;
}

// Velocity at a position for the backtrace: cell-centred 4-wide grid sampled at pos-0.5
// (the reference's normalized coordinate pos/N with a linear filter), or the three staggered
// face grids (u at (x, y-0.5, z-0.5) on (W+1)*H*D, ...).
template <class R>
struct RegularVel {
    const vec4_t<R>* vel4;
    Dims d;
    __device__ __forceinline__ R3<R> operator()(R px, R py, R pz) const {
        return sample4<R>(vel4, d.W, d.H, d.D, px - R(0.5), py - R(0.5), pz - R(0.5));
    }
};

// The 8 taps of one trilinear cell of a 4-wide grid, kept between samples: consecutive backtrace
// positions move a small fraction of a cell, so most samples of a path reuse the taps of the
// previous one and load nothing.  The arithmetic is sample4's, so results are identical.
template <class R>
struct Cell4Taps {
    int i0 = -1, i1 = -1, j0 = -1, j1 = -1, k0 = -1, k1 = -1;
    vec4_t<R> t000, t100, t010, t110, t001, t101, t011, t111;
    __device__ __forceinline__ bool same(int a0, int a1, int b0, int b1, int c0, int c1) const {
        return a0 == i0 && a1 == i1 && b0 == j0 && b1 == j1 && c0 == k0 && c1 == k1;
    }
    __device__ __forceinline__ void fetch(const vec4_t<R>* __restrict__ g, int nx, int ny, int a0, int a1, int b0,
                                          int b1, int c0, int c1) {
        if (same(a0, a1, b0, b1, c0, c1)) return;
        i0 = a0, i1 = a1, j0 = b0, j1 = b1, k0 = c0, k1 = c1;
        const int r00 = (c0 * ny + b0) * nx, r10 = (c0 * ny + b1) * nx;
        const int r01 = (c1 * ny + b0) * nx, r11 = (c1 * ny + b1) * nx;
        t000 = g[r00 + a0], t100 = g[r00 + a1], t010 = g[r10 + a0], t110 = g[r10 + a1];
        t001 = g[r01 + a0], t101 = g[r01 + a1], t011 = g[r11 + a0], t111 = g[r11 + a1];
    }
};

// One axis of RegularVelCached's filter: the coordinate limited as in axis_setup, its floor k and
// the weight a (the texels k and k+1 are clamped into [0, n) only when the cell's taps are loaded).
template <class R>
__device__ __forceinline__ int axis_floor(R x, int n, R& a) {
    x = zmin(zmax(x, R(-1)), (R)n);
    const R fx = zfloor(x);
    a = x - fx;
    return (int)fx;
}

// The 8 taps of one trilinear cell of a 4-wide grid in the form the x-lerps use: for each x pair
// the low tap lo and the difference hi - lo, so that lerp(lo, hi, ax) = lo + ax*(hi - lo) takes its
// subtraction once per cell instead of once per sample (the same operations on the same values:
// identical results).  The cell is keyed by the floors (kx, ky, kz) of the filter coordinates.
// The fourth component of a tap holds its cell's occupancy (nonzero = solid) in the grids the forward
// backtrace samples (st_mac_to_cell4 with occupancy): bit t of `solid` is tap t's, t = x + 2y + 4z
// over the taps 000, 100, 010, 110, 001, 101, 011, 111.
template <class R>
struct Cell4Lerp {
    int kx = INT_MIN, ky = INT_MIN, kz = INT_MIN;
    unsigned solid = 0;
    R3<R> a00, d00, a10, d10, a01, d01, a11, d11;  // pairs (i0,j0,k0)-(i1,..), (i0,j1,k0), (i0,j0,k1), (i0,j1,k1)
    __device__ __forceinline__ void fetch(const vec4_t<R>* __restrict__ g, const Dims& d, int x, int y, int z) {
        if (x == kx && y == ky && z == kz) return;
        kx = x, ky = y, kz = z;
        const int i0 = clampi(x, 0, d.W - 1), i1 = clampi(x + 1, 0, d.W - 1);
        const int j0 = clampi(y, 0, d.H - 1), j1 = clampi(y + 1, 0, d.H - 1);
        const int k0 = clampi(z, 0, d.D - 1), k1 = clampi(z + 1, 0, d.D - 1);
        const int r00 = (k0 * d.H + j0) * d.W, r10 = (k0 * d.H + j1) * d.W;
        const int r01 = (k1 * d.H + j0) * d.W, r11 = (k1 * d.H + j1) * d.W;
        const vec4_t<R> t000 = g[r00 + i0], t100 = g[r00 + i1], t010 = g[r10 + i0], t110 = g[r10 + i1];
        const vec4_t<R> t001 = g[r01 + i0], t101 = g[r01 + i1], t011 = g[r11 + i0], t111 = g[r11 + i1];
        auto pair = [](const vec4_t<R>& lo, const vec4_t<R>& hi, R3<R>& a, R3<R>& dd) {
            a = R3<R>{lo.x, lo.y, lo.z};
            dd = R3<R>{hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
        };
        pair(t000, t100, a00, d00);
        pair(t010, t110, a10, d10);
        pair(t001, t101, a01, d01);
        pair(t011, t111, a11, d11);
        solid = (t000.w != R(0)) | (t100.w != R(0)) << 1 | (t010.w != R(0)) << 2 | (t110.w != R(0)) << 3 |
                (t001.w != R(0)) << 4 | (t101.w != R(0)) << 5 | (t011.w != R(0)) << 6 | (t111.w != R(0)) << 7;
    }
};

// RegularVel with a per-thread tap cache (one instance per thread; not shared): consecutive
// backtrace positions move a small fraction of a cell, so most samples reuse the taps of the
// previous one and load nothing.  The arithmetic is sample4's, so results are identical.
template <class R>
struct RegularVelCached {
    const vec4_t<R>* vel4;
    Dims d;
    mutable Cell4Lerp<R> c;
    mutable R ax = R(0), ay = R(0), az = R(0);  // weights of the last prepared position
    // taps and weights of the trilinear cell around (px, py, pz)
    __device__ __forceinline__ void prepare(R px, R py, R pz) const {
        const int x = axis_floor(px - R(0.5), d.W, ax);
        const int y = axis_floor(py - R(0.5), d.H, ay);
        const int z = axis_floor(pz - R(0.5), d.D, az);
        c.fetch(vel4, d, x, y, z);
    }
    // Whether the cell containing the prepared position is solid.  For a position inside
    // [0,W]x[0,H]x[0,D], clamp(floor(p), 0, n-1) along an axis is the high tap exactly when the
    // weight is at least 1/2 (p - 1/2 and its fraction are exact in FP32 below 2^23), so the cell
    // the reference point-samples is one of the 8 taps.
    __device__ __forceinline__ bool solid_prepared() const {
        const int t = (ax >= R(0.5) ? 1 : 0) | (ay >= R(0.5) ? 2 : 0) | (az >= R(0.5) ? 4 : 0);
        return (c.solid >> t) & 1u;
    }
    __device__ __forceinline__ R3<R> operator()(R px, R py, R pz) const {
        prepare(px, py, pz);
        return value();
    }
    // the velocity at the prepared position
    __device__ __forceinline__ R3<R> value() const {
        R3<R> out;
        {
            const R c00 = c.a00.x + ax * c.d00.x, c10 = c.a10.x + ax * c.d10.x;
            const R c01 = c.a01.x + ax * c.d01.x, c11 = c.a11.x + ax * c.d11.x;
            out.x = lerp(lerp(c00, c10, ay), lerp(c01, c11, ay), az);
        }
        {
            const R c00 = c.a00.y + ax * c.d00.y, c10 = c.a10.y + ax * c.d10.y;
            const R c01 = c.a01.y + ax * c.d01.y, c11 = c.a11.y + ax * c.d11.y;
            out.y = lerp(lerp(c00, c10, ay), lerp(c01, c11, ay), az);
        }
        {
            const R c00 = c.a00.z + ax * c.d00.z, c10 = c.a10.z + ax * c.d10.z;
            const R c01 = c.a01.z + ax * c.d01.z, c11 = c.a11.z + ax * c.d11.z;
            out.z = lerp(lerp(c00, c10, ay), lerp(c01, c11, ay), az);
        }
        return out;
    }
};

template <class R>
struct MacVel {
    const R* u;
    const R* v;
    const R* w;
    Dims d;
    __device__ __forceinline__ R3<R> operator()(R px, R py, R pz) const {
        R3<R> r;
        r.x = sample3<R>(u, d.W + 1, d.H, d.D, px, py - R(0.5), pz - R(0.5));
        r.y = sample3<R>(v, d.W, d.H + 1, d.D, px - R(0.5), py, pz - R(0.5));
        r.z = sample3<R>(w, d.W, d.H, d.D + 1, px - R(0.5), py - R(0.5), pz);
        return r;
    }
};


// A velocity sampler negated (the BFECC back-advection, ZQ_CUDA_Advection3D.cu:997-1019, negates the
// field before sampling it; negating the interpolated value gives the same bits).
template <class R, class Vel>
struct NegVel {
    Vel v;
    __device__ __forceinline__ R3<R> operator()(R px, R py, R pz) const {
        const R3<R> r = v(px, py, pz);
        return R3<R>{-r.x, -r.y, -r.z};
    }
};

// True while the backtrace may take another step from (px,py,pz): inside [0,W]x[0,H]x[0,D] and
// not in a solid cell (point sample, clamp).  The occupancy byte of the cell last tested is kept:
// consecutive backtrace points usually share a cell, so most tests load nothing.
struct OccCache {
    int o = -1;
    bool open = false;
};

template <class R>
__device__ __forceinline__ bool trace_open_cached(const uint8_t* __restrict__ occ, const Dims& d, R px, R py, R pz,
                                                  OccCache& c) {
    if (!(px >= R(0) && px <= (R)d.W && py >= R(0) && py <= (R)d.H && pz >= R(0) && pz <= (R)d.D)) return false;
    const int oi = clampi((int)zfloor(px), 0, d.W - 1);
    const int oj = clampi((int)zfloor(py), 0, d.H - 1);
    const int ok = clampi((int)zfloor(pz), 0, d.D - 1);
    const int o = cidx(d, oi, oj, ok);
    if (o != c.o) {
        c.o = o;
        c.open = !occ[o];
    }
    return c.open;
}

// Scaling of a backtrace displacement v*dtt into cells: (v*dtt)/h, the reference's division.  When h
// is a power of two (voxelSize = 1/W for W = 64, 128, 256, ...) the division is exact scaling by
// 2^k, so multiplying by inv = 1/h gives the same bits; inv = 0 selects the division.
template <class R>
struct StepScale {
    R h, inv;
    __device__ __forceinline__ R operator()(R x) const { return inv != R(0) ? x * inv : x / h; }
};

// The scale for voxel size h (host side): inv = 1/h when h is an exact power of two whose inverse
// is a normal number of the type (then x / h == x * inv for every x), else 0.
template <class R>
inline StepScale<R> step_scale(float h) {
    int e = 0;
    const double m = std::frexp((double)h, &e);  // h = m * 2^e, m in [0.5, 1)
    R inv = R(0);
    if (m == 0.5 && e > -100 && e < 100) inv = (R)std::ldexp(1.0, 1 - e);
    return StepScale<R>{(R)h, inv};
}

// Semi-Lagrangian backtrace with the reference's loop (ZQ_CUDA_Advection3D.cu:87-109): stop when
// the point leaves [0,W]x[0,H]x[0,D] or lands in a solid cell; otherwise remember it and step
// back by vel*dtt/h.  The end point is the last accepted position.  The reference samples the
// velocity at every new point and takes `steps` steps, but the point reached by the last step is
// never accepted (the loop ends first), so neither that step nor the velocity it uses is computed
// here: the accepted points, hence the result, are the same.
template <class R, class Vel>
__device__ __forceinline__ R3<R> backtrace(const Vel& V, const uint8_t* __restrict__ occ, const Dims& d,
                                           const StepScale<R>& sc, int steps, R dtt, R px, R py, R pz);

// The same loop for the cell-centred sampler over a grid whose taps carry the occupancy: a point's
// test reads the occupancy from the taps its velocity sample needs anyway (one cache lookup per
// point instead of two).  Same accepted points, same result.
template <class R>
__device__ __forceinline__ R3<R> backtrace_taps(const RegularVelCached<R>& V, const Dims& d, const StepScale<R>& sc,
                                                int steps, R dtt, R px, R py, R pz) {
    auto open_at = [&](R x, R y, R z) {
        if (!(x >= R(0) && x <= (R)d.W && y >= R(0) && y <= (R)d.H && z >= R(0) && z <= (R)d.D)) return false;
        V.prepare(x, y, z);
        return !V.solid_prepared();
    };
    R3<R> last{px, py, pz};
    bool open = open_at(px, py, pz);
    for (int istep = 0; istep < steps; istep++) {
        if (!open) break;
        last = R3<R>{px, py, pz};
        if (istep + 1 == steps) break;
        const R3<R> vel = V.value();
        px = px - sc(vel.x * dtt);
        py = py - sc(vel.y * dtt);
        pz = pz - sc(vel.z * dtt);
        open = open_at(px, py, pz);
    }
    return last;
}

template <class R, class Vel>
__device__ __forceinline__ R3<R> backtrace(const Vel& V, const uint8_t* __restrict__ occ, const Dims& d,
                                           const StepScale<R>& sc, int steps, R dtt, R px, R py, R pz) {
    if constexpr (std::is_same<Vel, RegularVelCached<R>>::value)
        return backtrace_taps<R>(V, d, sc, steps, dtt, px, py, pz);
    R3<R> last{px, py, pz};
    OccCache oc;
    bool open = trace_open_cached(occ, d, px, py, pz, oc);
    for (int istep = 0; istep < steps; istep++) {
        if (!open) break;
        last = R3<R>{px, py, pz};
        if (istep + 1 == steps) break;
        const R3<R> vel = V(px, py, pz);
        px = px - sc(vel.x * dtt);
        py = py - sc(vel.y * dtt);
        pz = pz - sc(vel.z * dtt);
        open = trace_open_cached(occ, d, px, py, pz, oc);
    }
    return last;
}

}  // namespace zq3d
