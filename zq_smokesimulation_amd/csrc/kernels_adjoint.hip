// kernels_adjoint.hip -- hand-written adjoint (vector-Jacobian products) of one smoke step (gfx950).
//
// No reference counterpart: the reference has no gradients.  Each kernel is the transpose of a
// forward kernel of kernels_force.hip / kernels_advect.hip / kernels_project.hip / kernels_misc.hip,
// templated on the real type R (float, or double for the FP64 check mode):
//   * force + attenuation: masks and coefficients of the face update, buoyancy, the density term,
//     the control force, and vorticity confinement through |w|, its normalized gradient N and
//     N x w (k_adj_force_cells, k_adj_vort_scale, k_adj_curl, k_adj_force_faces);
//   * semi-Lagrangian advection: the sampled field's adjoint is scattered to the 8 taps of the end
//     point, and the end point's adjoint is carried back along the stored backtrace path through
//     the velocity it sampled (k_adj_advect_*); the break decisions (domain, solids) are held
//     fixed, as the derivative of a step function is zero almost everywhere;
//   * MAC <-> cell averaging and the masked face apply (k_adj_cells_to_faces, k_adj_vel_to_faces).
// Scatters use hardware float/double atomics (global_atomic_add), so gradients are reproducible
// to rounding only.
#include "zq3d_adjoint.h"
#include "zq3d_device.h"

namespace zq3d {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid_for(int nx, int ny, int nz) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz); }

__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// LDS atomics (ds_add_f32 / ds_add_f64).  The window pointer travels through a struct, where the
// compiler no longer sees that it addresses LDS and would emit flat atomics; the explicit address
// space keeps them on the LDS path.
typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(3))) double lds_double;
__device__ __forceinline__ void lds_add(float* p, float v) {
    __hip_atomic_fetch_add((lds_float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(double* p, double v) {
    __hip_atomic_fetch_add((lds_double*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Texels and weights of one trilinear sample (the forward's axis_setup).
template <class R>
struct Tri {
    int i0, i1, j0, j1, k0, k1;
    R ax, ay, az;
};

template <class R>
__device__ __forceinline__ Tri<R> tri(int nx, int ny, int nz, R x, R y, R z) {
    Tri<R> t;
    axis_setup(x, nx, t.i0, t.i1, t.ax);
    axis_setup(y, ny, t.j0, t.j1, t.ay);
    axis_setup(z, nz, t.k0, t.k1, t.az);
    return t;
}

// g[taps] += a * weights
template <class R>
__device__ __forceinline__ void scatter(R* __restrict__ g, int nx, int ny, const Tri<R>& t, R a) {
    if (a == R(0)) return;
    const R bx = R(1) - t.ax, by = R(1) - t.ay, bz = R(1) - t.az;
    const int r00 = (t.k0 * ny + t.j0) * nx, r10 = (t.k0 * ny + t.j1) * nx;
    const int r01 = (t.k1 * ny + t.j0) * nx, r11 = (t.k1 * ny + t.j1) * nx;
    atomic_add(g + r00 + t.i0, a * bx * by * bz);
    atomic_add(g + r00 + t.i1, a * t.ax * by * bz);
    atomic_add(g + r10 + t.i0, a * bx * t.ay * bz);
    atomic_add(g + r10 + t.i1, a * t.ax * t.ay * bz);
    atomic_add(g + r01 + t.i0, a * bx * by * t.az);
    atomic_add(g + r01 + t.i1, a * t.ax * by * t.az);
    atomic_add(g + r11 + t.i0, a * bx * t.ay * t.az);
    atomic_add(g + r11 + t.i1, a * t.ax * t.ay * t.az);
}

// Gradient of the trilinear sample with respect to the texel coordinates.  A clamped axis has
// equal taps, so its derivative is zero, as the forward's is.
template <class R>
__device__ __forceinline__ R3<R> grad_taps(R t000, R t100, R t010, R t110, R t001, R t101, R t011, R t111,
                                           const Tri<R>& t) {
    const R c00 = lerp(t000, t100, t.ax), c10 = lerp(t010, t110, t.ax);
    const R c01 = lerp(t001, t101, t.ax), c11 = lerp(t011, t111, t.ax);
    const R c0 = lerp(c00, c10, t.ay), c1 = lerp(c01, c11, t.ay);
    R3<R> g;
    g.x = lerp(lerp(t100 - t000, t110 - t010, t.ay), lerp(t101 - t001, t111 - t011, t.ay), t.az);
    g.y = lerp(c10 - c00, c11 - c01, t.az);
    g.z = c1 - c0;
    return g;
}

template <class R>
__device__ __forceinline__ R3<R> grad3(const R* __restrict__ g, int nx, int ny, const Tri<R>& t) {
    const int r00 = (t.k0 * ny + t.j0) * nx, r10 = (t.k0 * ny + t.j1) * nx;
    const int r01 = (t.k1 * ny + t.j0) * nx, r11 = (t.k1 * ny + t.j1) * nx;
    return grad_taps<R>(g[r00 + t.i0], g[r00 + t.i1], g[r10 + t.i0], g[r10 + t.i1], g[r01 + t.i0], g[r01 + t.i1],
                        g[r11 + t.i0], g[r11 + t.i1], t);
}

// Velocity-field adjoints of the two samplers.  vjp(p, a) = J(p)^T a, where J is the Jacobian of
// the sampled velocity with respect to the position; put(p, a) adds a (one vector) to the
// velocity field's adjoint at the taps of p.
template <class R>
struct RegularVelAdj {
    const vec4_t<R>* vel4;
    R* vb;  // 3N SoA adjoint of the cell velocity
    Dims d;
    __device__ __forceinline__ R3<R> vjp(R px, R py, R pz, R3<R> a) const {
        const Tri<R> t = tri<R>(d.W, d.H, d.D, px - R(0.5), py - R(0.5), pz - R(0.5));
        const int nx = d.W, ny = d.H;
        const int r00 = (t.k0 * ny + t.j0) * nx, r10 = (t.k0 * ny + t.j1) * nx;
        const int r01 = (t.k1 * ny + t.j0) * nx, r11 = (t.k1 * ny + t.j1) * nx;
        const vec4_t<R> q000 = vel4[r00 + t.i0], q100 = vel4[r00 + t.i1], q010 = vel4[r10 + t.i0],
                        q110 = vel4[r10 + t.i1];
        const vec4_t<R> q001 = vel4[r01 + t.i0], q101 = vel4[r01 + t.i1], q011 = vel4[r11 + t.i0],
                        q111 = vel4[r11 + t.i1];
        // gradient of the scalar field a . v
        auto dot = [&](const vec4_t<R>& q) { return a.x * q.x + a.y * q.y + a.z * q.z; };
        return grad_taps<R>(dot(q000), dot(q100), dot(q010), dot(q110), dot(q001), dot(q101), dot(q011), dot(q111),
                            t);
    }
    __device__ __forceinline__ void put(R px, R py, R pz, R3<R> a) const {
        const Tri<R> t = tri<R>(d.W, d.H, d.D, px - R(0.5), py - R(0.5), pz - R(0.5));
        const size_t N = (size_t)d.W * d.H * d.D;
        scatter<R>(vb, d.W, d.H, t, a.x);
        scatter<R>(vb + N, d.W, d.H, t, a.y);
        scatter<R>(vb + 2 * N, d.W, d.H, t, a.z);
    }
};

template <class R>
struct MacVelAdj {
    const R* u;
    const R* v;
    const R* w;
    R* ub;  // adjoints of the three face grids
    R* vb;
    R* wb;
    Dims d;
    __device__ __forceinline__ R3<R> vjp(R px, R py, R pz, R3<R> a) const {
        const Tri<R> tu = tri<R>(d.W + 1, d.H, d.D, px, py - R(0.5), pz - R(0.5));
        const Tri<R> tv = tri<R>(d.W, d.H + 1, d.D, px - R(0.5), py, pz - R(0.5));
        const Tri<R> tw = tri<R>(d.W, d.H, d.D + 1, px - R(0.5), py - R(0.5), pz);
        const R3<R> gu = grad3<R>(u, d.W + 1, d.H, tu);
        const R3<R> gv = grad3<R>(v, d.W, d.H + 1, tv);
        const R3<R> gw = grad3<R>(w, d.W, d.H, tw);
        return R3<R>{a.x * gu.x + a.y * gv.x + a.z * gw.x, a.x * gu.y + a.y * gv.y + a.z * gw.y,
                     a.x * gu.z + a.y * gv.z + a.z * gw.z};
    }
    __device__ __forceinline__ void put(R px, R py, R pz, R3<R> a) const {
        scatter<R>(ub, d.W + 1, d.H, tri<R>(d.W + 1, d.H, d.D, px, py - R(0.5), pz - R(0.5)), a.x);
        scatter<R>(vb, d.W, d.H + 1, tri<R>(d.W, d.H + 1, d.D, px - R(0.5), py, pz - R(0.5)), a.y);
        scatter<R>(wb, d.W, d.H, tri<R>(d.W, d.H, d.D + 1, px - R(0.5), py - R(0.5), pz), a.z);
    }
};

// The forward backtrace again, keeping the accepted positions q_0 .. q_{n-1} (the end point is
// q_{n-1}; n == 0 means the start was rejected and the end point is the start).
template <class R, class Vel>
__device__ __forceinline__ int trace_path(const Vel& V, const uint8_t* __restrict__ occ, const Dims& d, R h,
                                          int steps, R dtt, R px, R py, R pz, R3<R>* q) {
    OccCache oc;
    bool open = trace_open_cached(occ, d, px, py, pz, oc);
    R3<R> vel = V(px, py, pz);
    int n = 0;
    int istep = 0;
    do {
        if (!open) break;
        q[n++] = R3<R>{px, py, pz};
        px = px - (vel.x * dtt) / h;
        py = py - (vel.y * dtt) / h;
        pz = pz - (vel.z * dtt) / h;
        open = trace_open_cached(occ, d, px, py, pz, oc);
        vel = V(px, py, pz);
        istep++;
    } while (istep < steps);
    return n;
}

// Carry the end point's adjoint g back along the path: q_{m+1} = q_m - c V(q_m), c = dtt/h.
template <class R, class VA>
__device__ __forceinline__ void reverse_path(const VA& A, const R3<R>* q, int n, R c, R3<R> g) {
    for (int m = n - 2; m >= 0; m--) {
        const R3<R> p = q[m];
        A.put(p.x, p.y, p.z, R3<R>{-c * g.x, -c * g.y, -c * g.z});
        const R3<R> jt = A.vjp(p.x, p.y, p.z, g);
        g = R3<R>{g.x - c * jt.x, g.y - c * jt.y, g.z - c * jt.z};
    }
}

// ---- tile-local accumulation of the cell-grid scatters ----
// The backtrace paths of a tile of neighbouring cells stay close together, so the taps they
// scatter into fill a small box.  A block first traces its paths and reduces the box of every tap
// it will touch; when the box fits CAP cells the scatters go to an LDS window (ds atomics) that is
// flushed to HBM once, one global atomic per nonzero entry, instead of one per tap per path step.
// Larger boxes (very fast flow) use global atomics directly.  NA arrays of W*H*D cells.
constexpr int TX = 8, TY = 8, TZ = 4;

template <class R, int NA, int CAP>
struct TileWin {
    R* s;  // LDS, NA * CAP
    R* const* g;
    Dims d;
    int ox, oy, oz, wx, wy;
    bool lds;
    __device__ __forceinline__ void add(int a, int i, int j, int k, R v) {
        if (lds) lds_add(&s[a * CAP + ((k - oz) * wy + (j - oy)) * wx + (i - ox)], v);
        else atomic_add(g[a] + cidx(d, i, j, k), v);
    }
    __device__ __forceinline__ void scatter(int a, const Tri<R>& t, R v) {
        if (v == R(0)) return;
        const R bx = R(1) - t.ax, by = R(1) - t.ay, bz = R(1) - t.az;
        add(a, t.i0, t.j0, t.k0, v * bx * by * bz);
        add(a, t.i1, t.j0, t.k0, v * t.ax * by * bz);
        add(a, t.i0, t.j1, t.k0, v * bx * t.ay * bz);
        add(a, t.i1, t.j1, t.k0, v * t.ax * t.ay * bz);
        add(a, t.i0, t.j0, t.k1, v * bx * by * t.az);
        add(a, t.i1, t.j0, t.k1, v * t.ax * by * t.az);
        add(a, t.i0, t.j1, t.k1, v * bx * t.ay * t.az);
        add(a, t.i1, t.j1, t.k1, v * t.ax * t.ay * t.az);
    }
};

__device__ __forceinline__ void box_grow(int* b, int i0, int i1, int j0, int j1, int k0, int k1) {
    b[0] = min(b[0], i0);
    b[1] = min(b[1], j0);
    b[2] = min(b[2], k0);
    b[3] = max(b[3], i1);
    b[4] = max(b[4], j1);
    b[5] = max(b[5], k1);
}

template <class R>
__device__ __forceinline__ Tri<R> cell_tri(const Dims& d, const R3<R>& p) {
    return tri<R>(d.W, d.H, d.D, p.x - R(0.5), p.y - R(0.5), p.z - R(0.5));
}

// Reduces the per-thread boxes, decides LDS vs global, zeroes the window; all threads call it.
template <class R, int NA, int CAP>
__device__ __forceinline__ bool tile_setup(TileWin<R, NA, CAP>& w, const int* mine, int* sbox, int tid) {
    if (tid == 0) {
        sbox[0] = sbox[1] = sbox[2] = 0x7fffffff;
        sbox[3] = sbox[4] = sbox[5] = -1;
    }
    __syncthreads();
    if (mine[3] >= 0) {
        atomicMin(&sbox[0], mine[0]);
        atomicMin(&sbox[1], mine[1]);
        atomicMin(&sbox[2], mine[2]);
        atomicMax(&sbox[3], mine[3]);
        atomicMax(&sbox[4], mine[4]);
        atomicMax(&sbox[5], mine[5]);
    }
    __syncthreads();
    if (sbox[3] < 0) return false;  // nothing to scatter in this tile
    w.ox = sbox[0];
    w.oy = sbox[1];
    w.oz = sbox[2];
    w.wx = sbox[3] - sbox[0] + 1;
    w.wy = sbox[4] - sbox[1] + 1;
    const int vol = w.wx * w.wy * (sbox[5] - sbox[2] + 1);
    w.lds = vol <= CAP;
    if (w.lds)
        for (int x = tid; x < NA * CAP; x += TX * TY * TZ) w.s[x] = R(0);
    __syncthreads();
    return true;
}

template <class R, int NA, int CAP>
__device__ __forceinline__ void tile_flush(TileWin<R, NA, CAP>& w, const int* sbox, int tid) {
    __syncthreads();
    if (!w.lds) return;
    const int wz = sbox[5] - sbox[2] + 1;
    const int vol = w.wx * w.wy * wz;
    for (int x = tid; x < vol; x += TX * TY * TZ) {
        const int i = w.ox + x % w.wx;
        const int j = w.oy + (x / w.wx) % w.wy;
        const int k = w.oz + x / (w.wx * w.wy);
        for (int a = 0; a < NA; a++) {
            const R v = w.s[a * CAP + x];
            if (v != R(0)) atomic_add(w.g[a] + cidx(w.d, i, j, k), v);
        }
    }
}

// Carry the end point's adjoint back along the path into the velocity adjoint (window arrays 0..2),
// one path point at a time, last to first.  Consecutive path points usually share a trilinear
// cell: the 8 taps are loaded once per cell (Cell4Taps) and the path's weighted contributions to
// them are summed in registers, reaching the window only when the path moves to another cell.
template <class R, int NA, int CAP>
struct RevTile {
    const vec4_t<R>* vel4;
    Dims d;
    TileWin<R, NA, CAP>* w;
    R c;
    R3<R> g;
    Cell4Taps<R> cc;
    R acc[8][3];
    __device__ __forceinline__ void flush() {
        if (cc.i0 < 0) return;
        const int ti[8] = {cc.i0, cc.i1, cc.i0, cc.i1, cc.i0, cc.i1, cc.i0, cc.i1};
        const int tj[8] = {cc.j0, cc.j0, cc.j1, cc.j1, cc.j0, cc.j0, cc.j1, cc.j1};
        const int tk[8] = {cc.k0, cc.k0, cc.k0, cc.k0, cc.k1, cc.k1, cc.k1, cc.k1};
#pragma unroll
        for (int t = 0; t < 8; t++)
#pragma unroll
            for (int a = 0; a < 3; a++)
                if (acc[t][a] != R(0)) w->add(a, ti[t], tj[t], tk[t], acc[t][a]);
    }
    __device__ __forceinline__ void step(const R3<R>& p) {
        const Tri<R> t = cell_tri<R>(d, p);
        if (!cc.same(t.i0, t.i1, t.j0, t.j1, t.k0, t.k1)) {
            flush();
            cc.fetch(vel4, d.W, d.H, t.i0, t.i1, t.j0, t.j1, t.k0, t.k1);
#pragma unroll
            for (int x = 0; x < 8; x++) acc[x][0] = acc[x][1] = acc[x][2] = R(0);
        }
        // put -c g at p: weights of the 8 taps (order t000 t100 t010 t110 t001 t101 t011 t111)
        const R bx = R(1) - t.ax, by = R(1) - t.ay, bz = R(1) - t.az;
        const R wt[8] = {bx * by * bz,   t.ax * by * bz,   bx * t.ay * bz,   t.ax * t.ay * bz,
                         bx * by * t.az, t.ax * by * t.az, bx * t.ay * t.az, t.ax * t.ay * t.az};
        const R gx = -c * g.x, gy = -c * g.y, gz = -c * g.z;
#pragma unroll
        for (int x = 0; x < 8; x++) {
            acc[x][0] += gx * wt[x];
            acc[x][1] += gy * wt[x];
            acc[x][2] += gz * wt[x];
        }
        // J^T g from the cached taps
        auto dot = [&](const vec4_t<R>& v) { return g.x * v.x + g.y * v.y + g.z * v.z; };
        const R3<R> jt = grad_taps<R>(dot(cc.t000), dot(cc.t100), dot(cc.t010), dot(cc.t110), dot(cc.t001),
                                      dot(cc.t101), dot(cc.t011), dot(cc.t111), t);
        g = R3<R>{g.x - c * jt.x, g.y - c * jt.y, g.z - c * jt.z};
    }
};

// trace_path that also grows box by the cell box of every accepted point, and sets *last to the
// last accepted point (the start when none is accepted, whose box then goes in instead): one pass
// over the path instead of a second one over the stored points.
template <class R, class Vel>
__device__ __forceinline__ int trace_box(const Vel& V, const uint8_t* __restrict__ occ, const Dims& d, R h,
                                         int steps, R dtt, R px, R py, R pz, R3<R>* q, R3<R>* last, int* box) {
    OccCache oc;
    bool open = trace_open_cached(occ, d, px, py, pz, oc);
    R3<R> vel = V(px, py, pz);
    int n = 0;
    int istep = 0;
    *last = R3<R>{px, py, pz};
    do {
        if (!open) break;
        const R3<R> p{px, py, pz};
        q[n++] = p;
        const Tri<R> t = cell_tri<R>(d, p);
        box_grow(box, t.i0, t.i1, t.j0, t.j1, t.k0, t.k1);
        *last = p;
        px = px - (vel.x * dtt) / h;
        py = py - (vel.y * dtt) / h;
        pz = pz - (vel.z * dtt) / h;
        open = trace_open_cached(occ, d, px, py, pz, oc);
        vel = V(px, py, pz);
        istep++;
    } while (istep < steps);
    if (n == 0) {
        const Tri<R> t = cell_tri<R>(d, *last);
        box_grow(box, t.i0, t.i1, t.j0, t.j1, t.k0, t.k1);
    }
    return n;
}

// rev.step on the stored points n-2 .. 0 (n >= 2), the next point's load issued before the current
// step so its latency (the points live in scratch memory) overlaps the step's work
template <class R, class Rev>
__device__ __forceinline__ void reverse_points(Rev& rev, const R3<R>* q, int n) {
    R3<R> cur = q[n - 2];
    for (int m = n - 2; m >= 0; m--) {
        const R3<R> nxt = q[m > 0 ? m - 1 : 0];
        rev.step(cur);
        cur = nxt;
    }
}

// inRegular_outRegular velocity advection adjoint, tiled (see k_adj_advect_cells).  The FP32 build
// is held to 80 VGPRs (one spilled), which lets six waves per SIMD run instead of five.
template <class R, int CAP>
__global__ __launch_bounds__(TX* TY* TZ) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 6 : 1))) void
k_adj_advect_cells_tile(RegularVel<R> V, R* __restrict__ vb, const uint8_t* __restrict__ occ,
                        const R* __restrict__ advb, R h, int steps, R dtt) {
    __shared__ R swin[3 * CAP];
    __shared__ int sbox[6];
    const Dims d = V.d;
    const int tid = threadIdx.x;
    const int i = blockIdx.x * TX + (tid % TX);
    const int j = blockIdx.y * TY + (tid / TX) % TY;
    const int k = blockIdx.z * TZ + tid / (TX * TY);
    const size_t N = (size_t)d.W * d.H * d.D;
    R* const gptr[3] = {vb, vb + N, vb + 2 * N};
    TileWin<R, 3, CAP> w{swin, gptr, d, 0, 0, 0, 0, 0, false};
    R3<R> q[ZQ3D_ADJ_MAX_STEPS];
    int n = 0;
    R3<R> ab{R(0), R(0), R(0)}, last{R(0), R(0), R(0)};
    int mine[6] = {0x7fffffff, 0x7fffffff, 0x7fffffff, -1, -1, -1};
    bool act = i < d.W && j < d.H && k < d.D;
    if (act) {
        const int o = cidx(d, i, j, k);
        ab = R3<R>{advb[o], advb[N + o], advb[2 * N + o]};
        act = !(ab.x == R(0) && ab.y == R(0) && ab.z == R(0));
    }
    if (act) {
        const RegularVelCached<R> Vc{V.vel4, d, {}};
        n = trace_box<R>(Vc, occ, d, h, steps, dtt, i + R(0.5), j + R(0.5), k + R(0.5), q, &last, mine);
    }
    if (!tile_setup<R, 3, CAP>(w, mine, sbox, tid)) return;
    if (act) {
        const Tri<R> t = cell_tri<R>(d, last);
        w.scatter(0, t, ab.x);
        w.scatter(1, t, ab.y);
        w.scatter(2, t, ab.z);
        if (n >= 2) {
            const RegularVelAdj<R> A{V.vel4, nullptr, d};
            RevTile<R, 3, CAP> rev{V.vel4, d, &w, dtt / h, A.vjp(last.x, last.y, last.z, ab), {}, {}};
            reverse_points(rev, q, n);
            rev.flush();
        }
    }
    tile_flush<R, 3, CAP>(w, sbox, tid);
}

// Scalar advection adjoint through the cell velocity, tiled: window arrays 0..2 velocity adjoint,
// 3 and 4 the adjoints of the advected T and rho.
template <class R, int CAP>
__global__ __launch_bounds__(TX* TY* TZ) void k_adj_advect_scalar_tile(RegularVel<R> V, R* __restrict__ vb,
                                                                       const uint8_t* __restrict__ occ,
                                                                       const R* __restrict__ Tin,
                                                                       const R* __restrict__ Rin,
                                                                       const R* __restrict__ Tb,
                                                                       const R* __restrict__ Rb, R* __restrict__ Tb_in,
                                                                       R* __restrict__ Rb_in, R h, int steps, R dtt) {
    __shared__ R swin[5 * CAP];
    __shared__ int sbox[6];
    const Dims d = V.d;
    const int tid = threadIdx.x;
    const int i = blockIdx.x * TX + (tid % TX);
    const int j = blockIdx.y * TY + (tid / TX) % TY;
    const int k = blockIdx.z * TZ + tid / (TX * TY);
    const size_t N = (size_t)d.W * d.H * d.D;
    R* const gptr[5] = {vb, vb + N, vb + 2 * N, Tb_in, Rb_in};
    TileWin<R, 5, CAP> w{swin, gptr, d, 0, 0, 0, 0, 0, false};
    R3<R> q[ZQ3D_ADJ_MAX_STEPS];
    int n = 0;
    R tb = R(0), rb = R(0);
    R3<R> last{R(0), R(0), R(0)};
    int mine[6] = {0x7fffffff, 0x7fffffff, 0x7fffffff, -1, -1, -1};
    bool act = i < d.W && j < d.H && k < d.D;
    if (act) {
        const int o = cidx(d, i, j, k);
        tb = Tb[o];
        rb = Rb[o];
        act = !(tb == R(0) && rb == R(0));
    }
    if (act) {
        const RegularVelCached<R> Vc{V.vel4, d, {}};
        n = trace_box<R>(Vc, occ, d, h, steps, dtt, i + R(0.5), j + R(0.5), k + R(0.5), q, &last, mine);
    }
    if (!tile_setup<R, 5, CAP>(w, mine, sbox, tid)) return;
    if (act) {
        const Tri<R> t = cell_tri<R>(d, last);
        w.scatter(3, t, tb);
        w.scatter(4, t, rb);
        if (n >= 2) {
            const R3<R> gt = grad3<R>(Tin, d.W, d.H, t), gr = grad3<R>(Rin, d.W, d.H, t);
            const R3<R> g{tb * gt.x + rb * gr.x, tb * gt.y + rb * gr.y, tb * gt.z + rb * gr.z};
            RevTile<R, 5, CAP> rev{V.vel4, d, &w, dtt / h, g, {}, {}};
            reverse_points(rev, q, n);
            rev.flush();
        }
    }
    tile_flush<R, 5, CAP>(w, sbox, tid);
}

// ---- scalar advection (Advect_Scalar_*, ZQ_CUDA_Advection3D.cu:379-469) ----
// Tb/Rb: adjoints of the outputs; Tb_in/Rb_in (zeroed by the caller): adjoints of the inputs.
template <class R, class Vel, class VA>
__global__ __launch_bounds__(BX* BY) void k_adj_advect_scalar(Vel V, VA A, const uint8_t* __restrict__ occ,
                                                              const R* __restrict__ Tin, const R* __restrict__ Rin,
                                                              const R* __restrict__ Tb, const R* __restrict__ Rb,
                                                              R* __restrict__ Tb_in, R* __restrict__ Rb_in, Dims d,
                                                              R h, int steps, R dtt) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const int o = cidx(d, i, j, k);
    const R tb = Tb[o], rb = Rb[o];
    if (tb == R(0) && rb == R(0)) return;
    R3<R> q[ZQ3D_ADJ_MAX_STEPS];
    const R sx = i + R(0.5), sy = j + R(0.5), sz = k + R(0.5);
    const int n = trace_path<R>(V, occ, d, h, steps, dtt, sx, sy, sz, q);
    const R3<R> last = n ? q[n - 1] : R3<R>{sx, sy, sz};
    const Tri<R> t = tri<R>(d.W, d.H, d.D, last.x - R(0.5), last.y - R(0.5), last.z - R(0.5));
    scatter<R>(Tb_in, d.W, d.H, t, tb);
    scatter<R>(Rb_in, d.W, d.H, t, rb);
    if (n < 2) return;
    const R3<R> gt = grad3<R>(Tin, d.W, d.H, t), gr = grad3<R>(Rin, d.W, d.H, t);
    const R3<R> g{tb * gt.x + rb * gr.x, tb * gt.y + rb * gr.y, tb * gt.z + rb * gr.z};
    reverse_path<R>(A, q, n, dtt / h, g);
}

// ---- velocity advection, inRegular_outRegular (:76-113 + the masked cell->face apply) ----
// advb: 3N SoA adjoint of the advected cell velocity; the sample and the path both read vel4.
template <class R>
__global__ __launch_bounds__(BX* BY) void k_adj_advect_cells(RegularVel<R> V, RegularVelAdj<R> A,
                                                             const uint8_t* __restrict__ occ,
                                                             const R* __restrict__ advb, Dims d, R h, int steps,
                                                             R dtt) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const int o = cidx(d, i, j, k);
    const size_t N = (size_t)d.W * d.H * d.D;
    const R3<R> ab{advb[o], advb[N + o], advb[2 * N + o]};
    if (ab.x == R(0) && ab.y == R(0) && ab.z == R(0)) return;
    R3<R> q[ZQ3D_ADJ_MAX_STEPS];
    const R sx = i + R(0.5), sy = j + R(0.5), sz = k + R(0.5);
    const int n = trace_path<R>(V, occ, d, h, steps, dtt, sx, sy, sz, q);
    const R3<R> last = n ? q[n - 1] : R3<R>{sx, sy, sz};
    A.put(last.x, last.y, last.z, ab);
    if (n < 2) return;
    reverse_path<R>(A, q, n, dtt / h, A.vjp(last.x, last.y, last.z, ab));
}

// ---- velocity advection from face centres (inMAC_outMAC :232-377, inRegular_outMAC :115-230) ----
// outb: face adjoint of the advected component (already masked); the sample reads component AXIS.
template <class R, int AXIS, class Vel, class VA>
__global__ __launch_bounds__(BX* BY) void k_adj_advect_faces(Vel V, VA A, const uint8_t* __restrict__ occ,
                                                             const R* __restrict__ outb, Dims d, R h, int steps,
                                                             R dtt) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    const R ob = outb[fi];
    if (ob == R(0)) return;
    const R sx = AXIS == 0 ? (R)i : i + R(0.5);
    const R sy = AXIS == 1 ? (R)j : j + R(0.5);
    const R sz = AXIS == 2 ? (R)k : k + R(0.5);
    R3<R> q[ZQ3D_ADJ_MAX_STEPS];
    const int n = trace_path<R>(V, occ, d, h, steps, dtt, sx, sy, sz, q);
    const R3<R> last = n ? q[n - 1] : R3<R>{sx, sy, sz};
    const R3<R> a{AXIS == 0 ? ob : R(0), AXIS == 1 ? ob : R(0), AXIS == 2 ? ob : R(0)};
    A.put(last.x, last.y, last.z, a);
    if (n < 2) return;
    reverse_path<R>(A, q, n, dtt / h, A.vjp(last.x, last.y, last.z, a));
}

// Face mask of the advection apply (Apply_Advect_Velocity_Result_*, :471-650)
template <int AXIS>
__device__ __forceinline__ bool adv_mask(const uint8_t* __restrict__ occ, const Dims& d, int i, int j, int k,
                                         int is_open, int& ci, int& step, bool& domain) {
    const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
    step = AXIS == 0 ? 1 : (AXIS == 1 ? d.W : d.W * d.H);
    ci = cidx(d, AXIS == 0 ? (i < d.W ? i : d.W - 1) : i, AXIS == 1 ? (j < d.H ? j : d.H - 1) : j,
              AXIS == 2 ? (k < d.D ? k : d.D - 1) : k);
    domain = (pos == 0 || pos == lim);
    return domain ? (is_open && !occ[ci]) : (!occ[ci] && !occ[ci - step]);
}

// Adjoint of the cell -> face average + apply: advb (cells, 3N SoA) gathers the masked face
// adjoints (1/2 per interior face, 1 per domain face).
template <class R>
__global__ __launch_bounds__(BX* BY) void k_adj_faces_to_cells(const R* __restrict__ ub, const R* __restrict__ vb,
                                                               const R* __restrict__ wb,
                                                               const uint8_t* __restrict__ occ, R* __restrict__ advb,
                                                               Dims d, int is_open) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const int o = cidx(d, i, j, k);
    const size_t N = (size_t)d.W * d.H * d.D;
    int ci, st;
    bool dom;
    R a = R(0);
    if (adv_mask<0>(occ, d, i, j, k, is_open, ci, st, dom)) a += (dom ? R(1) : R(0.5)) * ub[uidx(d, i, j, k)];
    if (adv_mask<0>(occ, d, i + 1, j, k, is_open, ci, st, dom)) a += (dom ? R(1) : R(0.5)) * ub[uidx(d, i + 1, j, k)];
    advb[o] = a;
    a = R(0);
    if (adv_mask<1>(occ, d, i, j, k, is_open, ci, st, dom)) a += (dom ? R(1) : R(0.5)) * vb[vidx(d, i, j, k)];
    if (adv_mask<1>(occ, d, i, j + 1, k, is_open, ci, st, dom)) a += (dom ? R(1) : R(0.5)) * vb[vidx(d, i, j + 1, k)];
    advb[N + o] = a;
    a = R(0);
    if (adv_mask<2>(occ, d, i, j, k, is_open, ci, st, dom)) a += (dom ? R(1) : R(0.5)) * wb[widx(d, i, j, k)];
    if (adv_mask<2>(occ, d, i, j, k + 1, is_open, ci, st, dom)) a += (dom ? R(1) : R(0.5)) * wb[widx(d, i, j, k + 1)];
    advb[2 * N + o] = a;
}

// Face adjoint after the apply: masked faces were overwritten (their incoming adjoint moves to the
// advected value: out = it, or zero here), unmasked ones pass through.  With CELLS the incoming
// face value also fed the cell velocity (MAC -> cell average), whose adjoint cb is added back.
template <class R, int AXIS, bool CELLS>
__global__ __launch_bounds__(BX* BY) void k_adj_vel_to_faces(R* __restrict__ face, R* __restrict__ outb,
                                                             const R* __restrict__ cb,
                                                             const uint8_t* __restrict__ occ, Dims d, int is_open) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    int ci, st;
    bool dom;
    const bool m = adv_mask<AXIS>(occ, d, i, j, k, is_open, ci, st, dom);
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    R x = face[fi];
    if (outb) outb[fi] = m ? x : R(0);
    if (m) x = R(0);
    if (CELLS) {
        const R* c = cb + (size_t)AXIS * d.W * d.H * d.D;
        const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
        const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
        // cells on both sides of the face that exist: cell pos-1 (face is its high face) and pos
        if (pos > 0) x += R(0.5) * c[pos < lim ? ci - st : ci];
        if (pos < lim) x += R(0.5) * c[ci];
    }
    face[fi] = x;
}

// ---- force + attenuation ----
// Face mask and coefficient of the force update (k_force_faces): both cells fluid (1/2 dt each),
// a domain face needs its inside cell (dt).
template <int AXIS>
__device__ __forceinline__ bool force_mask(const uint8_t* __restrict__ occ, const Dims& d, int i, int j, int k,
                                           bool& domain) {
    int ci, st;
    // the force mask is the advection mask with is_open = 1
    return adv_mask<AXIS>(occ, d, i, j, k, 1, ci, st, domain);
}

// Per cell: f_bar from the masked faces, the control gradient, T/rho adjoints, and the first half
// of the confinement adjoint (w_bar from N x w, the gradient adjoint g_bar).
template <class R>
__global__ __launch_bounds__(BX* BY) void k_adj_force_cells(const R* __restrict__ ub, const R* __restrict__ vb,
                                                            const R* __restrict__ wb,
                                                            const uint8_t* __restrict__ occ,
                                                            const vec4_t<R>* __restrict__ vort, R* __restrict__ Tb,
                                                            R* __restrict__ Rb, R* __restrict__ ctrl_b,
                                                            R* __restrict__ gb, R* __restrict__ wbar, Dims d, R dt,
                                                            R dth, R aV, R aT, R aR, R beta, R alpha, R ch) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const int o = cidx(d, i, j, k);
    const size_t N = (size_t)d.W * d.H * d.D;
    const int HW = d.W * d.H;
    bool dom;
    R fx = R(0), fy = R(0), fz = R(0);
    if (force_mask<0>(occ, d, i, j, k, dom)) fx += (dom ? dt : dth) * (aV * ub[uidx(d, i, j, k)]);
    if (force_mask<0>(occ, d, i + 1, j, k, dom)) fx += (dom ? dt : dth) * (aV * ub[uidx(d, i + 1, j, k)]);
    if (force_mask<1>(occ, d, i, j, k, dom)) fy += (dom ? dt : dth) * (aV * vb[vidx(d, i, j, k)]);
    if (force_mask<1>(occ, d, i, j + 1, k, dom)) fy += (dom ? dt : dth) * (aV * vb[vidx(d, i, j + 1, k)]);
    if (force_mask<2>(occ, d, i, j, k, dom)) fz += (dom ? dt : dth) * (aV * wb[widx(d, i, j, k)]);
    if (force_mask<2>(occ, d, i, j, k + 1, dom)) fz += (dom ? dt : dth) * (aV * wb[widx(d, i, j, k + 1)]);
    if (ctrl_b) {
        ctrl_b[(size_t)o * 3 + 0] = fx;
        ctrl_b[(size_t)o * 3 + 1] = fy;
        ctrl_b[(size_t)o * 3 + 2] = fz;
    }
    Tb[o] = aT * Tb[o] + beta * fy;
    Rb[o] = aR * Rb[o] - alpha * fy;
    // confinement f = ch (N x w): w_bar = ch (f_bar x N), N_bar = ch (w x f_bar)
    const R s0 = vort[o].w;
    R gx, gy, gz;
    if (i == 0) gx = vort[o + 1].w - s0;
    else if (i == d.W - 1) gx = s0 - vort[o - 1].w;
    else gx = R(0.5) * (vort[o + 1].w - vort[o - 1].w);
    if (j == 0) gy = vort[o + d.W].w - s0;
    else if (j == d.H - 1) gy = s0 - vort[o - d.W].w;
    else gy = R(0.5) * (vort[o + d.W].w - vort[o - d.W].w);
    if (k == 0) gz = vort[o + HW].w - s0;
    else if (k == d.D - 1) gz = s0 - vort[o - HW].w;
    else gz = R(0.5) * (vort[o + HW].w - vort[o - HW].w);
    const R len = zsqrt(gx * gx + gy * gy + gz * gz);
    R nx = gx, ny = gy, nz = gz;
    if (len != R(0)) {
        nx = gx / len;
        ny = gy / len;
        nz = gz / len;
    }
    const vec4_t<R> wv = vort[o];
    wbar[o] = ch * (fy * nz - fz * ny);
    wbar[N + o] = ch * (fz * nx - fx * nz);
    wbar[2 * N + o] = ch * (fx * ny - fy * nx);
    const R bx = ch * (wv.y * fz - wv.z * fy);
    const R by = ch * (wv.z * fx - wv.x * fz);
    const R bz = ch * (wv.x * fy - wv.y * fx);
    R ox = bx, oy = by, oz = bz;
    if (len != R(0)) {
        const R dn = nx * bx + ny * by + nz * bz;
        ox = (bx - nx * dn) / len;
        oy = (by - ny * dn) / len;
        oz = (bz - nz * dn) / len;
    }
    gb[o] = ox;
    gb[N + o] = oy;
    gb[2 * N + o] = oz;
}

// Coefficient with which |w| at index c enters the gradient at index cp along an axis of n cells
// (one-sided at the two ends, half central differences inside).
__device__ __forceinline__ float grad_coef(int cp, int c, int n) {
    if (cp == 0) return c == 1 ? 1.0f : (c == 0 ? -1.0f : 0.0f);
    if (cp == n - 1) return c == n - 1 ? 1.0f : (c == n - 2 ? -1.0f : 0.0f);
    return c == cp + 1 ? 0.5f : (c == cp - 1 ? -0.5f : 0.0f);
}

// |w|_bar by the transposed gradient stencil; w_bar += |w|_bar w/|w| on interior cells; zero on the
// border, where w is the constant 0.
template <class R>
__global__ __launch_bounds__(BX* BY) void k_adj_vort_scale(const vec4_t<R>* __restrict__ vort,
                                                           const R* __restrict__ gb, R* __restrict__ wbar, Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const int o = cidx(d, i, j, k);
    const size_t N = (size_t)d.W * d.H * d.D;
    const bool interior = i >= 1 && i < d.W - 1 && j >= 1 && j < d.H - 1 && k >= 1 && k < d.D - 1;
    if (!interior) {
        wbar[o] = wbar[N + o] = wbar[2 * N + o] = R(0);
        return;
    }
    const int HW = d.W * d.H;
    R sb = R(0);
    for (int cp = i - 1; cp <= i + 1; cp++)
        if (cp >= 0 && cp < d.W) sb += (R)grad_coef(cp, i, d.W) * gb[o + (cp - i)];
    for (int cp = j - 1; cp <= j + 1; cp++)
        if (cp >= 0 && cp < d.H) sb += (R)grad_coef(cp, j, d.H) * gb[N + o + (cp - j) * d.W];
    for (int cp = k - 1; cp <= k + 1; cp++)
        if (cp >= 0 && cp < d.D) sb += (R)grad_coef(cp, k, d.D) * gb[2 * N + o + (size_t)(cp - k) * HW];
    const vec4_t<R> wv = vort[o];
    if (wv.w > R(0)) {
        const R s = sb / wv.w;
        wbar[o] += s * wv.x;
        wbar[N + o] += s * wv.y;
        wbar[2 * N + o] += s * wv.z;
    }
}

// Transposed curl: cell-velocity adjoint cb from w_bar of the neighbours (w = curl / 2h on
// interior cells; w_bar is zero elsewhere).
template <class R>
__global__ __launch_bounds__(BX* BY) void k_adj_curl(const R* __restrict__ wbar, R* __restrict__ cb, Dims d,
                                                     R two_h) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.W || j >= d.H) return;
    const int o = cidx(d, i, j, k);
    const size_t N = (size_t)d.W * d.H * d.D;
    const int HW = d.W * d.H;
    const R* wx = wbar;
    const R* wy = wbar + N;
    const R* wz = wbar + 2 * N;
    auto at = [&](const R* g, int di, int dj, int dk) -> R {
        const int ii = i + di, jj = j + dj, kk = k + dk;
        if (ii < 0 || ii >= d.W || jj < 0 || jj >= d.H || kk < 0 || kk >= d.D) return R(0);
        return g[o + di + dj * d.W + dk * HW];
    };
    // wy has +cu(k+1) - cu(k-1), wz has -cu(j+1) + cu(j-1)
    cb[o] = (at(wy, 0, 0, -1) - at(wy, 0, 0, 1) - at(wz, 0, -1, 0) + at(wz, 0, 1, 0)) / two_h;
    // wx has -cv(k+1) + cv(k-1), wz has +cv(i+1) - cv(i-1)
    cb[N + o] = (-at(wx, 0, 0, -1) + at(wx, 0, 0, 1) + at(wz, -1, 0, 0) - at(wz, 1, 0, 0)) / two_h;
    // wx has +cw(j+1) - cw(j-1), wy has -cw(i+1) + cw(i-1)
    cb[2 * N + o] = (at(wx, 0, -1, 0) - at(wx, 0, 1, 0) - at(wy, -1, 0, 0) + at(wy, 1, 0, 0)) / two_h;
}

// Face adjoint of the force stage: the face update's own factor (aV on masked faces) plus half of
// the cell-velocity adjoint of the two cells the face borders.
template <class R, int AXIS>
__global__ __launch_bounds__(BX* BY) void k_adj_force_faces(R* __restrict__ face, const R* __restrict__ cb,
                                                            const uint8_t* __restrict__ occ, Dims d, R aV) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    int ci, st;
    bool dom;
    const bool m = adv_mask<AXIS>(occ, d, i, j, k, 1, ci, st, dom);
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    R x = face[fi];
    if (m) x = aV * x;
    const R* c = cb + (size_t)AXIS * d.W * d.H * d.D;
    const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
    if (pos > 0) x += R(0.5) * c[pos < lim ? ci - st : ci];
    if (pos < lim) x += R(0.5) * c[ci];
    face[fi] = x;
}

inline dim3 tile_grid(const Dims& d) { return dim3((d.W + TX - 1) / TX, (d.H + TY - 1) / TY, (d.D + TZ - 1) / TZ); }

// LDS window capacity in cells of a tile kernel with NA window arrays: at most 2048 (FP32) / 1024
// (FP64) cells, and few enough that four 256-thread blocks (four waves per SIMD) fit the CU's
// 160 KiB of LDS with their 24-byte box.  For the five-array scalar window that is 2044 / 1022
// cells: 2048 would need 24 bytes more than a quarter and leave only three blocks on a CU.
template <class R, int NA>
constexpr int tile_cap() {
    return (40960 - 64) / (NA * (int)sizeof(R)) < 8192 / (int)sizeof(R) ? (40960 - 64) / (NA * (int)sizeof(R))
                                                                          : 8192 / (int)sizeof(R);
}
static_assert(tile_cap<float, 5>() == 2044 && tile_cap<float, 3>() == 2048, "FP32 window capacities");
static_assert(tile_cap<double, 5>() == 1022 && tile_cap<double, 3>() == 1024, "FP64 window capacities");

}  // namespace

template <class R>
void launch_adj_force(const Dims& d, R* ub, R* vb, R* wb, R* Tb, R* Rb, R* ctrl_b, const uint8_t* occ,
                      const vec4_t<R>* vort, const ForceParams& fp, float aV, float aT, float aR, R* gb, R* wbar,
                      R* cb, hipStream_t s) {
    const dim3 blk(BX, BY);
    const R ch = (R)(fp.eps * fp.h);
    const R dth = (R)(fp.dt * 0.5f);
    hipLaunchKernelGGL(k_adj_force_cells<R>, grid_for(d.W, d.H, d.D), blk, 0, s, ub, vb, wb, occ, vort, Tb, Rb, ctrl_b,
                       gb, wbar, d, (R)fp.dt, dth, (R)aV, (R)aT, (R)aR, (R)fp.beta, (R)fp.alpha, ch);
    hipLaunchKernelGGL(k_adj_vort_scale<R>, grid_for(d.W, d.H, d.D), blk, 0, s, vort, gb, wbar, d);
    hipLaunchKernelGGL(k_adj_curl<R>, grid_for(d.W, d.H, d.D), blk, 0, s, wbar, cb, d, R(2) * (R)fp.h);
    hipLaunchKernelGGL((k_adj_force_faces<R, 0>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, ub, cb, occ, d, (R)aV);
    hipLaunchKernelGGL((k_adj_force_faces<R, 1>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, vb, cb, occ, d, (R)aV);
    hipLaunchKernelGGL((k_adj_force_faces<R, 2>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, wb, cb, occ, d, (R)aV);
}

template <class R>
void launch_adj_advect_velocity(const Dims& d, const R* u, const R* v, const R* w, R* ub, R* vb, R* wb,
                                const uint8_t* occ, int is_open, int mode, const AdvParams& ap,
                                const vec4_t<R>* vel4, R* cellb, R* advb, R* outb, hipStream_t s) {
    const dim3 blk(BX, BY);
    const size_t N = (size_t)d.W * d.H * d.D;
    const size_t Nu = (size_t)(d.W + 1) * d.H * d.D, Nv = (size_t)d.W * (d.H + 1) * d.D;
    const R h = (R)ap.h, dtt = (R)ap.dtt;
    if (mode == 0) {
        // advb = transposed apply; cellb = velocity-field adjoint (sample + path); faces get both
        hipLaunchKernelGGL(k_adj_faces_to_cells<R>, grid_for(d.W, d.H, d.D), blk, 0, s, ub, vb, wb, occ, advb, d,
                           is_open);
        launch_fill_adj<R>(cellb, 3 * N, s);
        const RegularVel<R> V{vel4, d};
        hipLaunchKernelGGL((k_adj_advect_cells_tile<R, tile_cap<R, 3>()>), tile_grid(d), dim3(TX * TY * TZ), 0, s, V,
                           cellb, occ, advb, h, ap.steps, dtt);
        hipLaunchKernelGGL((k_adj_vel_to_faces<R, 0, true>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, ub, (R*)nullptr,
                           cellb, occ, d, is_open);
        hipLaunchKernelGGL((k_adj_vel_to_faces<R, 1, true>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, vb, (R*)nullptr,
                           cellb, occ, d, is_open);
        hipLaunchKernelGGL((k_adj_vel_to_faces<R, 2, true>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, wb, (R*)nullptr,
                           cellb, occ, d, is_open);
        return;
    }
    // face-centred modes: outb = masked face adjoint (the advected values' adjoint); faces keep
    // their unmasked part, then receive the velocity-field adjoint
    R* ob_u = outb;
    R* ob_v = outb + Nu;
    R* ob_w = outb + Nu + Nv;
    hipLaunchKernelGGL((k_adj_vel_to_faces<R, 0, false>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, ub, ob_u,
                       (const R*)nullptr, occ, d, is_open);
    hipLaunchKernelGGL((k_adj_vel_to_faces<R, 1, false>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, vb, ob_v,
                       (const R*)nullptr, occ, d, is_open);
    hipLaunchKernelGGL((k_adj_vel_to_faces<R, 2, false>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, wb, ob_w,
                       (const R*)nullptr, occ, d, is_open);
    if (mode == 2) {
        const MacVel<R> V{u, v, w, d};
        const MacVelAdj<R> A{u, v, w, ub, vb, wb, d};
        hipLaunchKernelGGL((k_adj_advect_faces<R, 0, MacVel<R>, MacVelAdj<R>>), grid_for(d.W + 1, d.H, d.D), blk, 0, s,
                           V, A, occ, ob_u, d, h, ap.steps, dtt);
        hipLaunchKernelGGL((k_adj_advect_faces<R, 1, MacVel<R>, MacVelAdj<R>>), grid_for(d.W, d.H + 1, d.D), blk, 0, s,
                           V, A, occ, ob_v, d, h, ap.steps, dtt);
        hipLaunchKernelGGL((k_adj_advect_faces<R, 2, MacVel<R>, MacVelAdj<R>>), grid_for(d.W, d.H, d.D + 1), blk, 0, s,
                           V, A, occ, ob_w, d, h, ap.steps, dtt);
        return;
    }
    // mode 4: samples the cell velocity; its adjoint goes back to the faces through the average
    launch_fill_adj<R>(cellb, 3 * N, s);
    const RegularVel<R> V{vel4, d};
    const RegularVelAdj<R> A{vel4, cellb, d};
    hipLaunchKernelGGL((k_adj_advect_faces<R, 0, RegularVel<R>, RegularVelAdj<R>>), grid_for(d.W + 1, d.H, d.D), blk, 0,
                       s, V, A, occ, ob_u, d, h, ap.steps, dtt);
    hipLaunchKernelGGL((k_adj_advect_faces<R, 1, RegularVel<R>, RegularVelAdj<R>>), grid_for(d.W, d.H + 1, d.D), blk, 0,
                       s, V, A, occ, ob_v, d, h, ap.steps, dtt);
    hipLaunchKernelGGL((k_adj_advect_faces<R, 2, RegularVel<R>, RegularVelAdj<R>>), grid_for(d.W, d.H, d.D + 1), blk, 0,
                       s, V, A, occ, ob_w, d, h, ap.steps, dtt);
    launch_add_cells_to_faces<R>(d, cellb, ub, vb, wb, s);
}

template <class R>
void launch_adj_advect_scalar(const Dims& d, const R* u, const R* v, const R* w, R* ub, R* vb, R* wb,
                              const uint8_t* occ, const R* Tin, const R* Rin, const R* Tb, const R* Rb, R* Tb_in,
                              R* Rb_in, int mac, const AdvParams& ap, const vec4_t<R>* vel4, R* cellb,
                              hipStream_t s) {
    const dim3 blk(BX, BY);
    const size_t N = (size_t)d.W * d.H * d.D;
    const R h = (R)ap.h, dtt = (R)ap.dtt;
    launch_fill_adj<R>(Tb_in, N, s);
    launch_fill_adj<R>(Rb_in, N, s);
    if (mac) {
        const MacVel<R> V{u, v, w, d};
        const MacVelAdj<R> A{u, v, w, ub, vb, wb, d};
        hipLaunchKernelGGL((k_adj_advect_scalar<R, MacVel<R>, MacVelAdj<R>>), grid_for(d.W, d.H, d.D), blk, 0, s, V, A,
                           occ, Tin, Rin, Tb, Rb, Tb_in, Rb_in, d, h, ap.steps, dtt);
        return;
    }
    launch_fill_adj<R>(cellb, 3 * N, s);
    const RegularVel<R> V{vel4, d};
    hipLaunchKernelGGL((k_adj_advect_scalar_tile<R, tile_cap<R, 5>()>), tile_grid(d), dim3(TX * TY * TZ), 0, s, V, cellb,
                       occ, Tin, Rin, Tb, Rb, Tb_in, Rb_in, h, ap.steps, dtt);
    launch_add_cells_to_faces<R>(d, cellb, ub, vb, wb, s);
}

namespace {

// faces += 1/2 (cell adjoint of the two cells they border): the transposed MAC -> cell average
template <class R, int AXIS>
__global__ __launch_bounds__(BX* BY) void k_add_cells_to_faces(R* __restrict__ face, const R* __restrict__ cb,
                                                               Dims d) {
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k = blockIdx.z;
    const int nx = d.W + (AXIS == 0), ny = d.H + (AXIS == 1);
    if (i >= nx || j >= ny) return;
    const int pos = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int lim = AXIS == 0 ? d.W : (AXIS == 1 ? d.H : d.D);
    const int step = AXIS == 0 ? 1 : (AXIS == 1 ? d.W : d.W * d.H);
    const int ci = cidx(d, AXIS == 0 ? (i < d.W ? i : d.W - 1) : i, AXIS == 1 ? (j < d.H ? j : d.H - 1) : j,
                        AXIS == 2 ? (k < d.D ? k : d.D - 1) : k);
    const R* c = cb + (size_t)AXIS * d.W * d.H * d.D;
    const int fi = AXIS == 0 ? uidx(d, i, j, k) : (AXIS == 1 ? vidx(d, i, j, k) : widx(d, i, j, k));
    R x = face[fi];
    if (pos > 0) x += R(0.5) * c[pos < lim ? ci - step : ci];
    if (pos < lim) x += R(0.5) * c[ci];
    face[fi] = x;
}

template <class T>
__global__ __launch_bounds__(256) void k_zero(T* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t x = i; x < n; x += stride) p[x] = T(0);
}

}  // namespace

template <class R>
void launch_add_cells_to_faces(const Dims& d, const R* cb, R* ub, R* vb, R* wb, hipStream_t s) {
    const dim3 blk(BX, BY);
    hipLaunchKernelGGL((k_add_cells_to_faces<R, 0>), grid_for(d.W + 1, d.H, d.D), blk, 0, s, ub, cb, d);
    hipLaunchKernelGGL((k_add_cells_to_faces<R, 1>), grid_for(d.W, d.H + 1, d.D), blk, 0, s, vb, cb, d);
    hipLaunchKernelGGL((k_add_cells_to_faces<R, 2>), grid_for(d.W, d.H, d.D + 1), blk, 0, s, wb, cb, d);
}

template <class R>
void launch_fill_adj(R* p, size_t n, hipStream_t s) {
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero<R>, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
}

#define ZQ3D_ADJ_INSTANTIATE(R)                                                                                        \
    template void launch_adj_force<R>(const Dims&, R*, R*, R*, R*, R*, R*, const uint8_t*, const vec4_t<R>*,           \
                                      const ForceParams&, float, float, float, R*, R*, R*, hipStream_t);               \
    template void launch_adj_advect_velocity<R>(const Dims&, const R*, const R*, const R*, R*, R*, R*,                 \
                                                const uint8_t*, int, int, const AdvParams&, const vec4_t<R>*, R*, R*,  \
                                                R*, hipStream_t);                                                      \
    template void launch_adj_advect_scalar<R>(const Dims&, const R*, const R*, const R*, R*, R*, R*, const uint8_t*,  \
                                              const R*, const R*, const R*, const R*, R*, R*, int, const AdvParams&,   \
                                              const vec4_t<R>*, R*, hipStream_t);                                      \
    template void launch_add_cells_to_faces<R>(const Dims&, const R*, R*, R*, R*, hipStream_t);                        \
    template void launch_fill_adj<R>(R*, size_t, hipStream_t);
ZQ3D_ADJ_INSTANTIATE(float)
ZQ3D_ADJ_INSTANTIATE(double)
#undef ZQ3D_ADJ_INSTANTIATE

}  // namespace zq3d
