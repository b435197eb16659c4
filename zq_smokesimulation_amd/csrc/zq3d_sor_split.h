// zq3d_sor_split.h -- colour-split storage of the SOR and the one-pass red-black iteration of one
// tile, shared by the single-grid kernel (kernels_project.hip) and the ensemble kernel
// (kernels_ensemble.hip).
#pragma once

#include "zq3d_device.h"
#include "zq3d_stencils.h"

namespace zq3d {

// ---- colour-split layout ----
// Colour c = (x+y+z)&1 cells of row (y,z) sit at x = 2*xh + par with par = (c+y+z)&1, stored at
// xh (Wh = (W+1)/2 slots per row; a slot past the row end is not a cell and holds 0).  Every
// neighbour of a colour-c cell has the other colour: x-1 and x+1 are slots xh-1+par and xh+par of
// the same row, y+-1 and z+-1 are slot xh of those rows.  Pressure arrays are padded by one slot
// (zero) on every side: slot xh of a row sits at SPLIT_XOFF + xh of split_row(Wh) floats.  (Rows
// padded to a multiple of 32 floats with the interior on a 128-byte line measured slower at 256^3:
// 51 against 46 us per one-launch iteration.)

constexpr int SPLIT_XOFF = 1;

__host__ __device__ inline int split_row(int Wh) { return Wh + 2; }

__device__ __forceinline__ int split_q(int Wh, int H, int xh, int y, int z) {
    return ((z + 1) * (H + 2) + (y + 1)) * split_row(Wh) + SPLIT_XOFF + xh;
}

// One full red-black iteration of one tile on split storage: src pair (s0 red, s1 black) -> dst
// pair (d0, d1), read-only sources so workgroups never see each other's updates.  The calling
// workgroup (NW waves, every thread calling) owns a tile of TX slots x TY rows and walks
// planes z0..z1-1.  At step t it stages the old black plane t+2 (tile + 2-cell halo) into an LDS
// ring, updates red at plane t on the tile + 1-cell halo (the halo red values are recomputed, not
// exchanged) into a second ring, then updates black at plane t-1 from the new red planes t-2..t.
// The global loads of step t+1 are issued at the top of step t into registers, so they are in
// flight while step t computes.  Per cell the arithmetic and the order of the six-term sum are
// those of a single-colour sweep (k_sor_split): the result is bit-identical to two such sweeps.
// m is the mean subtracted from the right-hand side.  A caller that runs several tiles in one
// workgroup must synchronise the workgroup between them (the LDS rings are reused).
// (TX is the widest tile; a row's last tile may be narrower.)
// GATE: dmax accumulates, as FP32 bits, max |new - old| over the cells this tile updates (red on its
// own slots and planes, black), the size of the iteration's last updates that decides whether a
// gated pass goes on.  ZS: the source pair is taken as all zeros and not read (the first iteration
// of a pass that starts from a zero correction).
constexpr int SPLIT_TX = 64;  // tile width of the single-grid kernel

// Raw buffer access with the hardware's range check (gfx9 buffer resource, byte offsets): a load at
// an offset past the buffer's end returns 0 and a store there is dropped, so a lane with nothing to
// load or store passes SPLIT_NONE instead of branching (no exec-mask juggling, 32-bit offsets).
constexpr uint32_t SPLIT_NONE = 0x80000000u;  // past the end of every buffer the tile touches (< 2 GiB)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t split_buf(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
// off: the lane's byte offset (SPLIT_NONE: nothing); plane: a byte offset common to the wave (a
// plane's start), added by the hardware outside the range check, which sees the lane's part only.
__device__ __forceinline__ float split_ldf(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t plane = 0) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)plane, 0));
}
__device__ __forceinline__ int split_ldb(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t plane = 0) {
    return (int)__builtin_amdgcn_raw_buffer_load_b8(r, (int)off, (int)plane, 0);
}
__device__ __forceinline__ void split_stf(__amdgpu_buffer_rsrc_t r, uint32_t off, float v, uint32_t plane = 0) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)off, (int)plane, 0);
}

struct SplitCellIn {
    float p, b;
    int n;  // negative: not a cell (outside the grid or past the row's end)
};

// Per-thread items of one plane: every index except the plane's own stride is fixed for the whole
// walk, so it is computed once; a colour's cells sit at even or odd x depending on the plane's
// parity, hence two validity bits (for the last slot of a row of odd width).
struct SplitItem {
    int off;   // (ys+1)*split_row(Wh) + SPLIT_XOFF + xs: padded split index without the plane term
    int oh;    // ys*Wh + xs: unpadded split index without the plane term
    int lds;   // LDS index of the item within a ring plane (red and black items: in the red ring)
    int kix;   // red: K-ring index of (xr, yr+1); black: K-ring index of the cell itself
    int bits;  // 1: a cell when the colour's parity is 0, 2: when it is 1, 4: tile-owned, 8: ys odd
};

template <int TX, int TY, int NW, bool GATE = false, bool ZS = false>
__device__ __forceinline__ void sor_split_rb_tile(const uint8_t* __restrict__ n0, const uint8_t* __restrict__ n1,
                                                  const float* __restrict__ b0, const float* __restrict__ b1, float m,
                                                  const float* __restrict__ s0, const float* __restrict__ s1,
                                                  float* __restrict__ d0, float* __restrict__ d1, const Dims& d,
                                                  int Wh, int x0, int y0, int z0, int z1, float omega, float omw,
                                                  unsigned* dmax = nullptr) {
    constexpr int KX = TX + 4, KY = TY + 4, RX = TX + 2, RY = TY + 2;  // ring plane row strides, rows
    constexpr int KPL = KY * KX, RPL = RY * RX;                        // floats per ring plane
    constexpr int NT = 64 * NW;
    constexpr int NK = (KY * KX + NT - 1) / NT;  // staged items per thread
    constexpr int NR = (RY * RX + NT - 1) / NT;  // red items per thread
    constexpr int NB = (TY * TX + NT - 1) / NT;  // black items per thread
    // old black planes t-1 .. t+2 (slot x0-2+c, row y0-2+r) and new red planes t-2 .. t (slot
    // x0-1+c, row y0-1+r); 1/n for n = 1..6
    __shared__ float Ks[5 * KPL + 1];  // + 1: where items outside the tile's region write
    __shared__ float Rs[4 * RPL + 1];
    __shared__ float rtab[8];
    // the tile's own width (the last tile of a row may be narrower); items are numbered row by row
    // over the tile's region so a wave reads contiguous runs of slots
    const int tw = min(TX, Wh - x0), kw = tw + 4, rw = tw + 2;
    const int SP = split_row(Wh), SZ = SP * (d.H + 2), HW = d.H * Wh;
    const uint32_t PB = (uint32_t)SZ * (uint32_t)(d.D + 2) * 4u, CB = (uint32_t)HW * (uint32_t)d.D;
    const __amdgpu_buffer_rsrc_t rs0 = split_buf(s0, PB), rs1 = split_buf(s1, PB), rd0 = split_buf(d0, PB),
                                 rd1 = split_buf(d1, PB), rn0 = split_buf(n0, CB), rn1 = split_buf(n1, CB),
                                 rb0 = split_buf(b0, 4u * CB), rb1 = split_buf(b1, 4u * CB);

    // ---- items, computed once ----
    // staged old black: KY rows x (tw + 4) columns, slot x0-2+c, row y0-2+r
    SplitItem ki[NK];
#pragma unroll
    for (int i = 0; i < NK; i++) {
        const int f = threadIdx.x + NT * i, r = f / kw, c = f % kw;
        const int xs = x0 - 2 + c, ys = y0 - 2 + r;
        const bool mine = r < KY;
        const bool ok = mine && xs >= -1 && xs <= Wh && ys >= -1 && ys <= d.H;
        ki[i] = SplitItem{(ys + 1) * SP + SPLIT_XOFF + xs, 0, mine ? r * KX + c : -1, 0, ok ? 3 : 0};
    }
    // red: RY rows x (tw + 2) columns, slot x0-1+xr, row y0-1+yr
    SplitItem ri[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int f = threadIdx.x + NT * i, yr = f / rw, xr = f % rw;
        const int xs = x0 - 1 + xr, ys = y0 - 1 + yr;
        const bool mine = yr < RY;
        const bool in = mine && xs >= 0 && xs < Wh && ys >= 0 && ys < d.H;
        int bits = 0;
        if (in && 2 * xs < d.W) bits |= 1;
        if (in && 2 * xs + 1 < d.W) bits |= 2;
        if (in && xr >= 1 && xr <= tw && yr >= 1 && yr <= TY) bits |= 4;
        if (ys & 1) bits |= 8;
        ri[i] = SplitItem{(ys + 1) * SP + SPLIT_XOFF + xs, ys * Wh + xs, mine ? yr * RX + xr : -1,
                          mine ? (yr + 1) * KX + xr : KX, bits};
    }
    // black: TY rows x tw columns, slot x0+xb, row y0+yb
    SplitItem bi[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int f = threadIdx.x + NT * i, yb = f / tw, xb = f % tw;
        const int xs = x0 + xb, ys = y0 + yb;
        const bool mine = yb < TY;
        const bool in = mine && ys < d.H;
        int bits = 0;
        if (in && 2 * xs < d.W) bits |= 1;
        if (in && 2 * xs + 1 < d.W) bits |= 2;
        if (ys & 1) bits |= 8;
        bi[i] = SplitItem{(ys + 1) * SP + SPLIT_XOFF + xs, ys * Wh + xs, mine ? yb * RX + xb : -1,
                          mine ? (yb + 2) * KX + xb + 2 : 0, bits};
    }
    if (threadIdx.x < 8) rtab[threadIdx.x] = threadIdx.x >= 1 && threadIdx.x <= 6 ? rcp_count(threadIdx.x) : 0.0f;

    // ---- per-plane work; loads are branch-free (a lane with nothing to fetch reads index 0, a
    // padding slot holding 0) so no path leaves a register half-written by a load ----
    auto load_k = [&](int z, float (&o)[NK]) {
        const bool zin = z >= -1 && z <= d.D;
        const int zo = (z + 1) * SZ;
#pragma unroll
        for (int i = 0; i < NK; i++)
            o[i] = ZS ? 0.0f
                      : split_ldf(rs1, (zin && ki[i].bits) ? 4u * (uint32_t)ki[i].off : SPLIT_NONE,
                                  zin ? 4u * (uint32_t)zo : 0u);
    };
    auto store_k = [&](int z, const float (&o)[NK]) {
        const int K = ((z + 10) % 5) * KPL;
#pragma unroll
        for (int i = 0; i < NK; i++) Ks[ki[i].lds >= 0 ? K + ki[i].lds : 5 * KPL] = o[i];
    };
    // a colour-c cell of row ys, plane t has parity (c + ys + t) & 1
    auto cell = [](const SplitItem& it, int par) { return (it.bits >> par) & 1; };
    auto load_red = [&](int t, SplitCellIn (&o)[NR]) {
        const bool tin = t >= 0 && t < d.D;
        const int zo = (t + 1) * SZ, zh = t * HW;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const bool ok = tin && cell(ri[i], ((ri[i].bits >> 3) ^ t) & 1);
            const uint32_t oh = (uint32_t)ri[i].oh;
            const int n = split_ldb(rn0, ok ? oh : SPLIT_NONE, tin ? (uint32_t)zh : 0u);
            const float bv = split_ldf(rb0, ok ? 4u * oh : SPLIT_NONE, tin ? 4u * (uint32_t)zh : 0u);
            const float pv = ZS ? 0.0f
                                : split_ldf(rs0, ok ? 4u * (uint32_t)ri[i].off : SPLIT_NONE,
                                            tin ? 4u * (uint32_t)zo : 0u);
            // not a cell: the sign bit (the load ran and read 0, so no branch is needed around it)
            o[i] = SplitCellIn{pv, bv, (int)((uint32_t)n | (ok ? 0u : 0x80000000u))};
        }
    };
    auto load_black = [&](int t, SplitCellIn (&o)[NB]) {
        const bool tin = t >= z0 && t < z1;
        const int zh = t * HW;
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const bool ok = tin && cell(bi[i], ((bi[i].bits >> 3) ^ t ^ 1) & 1);
            const uint32_t oh = (uint32_t)bi[i].oh;
            const int n = split_ldb(rn1, ok ? oh : SPLIT_NONE, tin ? (uint32_t)zh : 0u);
            const float bv = split_ldf(rb1, ok ? 4u * oh : SPLIT_NONE, tin ? 4u * (uint32_t)zh : 0u);
            o[i] = SplitCellIn{0.0f, bv, (int)((uint32_t)n | (ok ? 0u : 0x80000000u))};
        }
    };
    // The updates are evaluated for every item a lane holds and selected afterwards (n > 0:
    // updated, n == 0: kept, n < 0: not a cell); the LDS reads stay inside the ring planes.
    auto red = [&](const SplitItem& it, int t, const SplitCellIn& c, bool own_plane) {
        const int par = ((it.bits >> 3) ^ t) & 1;
        const float* K0 = Ks + ((t + 10) % 5) * KPL;
        const float* Km = Ks + ((t + 9) % 5) * KPL;
        const float* Kp = Ks + ((t + 11) % 5) * KPL;
        const int kc = it.kix;  // K[yr+1][xr]
        const float sum = K0[kc + par] + K0[kc + 1 + par] + K0[kc + 1 - KX] + K0[kc + 1 + KX] + Km[kc + 1] + Kp[kc + 1];
        const float bv = c.b - m;
        const float pg = (sum - bv) * rtab[c.n & 7];
        const float upd = omw * c.p + omega * pg;
        const float pc = c.n > 0 ? upd : c.p;
        const bool owned = own_plane && (it.bits & 4);
        split_stf(rd0, c.n >= 0 && owned ? 4u * (uint32_t)it.off : SPLIT_NONE, pc, 4u * (uint32_t)((t + 1) * SZ));
        if (GATE) *dmax = max(*dmax, c.n > 0 && owned ? __float_as_uint(fabsf(pc - c.p)) : 0u);
        Rs[it.lds >= 0 ? ((t + 8) & 3) * RPL + it.lds : 4 * RPL] = c.n >= 0 ? pc : 0.0f;
    };
    auto black = [&](const SplitItem& it, int t, const SplitCellIn& c) {
        const int par = ((it.bits >> 3) ^ t ^ 1) & 1;
        const float p0 = Ks[((t + 10) % 5) * KPL + it.kix];
        const float* R0 = Rs + ((t + 8) & 3) * RPL;
        const float* Rm = Rs + ((t + 7) & 3) * RPL;
        const float* Rp = Rs + ((t + 9) & 3) * RPL;
        const int rc = it.lds + RX;  // R[yb+1][xb]
        const float sum = R0[rc + par] + R0[rc + 1 + par] + R0[rc + 1 - RX] + R0[rc + 1 + RX] + Rm[rc + 1] + Rp[rc + 1];
        const float bv = c.b - m;
        const float pg = (sum - bv) * rtab[c.n & 7];
        const float upd = omw * p0 + omega * pg;
        split_stf(rd1, c.n >= 0 ? 4u * (uint32_t)it.off : SPLIT_NONE, c.n > 0 ? upd : p0, 4u * (uint32_t)((t + 1) * SZ));
        if (GATE) *dmax = max(*dmax, c.n > 0 ? __float_as_uint(fabsf(upd - p0)) : 0u);
    };

    float kc[NK];
    SplitCellIn rc[NR], bc[NB];
    load_k(z0 - 2, kc);
    store_k(z0 - 2, kc);
    load_k(z0 - 1, kc);
    store_k(z0 - 1, kc);
    load_k(z0, kc);
    store_k(z0, kc);
    load_k(z0 + 1, kc);
    load_red(z0 - 1, rc);
    load_black(z0 - 2, bc);  // not a chunk plane: no cells
    // Land the prologue's loads here: with them still pending at the loop entry, the wait inserted
    // before their first use in the loop would also drain the next step's prefetches every step.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int t = z0 - 1; t <= z1; t++) {
        // loads for step t+1 (at t == z1 every lane's guards skip them); unconditional, like the
        // hand-over below, so no path reaches the loop top with a prefetch register still pending
        float kn[NK];
        SplitCellIn rn[NR], bn[NB];
        load_k(t + 3, kn);
        load_red(t + 1, rn);
        load_black(t, bn);
        store_k(t + 2, kc);
        const bool own = t >= z0 && t < z1;
        // whole waves whose items lie outside the tile's region (narrow tiles) skip them
#pragma unroll
        for (int i = 0; i < NR; i++)
            if (ri[i].lds >= 0) red(ri[i], t, rc[i], own);
        __syncthreads();
        if (t - 1 >= z0) {
#pragma unroll
            for (int i = 0; i < NB; i++)
                if (bi[i].lds >= 0) black(bi[i], t - 1, bc[i]);
        }
#pragma unroll
        for (int i = 0; i < NK; i++) kc[i] = kn[i];
#pragma unroll
        for (int i = 0; i < NR; i++) rc[i] = rn[i];
#pragma unroll
        for (int i = 0; i < NB; i++) bc[i] = bn[i];
    }
}


}  // namespace zq3d

namespace zq3d {

// Two full red-black iterations of one tile in one pass over memory: src pair -> dst pair.  Four
// stages run at every plane step t, each one plane behind the previous and separated by a
// workgroup barrier:
//   red1 at plane t (tile + 3-cell halo) from the old black plane ring K (tile + 4),
//   black1 at plane t-1 (tile + 2) from the red1 ring R1,
//   red2 at plane t-2 (tile + 1) from the black1 ring B1 (written to d0 on the tile),
//   black2 at plane t-3 (tile) from the red2 ring R2 (written to d1).
// Each stage's own previous value is its colour's value one iteration earlier, two cells further
// out in that colour's ring (old red comes from s0).  With a barrier between stages, rings of 4
// (K) and 3 (R1, B1, R2) planes suffice: a plane is overwritten only after every stage that reads
// it has passed the barrier that precedes the writer.  A stage runs only on the planes a later
// stage needs (red1 z0-3 .. z1+2, black1 z0-2 .. z1+1, red2 z0-1 .. z1, black2 z0 .. z1-1), so no
// stage reads a ring plane that was not written in this walk.  Per cell the arithmetic is that of
// a single-colour sweep: the result is bit-identical to four single-colour sweeps.
template <int TX, int TY, int NW>
__device__ __forceinline__ void sor_split_rb2_tile(const uint8_t* __restrict__ n0, const uint8_t* __restrict__ n1,
                                                   const float* __restrict__ b0, const float* __restrict__ b1, float m,
                                                   const float* __restrict__ s0, const float* __restrict__ s1,
                                                   float* __restrict__ d0, float* __restrict__ d1, const Dims& d,
                                                   int Wh, int x0, int y0, int z0, int z1, float omega, float omw) {
    constexpr int NT = 64 * NW;
    // ring of halo h: (TY + 2h) rows x (TX + 2h) columns per plane, origin (x0 - h, y0 - h)
    constexpr int XK = TX + 8, YK = TY + 8, X1 = TX + 6, Y1 = TY + 6, X2 = TX + 4, Y2 = TY + 4, X3 = TX + 2,
                  Y3 = TY + 2;
    constexpr int PK = XK * YK, P1 = X1 * Y1, P2 = X2 * Y2, P3 = X3 * Y3;
    constexpr int NK = (PK + NT - 1) / NT, N1 = (P1 + NT - 1) / NT, N2 = (P2 + NT - 1) / NT,
                  N3 = (P3 + NT - 1) / NT, N4 = (TX * TY + NT - 1) / NT;
    __shared__ float Ks[4 * PK];
    __shared__ float R1s[3 * P1];
    __shared__ float B1s[3 * P2];
    __shared__ float R2s[3 * P3];
    __shared__ float rtab[8];
    const int SP = split_row(Wh), SZ = SP * (d.H + 2), HW = d.H * Wh;
    const int tw = min(TX, Wh - x0);
    if (threadIdx.x < 8) rtab[threadIdx.x] = threadIdx.x >= 1 && threadIdx.x <= 6 ? rcp_count(threadIdx.x) : 0.0f;

    // An item is a (row, column) of a stage's region, numbered row by row over the tile's own
    // width, packed in one int: column in bits 0-7, row in bits 8-15, flags in bits 16-19 (1 / 2: a
    // cell when the colour's parity is 0 / 1, 4: tile-owned, 8: ys odd); -1 when the thread has no
    // such item.
    static_assert(TX + 8 <= 256 && TY + 8 <= 256, "row and column fit 8 bits");
    auto make = [&](int i, int h, int rows, int& r, int& c) {
        const int w = tw + 2 * h;
        const int f = threadIdx.x + NT * i;
        r = f / w;
        c = f % w;
        return r < rows;
    };
    auto cell_item = [&](int i, int h) -> int {
        int r, c;
        if (!make(i, h, TY + 2 * h, r, c)) return -1;
        const int xs = x0 - h + c, ys = y0 - h + r;
        const bool in = xs >= 0 && xs < Wh && ys >= 0 && ys < d.H;
        int fl = 0;
        if (in && 2 * xs < d.W) fl |= 1;
        if (in && 2 * xs + 1 < d.W) fl |= 2;
        if (in && r >= h && r < TY + h && c >= h && c < tw + h) fl |= 4;
        if (ys & 1) fl |= 8;
        return c | (r << 8) | (fl << 16);
    };
    int kI[NK], i1[N1], i2[N2], i3[N3], i4[N4];
#pragma unroll
    for (int i = 0; i < NK; i++) {
        int r, c;
        const bool mine = make(i, 4, YK, r, c);
        const int xs = x0 - 4 + c, ys = y0 - 4 + r;
        const bool ok = mine && xs >= -1 && xs <= Wh && ys >= -1 && ys <= d.H;
        kI[i] = mine ? c | (r << 8) | ((ok ? 1 : 0) << 16) : -1;
    }
#pragma unroll
    for (int i = 0; i < N1; i++) i1[i] = cell_item(i, 3);
#pragma unroll
    for (int i = 0; i < N2; i++) i2[i] = cell_item(i, 2);
#pragma unroll
    for (int i = 0; i < N3; i++) i3[i] = cell_item(i, 1);
#pragma unroll
    for (int i = 0; i < N4; i++) i4[i] = cell_item(i, 0);
    auto row = [](int it) { return (it >> 8) & 0xff; };
    auto col = [](int it) { return it & 0xff; };
    auto flags = [](int it) { return (it >> 16) & 0xf; };

    // ---- loads (branch-free: a lane with nothing to fetch reads index 0, a zero padding slot) ----
    auto load_k = [&](int z, float (&o)[NK]) {
        const bool zin = z >= -1 && z <= d.D;
#pragma unroll
        for (int i = 0; i < NK; i++) {
            const int xs = x0 - 4 + col(kI[i]), ys = y0 - 4 + row(kI[i]);
            o[i] = s1[(zin && kI[i] >= 0 && flags(kI[i])) ? (z + 1) * SZ + (ys + 1) * SP + SPLIT_XOFF + xs : 0];
        }
    };
    auto store_k = [&](int z, const float (&o)[NK]) {
        float* K = Ks + ((z + 8) & 3) * PK;
#pragma unroll
        for (int i = 0; i < NK; i++)
            if (kI[i] >= 0) K[row(kI[i]) * XK + col(kI[i])] = o[i];
    };
    // n, b (and for red1 the old value) of a stage's items at plane t; plane outside [lo, hi]: none
    auto load_cells = [&](auto& items, auto& o, int h, int colour, int t, int lo, int hi, bool with_p) {
        const bool tin = t >= lo && t <= hi && t >= 0 && t < d.D;
        const uint8_t* nn = colour ? n1 : n0;
        const float* bb = colour ? b1 : b0;
#pragma unroll
        for (int i = 0; i < (int)(sizeof(o) / sizeof(o[0])); i++) {
            const int it = items[i];
            const int xs = x0 - h + col(it), ys = y0 - h + row(it);
            const int par = ((flags(it) >> 3) ^ t ^ colour) & 1;
            const bool ok = tin && it >= 0 && ((flags(it) >> par) & 1);
            const int oh = ok ? t * HW + ys * Wh + xs : 0;
            const int n = nn[oh];
            const float bv = bb[oh];
            const float pv = with_p ? s0[ok ? (t + 1) * SZ + (ys + 1) * SP + SPLIT_XOFF + xs : 0] : 0.0f;
            o[i] = SplitCellIn{pv, bv, ok ? n : -1};
        }
    };
    // one cell of a stage: in = the other colour's ring (halo h+1) at planes t-1, t, t+1 (slot
    // bases), own = the cell's previous value; writes the stage's ring (halo h) and, on the tile,
    // the destination
    auto update = [&](int it, int h, int colour, int t, const SplitCellIn& c, const float* Im, const float* I0,
                      const float* Ip, int Xin, float own, float* out, int Xout, float* dst, bool to_dst) {
        const int r = row(it), cc = col(it);
        const int par = ((flags(it) >> 3) ^ t ^ colour) & 1;
        const int base = (r + 1) * Xin + cc + 1;
        const float sum = I0[base - 1 + par] + I0[base + par] + I0[base - Xin] + I0[base + Xin] + Im[base] + Ip[base];
        const float bv = c.b - m;
        const float pg = (sum - bv) * rtab[c.n & 7];
        const float upd = omw * own + omega * pg;
        const float pc = c.n > 0 ? upd : own;
        if (out) out[r * Xout + cc] = c.n >= 0 ? pc : 0.0f;
        if (to_dst && c.n >= 0 && (flags(it) & 4)) {
            const int xs = x0 - h + cc, ys = y0 - h + r;
            dst[(t + 1) * SZ + (ys + 1) * SP + SPLIT_XOFF + xs] = pc;
        }
    };
    auto slot3 = [](int z) { return (z + 9) % 3; };
    auto kslot = [](int z) { return (z + 8) & 3; };

    // prefetched one step ahead: the first touches of a plane (old black, red1's and black1's cells);
    // red2 and black2 read counts and right-hand sides red1 and black1 brought into L2 1-2 steps earlier
    float kc[NK];
    SplitCellIn c1[N1], c2[N2];
    load_k(z0 - 4, kc);
    store_k(z0 - 4, kc);
    load_k(z0 - 3, kc);
    store_k(z0 - 3, kc);
    load_k(z0 - 2, kc);
    store_k(z0 - 2, kc);
    load_k(z0 - 1, kc);
    load_cells(i1, c1, 3, 0, z0 - 3, z0 - 3, z1 + 2, true);
    load_cells(i2, c2, 2, 1, z0 - 4, z0 - 2, z1 + 1, false);  // not a needed plane: no cells
    // land the prologue's loads before the loop (see sor_split_rb_tile)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int t = z0 - 3; t <= z1 + 2; t++) {
        float kn[NK];
        SplitCellIn e1[N1], e2[N2];
        load_k(t + 3, kn);
        load_cells(i1, e1, 3, 0, t + 1, z0 - 3, z1 + 2, true);
        load_cells(i2, e2, 2, 1, t, z0 - 2, z1 + 1, false);
        store_k(t + 2, kc);
        {  // red1 at t
            const float *Im = Ks + kslot(t - 1) * PK, *I0 = Ks + kslot(t) * PK, *Ip = Ks + kslot(t + 1) * PK;
            float* out = R1s + slot3(t) * P1;
#pragma unroll
            for (int i = 0; i < N1; i++)
                if (i1[i] >= 0) update(i1[i], 3, 0, t, c1[i], Im, I0, Ip, XK, c1[i].p, out, X1, nullptr, false);
        }
        __syncthreads();
        if (t - 1 >= z0 - 2 && t - 1 <= z1 + 1) {  // black1 at t-1; own: old black, K two cells out
            const int tb = t - 1;
            const float *Im = R1s + slot3(tb - 1) * P1, *I0 = R1s + slot3(tb) * P1, *Ip = R1s + slot3(tb + 1) * P1;
            const float* own = Ks + kslot(tb) * PK;
            float* out = B1s + slot3(tb) * P2;
#pragma unroll
            for (int i = 0; i < N2; i++)
                if (i2[i] >= 0)
                    update(i2[i], 2, 1, tb, c2[i], Im, I0, Ip, X1, own[(row(i2[i]) + 2) * XK + col(i2[i]) + 2], out,
                           X2, nullptr, false);
        }
        __syncthreads();
        if (t - 2 >= z0 - 1 && t - 2 <= z1) {  // red2 at t-2; own: red1, R1 two cells out
            const int tr = t - 2;
            const float *Im = B1s + slot3(tr - 1) * P2, *I0 = B1s + slot3(tr) * P2, *Ip = B1s + slot3(tr + 1) * P2;
            const float* own = R1s + slot3(tr) * P1;
            float* out = R2s + slot3(tr) * P3;
            const bool owned_plane = tr >= z0 && tr < z1;
            SplitCellIn c3[N3];
            load_cells(i3, c3, 1, 0, tr, z0 - 1, z1, false);
#pragma unroll
            for (int i = 0; i < N3; i++)
                if (i3[i] >= 0)
                    update(i3[i], 1, 0, tr, c3[i], Im, I0, Ip, X2, own[(row(i3[i]) + 2) * X1 + col(i3[i]) + 2], out,
                           X3, d0, owned_plane);
        }
        __syncthreads();
        if (t - 3 >= z0 && t - 3 < z1) {  // black2 at t-3; own: black1, B1 two cells out
            const int tb = t - 3;
            const float *Im = R2s + slot3(tb - 1) * P3, *I0 = R2s + slot3(tb) * P3, *Ip = R2s + slot3(tb + 1) * P3;
            const float* own = B1s + slot3(tb) * P2;
            SplitCellIn c4[N4];
            load_cells(i4, c4, 0, 1, tb, z0, z1 - 1, false);
#pragma unroll
            for (int i = 0; i < N4; i++)
                if (i4[i] >= 0)
                    update(i4[i], 0, 1, tb, c4[i], Im, I0, Ip, X3, own[(row(i4[i]) + 2) * X2 + col(i4[i]) + 2],
                           nullptr, 0, d1, true);
        }
#pragma unroll
        for (int i = 0; i < NK; i++) kc[i] = kn[i];
#pragma unroll
        for (int i = 0; i < N1; i++) c1[i] = e1[i];
#pragma unroll
        for (int i = 0; i < N2; i++) c2[i] = e2[i];
    }
}

}  // namespace zq3d
