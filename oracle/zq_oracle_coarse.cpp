// zq_oracle_coarse.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's coarse + sub-grid projection (COARSE_OPEN_POISSON_SOR, the
// default globalGridSolver, ZQ_SmokeSimulationUtils3D.cpp:3688-3795): coarse face open fractions
// and coarse occupancy, fine -> coarse face averages, the face-fraction Poisson solve on the
// coarse grid, coarse -> fine correction, and the sub-projection of the regular blocks.  The
// sub-block operator and the blocks holding solid cells are restated in zq_oracle.py (numpy).
// Only tests/ may load this; the product library never links or calls it.
#include <cstdint>
#include <cstring>
#include <vector>

#include "zq_oracle.h"

namespace {

struct CD {  // fine grid W, H, D and block size s; coarse Wc, Hc, Dc
    int W, H, D, s, Wc, Hc, Dc;
    size_t c(int i, int j, int k) const { return ((size_t)k * H + j) * W + i; }
    size_t fu(int i, int j, int k) const { return ((size_t)k * H + j) * (W + 1) + i; }
    size_t fv(int i, int j, int k) const { return ((size_t)k * (H + 1) + j) * W + i; }
    size_t fw(int i, int j, int k) const { return ((size_t)k * H + j) * W + i; }
    size_t cc(int i, int j, int k) const { return ((size_t)k * Hc + j) * Wc + i; }
    size_t cu(int i, int j, int k) const { return ((size_t)k * Hc + j) * (Wc + 1) + i; }
    size_t cv(int i, int j, int k) const { return ((size_t)k * (Hc + 1) + j) * Wc + i; }
    size_t cw(int i, int j, int k) const { return ((size_t)k * Hc + j) * Wc + i; }
    size_t cp(int i, int j, int k) const { return ((size_t)(k + 1) * (Hc + 2) + (j + 1)) * (Wc + 2) + (i + 1); }
};

CD make(int W, int H, int D, int s) { return CD{W, H, D, s, W / s, H / s, D / s}; }

}  // namespace

extern "C" {

// Coarse face open fractions 1 - (solid sub-faces)/s^2, where a sub-face of an interior coarse
// face is solid when either fine cell beside it is, and of a domain face when the inside cell is;
// a coarse cell is solid when all six of its fractions are 0 (ZQ_InitCoarseGrid3D,
// ZQ_SmokeSimulationUtils3D.cpp:1666-1860).  regular[b] = 1 when block b holds no solid fine cell
// (ZQ_SubProjection3D, :3318-3340).
void orc_coarse_setup(int W, int H, int D, int s, const uint8_t* occ, float* rU, float* rV, float* rW,
                      uint8_t* cocc, uint8_t* regular) {
    const CD d = make(W, H, D, s);
    const int sf = s * s;
    for (int k = 0; k < d.Dc; k++)
        for (int j = 0; j < d.Hc; j++)
            for (int i = 0; i <= d.Wc; i++) {
                int n = 0;
                for (int sk = 0; sk < s; sk++)
                    for (int sj = 0; sj < s; sj++) {
                        const int y = j * s + sj, z = k * s + sk;
                        bool solid;
                        if (i == 0) solid = occ[d.c(0, y, z)];
                        else if (i == d.Wc) solid = occ[d.c(W - 1, y, z)];
                        else solid = occ[d.c(i * s, y, z)] || occ[d.c(i * s - 1, y, z)];
                        n += solid;
                    }
                rU[d.cu(i, j, k)] = 1.0f - (float)n / sf;
            }
    for (int k = 0; k < d.Dc; k++)
        for (int j = 0; j <= d.Hc; j++)
            for (int i = 0; i < d.Wc; i++) {
                int n = 0;
                for (int sk = 0; sk < s; sk++)
                    for (int si = 0; si < s; si++) {
                        const int x = i * s + si, z = k * s + sk;
                        bool solid;
                        if (j == 0) solid = occ[d.c(x, 0, z)];
                        else if (j == d.Hc) solid = occ[d.c(x, H - 1, z)];
                        else solid = occ[d.c(x, j * s, z)] || occ[d.c(x, j * s - 1, z)];
                        n += solid;
                    }
                rV[d.cv(i, j, k)] = 1.0f - (float)n / sf;
            }
    for (int k = 0; k <= d.Dc; k++)
        for (int j = 0; j < d.Hc; j++)
            for (int i = 0; i < d.Wc; i++) {
                int n = 0;
                for (int sj = 0; sj < s; sj++)
                    for (int si = 0; si < s; si++) {
                        const int x = i * s + si, y = j * s + sj;
                        bool solid;
                        if (k == 0) solid = occ[d.c(x, y, 0)];
                        else if (k == d.Dc) solid = occ[d.c(x, y, D - 1)];
                        else solid = occ[d.c(x, y, k * s)] || occ[d.c(x, y, k * s - 1)];
                        n += solid;
                    }
                rW[d.cw(i, j, k)] = 1.0f - (float)n / sf;
            }
    for (int k = 0; k < d.Dc; k++)
        for (int j = 0; j < d.Hc; j++)
            for (int i = 0; i < d.Wc; i++) {
                cocc[d.cc(i, j, k)] = rU[d.cu(i, j, k)] == 0 && rU[d.cu(i + 1, j, k)] == 0 && rV[d.cv(i, j, k)] == 0 &&
                                      rV[d.cv(i, j + 1, k)] == 0 && rW[d.cw(i, j, k)] == 0 &&
                                      rW[d.cw(i, j, k + 1)] == 0;
                int solid = 0;
                for (int sk = 0; sk < s; sk++)
                    for (int sj = 0; sj < s; sj++)
                        for (int si = 0; si < s; si++) solid |= occ[d.c(i * s + si, j * s + sj, k * s + sk)];
                regular[d.cc(i, j, k)] = !solid;
            }
}

// 1/c of a fluid coarse cell for the face-fraction operator below (0 when c == 0): c is the count
// of out-of-domain sides that are open, then + r(x-) + r(x+) + r(y-) + r(y+) + r(z-) + r(z+) over
// the in-domain sides, in that order (ZQ_InitMatOpenPoisson3D, ZQ_SmokeSimulationUtils3D.cpp:570-589,
// where every out-of-domain side is open).  A solid neighbour's face has fraction 0.
float orc_coarse_inv_diag(int Wc, int Hc, int Dc, int i, int j, int k, const float* rU, const float* rV,
                          const float* rW, int open_mask, int flux) {
    const CD d{Wc, Hc, Dc, 1, Wc, Hc, Dc};
    if (flux) {
        // the flux KKT systems (ZQ_InitMatOpenFlux3D :749-936, ZQ_InitMatClosedFlux3D :937-1150):
        // every side by its own fraction, an out-of-domain side only when it is open
        float c = 0.0f;
        if (i > 0 || (open_mask & ORC_OPEN_XLO)) c += rU[d.cu(i, j, k)];
        if (i < Wc - 1 || (open_mask & ORC_OPEN_XHI)) c += rU[d.cu(i + 1, j, k)];
        if (j > 0 || (open_mask & ORC_OPEN_YLO)) c += rV[d.cv(i, j, k)];
        if (j < Hc - 1 || (open_mask & ORC_OPEN_YHI)) c += rV[d.cv(i, j + 1, k)];
        if (k > 0 || (open_mask & ORC_OPEN_ZLO)) c += rW[d.cw(i, j, k)];
        if (k < Dc - 1 || (open_mask & ORC_OPEN_ZHI)) c += rW[d.cw(i, j, k + 1)];
        return c > 0.0f ? 1.0f / c : 0.0f;
    }
    int nopen = 0;
    nopen += (i == 0) && (open_mask & ORC_OPEN_XLO);
    nopen += (i == Wc - 1) && (open_mask & ORC_OPEN_XHI);
    nopen += (j == 0) && (open_mask & ORC_OPEN_YLO);
    nopen += (j == Hc - 1) && (open_mask & ORC_OPEN_YHI);
    nopen += (k == 0) && (open_mask & ORC_OPEN_ZLO);
    nopen += (k == Dc - 1) && (open_mask & ORC_OPEN_ZHI);
    float c = (float)nopen;
    if (i > 0) c += rU[d.cu(i, j, k)];
    if (i < Wc - 1) c += rU[d.cu(i + 1, j, k)];
    if (j > 0) c += rV[d.cv(i, j, k)];
    if (j < Hc - 1) c += rV[d.cv(i, j + 1, k)];
    if (k > 0) c += rW[d.cw(i, j, k)];
    if (k < Dc - 1) c += rW[d.cw(i, j, k + 1)];
    return c > 0.0f ? 1.0f / c : 0.0f;
}

// Coarse face = sum of its s x s fine sub-faces (outer loop the slower axis) / s^2
// (ZQ_MapFine2Coarse3D, ZQ_SmokeSimulationUtils3D.cpp:2170-2258).
void orc_fine_to_coarse(int W, int H, int D, int s, const float* u, const float* v, const float* w, float* cu,
                        float* cv, float* cw) {
    const CD d = make(W, H, D, s);
    const int sf = s * s;
    for (int k = 0; k < d.Dc; k++)
        for (int j = 0; j < d.Hc; j++)
            for (int i = 0; i <= d.Wc; i++) {
                float a = 0.0f;
                for (int sk = 0; sk < s; sk++)
                    for (int sj = 0; sj < s; sj++) a += u[d.fu(i * s, j * s + sj, k * s + sk)];
                cu[d.cu(i, j, k)] = a / sf;
            }
    for (int k = 0; k < d.Dc; k++)
        for (int j = 0; j <= d.Hc; j++)
            for (int i = 0; i < d.Wc; i++) {
                float a = 0.0f;
                for (int sk = 0; sk < s; sk++)
                    for (int si = 0; si < s; si++) a += v[d.fv(i * s + si, j * s, k * s + sk)];
                cv[d.cv(i, j, k)] = a / sf;
            }
    for (int k = 0; k <= d.Dc; k++)
        for (int j = 0; j < d.Hc; j++)
            for (int i = 0; i < d.Wc; i++) {
                float a = 0.0f;
                for (int sj = 0; sj < s; sj++)
                    for (int si = 0; si < s; si++) a += w[d.fw(i * s + si, j * s + sj, k * s)];
                cw[d.cw(i, j, k)] = a / sf;
            }
}

// Face-fraction Poisson on the coarse grid, the system of the reference's matrix build
// (ZQ_InitMatOpenPoisson3D, ZQ_SmokeSimulationUtils3D.cpp:476-592): for a fluid coarse cell,
//   sum over in-domain sides of r_f * p_nb  -  c * p  =  div,
//   c = (out-of-domain sides) + sum of r_f over the in-domain sides,
// div the unscaled coarse divergence (u[i+1]-u[i] + v[j+1]-v[j] + w[k+1]-w[k], left to right).
// The reference's COARSE_*_POISSON_SOR solvers copy the coarse faces, project the copy with
// SolveOpenPoissonRedBlackwithFaceRatio3D_MAC (maxIter iterations) and keep delta = projected -
// original (:2611-2642; COARSE_CLOSED_POISSON_SOR takes the same call, :3732-3738).  That solver
// lives in the un-vendored ZQlibCuda: parity unpinned.  Restated here as the full-grid solver of
// this library restricted to face weights: red-black SOR from p = 0, colour (i+j+k)&1, colour 0
// first; s = r(x-) p(x-) + r(x+) p(x+) + r(y-) p(y-) + r(y+) p(y+) + r(z-) p(z-) + r(z+) p(z+)
// (ghosts and solid cells hold 0), p = (1-omega) p + omega (s - div) * (1/c); cells with c == 0
// stay 0.  Weights of 1 everywhere reduce it, bit for bit, to orc_sor on the same grid.  The
// projection then moves a face with fraction > 0 by g = p_r - p_l (p = 0 outside an open domain
// face; closed domain faces and faces with fraction 0 stay), as ZQ_AdjustOpenVelocity3D does
// (:2794-2865), and delta = (u - g) - u.  open_mask selects the open domain faces (the reference:
// all six).  p: (Wc+2)(Hc+2)(Dc+2), zero on entry.  vol (nullable): closed domains, the
// divergence's mean per unit fluid volume is removed first (the *_CLOSED_FLUX_SOR solvers,
// :2465-2520, 2734-2792, whose ZQlibCuda solves are absent: the weighting is this restatement's).
// flux != 0 takes the flux form's diagonal (orc_coarse_inv_diag).
void orc_coarse_project(int Wc, int Hc, int Dc, const float* cu, const float* cv, const float* cw,
                        const uint8_t* cocc, const float* rU, const float* rV, const float* rW, int open_mask,
                        float omega, int iters, float* p, float* du, float* dv, float* dw, const float* vol,
                        int flux) {
    const CD d{Wc, Hc, Dc, 1, Wc, Hc, Dc};
    const size_t N = (size_t)Wc * Hc * Dc;
    std::vector<float> div(N, 0.0f), inv(N, 0.0f);
    for (int k = 0; k < Dc; k++)
        for (int j = 0; j < Hc; j++)
            for (int i = 0; i < Wc; i++) {
                const size_t o = d.cc(i, j, k);
                if (cocc[o]) continue;
                div[o] = cu[d.cu(i + 1, j, k)] - cu[d.cu(i, j, k)] + cv[d.cv(i, j + 1, k)] - cv[d.cv(i, j, k)] +
                         cw[d.cw(i, j, k + 1)] - cw[d.cw(i, j, k)];
                inv[o] = orc_coarse_inv_diag(Wc, Hc, Dc, i, j, k, rU, rV, rW, open_mask, flux);
            }
    if (vol) {
        // closed domains: div -= mean * vol, mean = sum(div) / sum(vol) over non-solid cells, FP64
        double sd = 0.0, sv = 0.0;
        for (size_t o = 0; o < N; o++)
            if (!cocc[o]) {
                sd += (double)div[o];
                sv += (double)vol[o];
            }
        const double mean = sv > 0.0 ? sd / sv : 0.0;
        for (size_t o = 0; o < N; o++)
            if (!cocc[o]) div[o] = (float)((double)div[o] - mean * (double)vol[o]);
    }
    const size_t sy = (size_t)(Wc + 2), sz = (size_t)(Wc + 2) * (Hc + 2);
    const float omw = 1.0f - omega;
    for (int it = 0; it < iters; it++)
        for (int colour = 0; colour < 2; colour++)
            for (int k = 0; k < Dc; k++)
                for (int j = 0; j < Hc; j++)
                    for (int i = (colour + j + k) & 1; i < Wc; i += 2) {
                        const size_t o = d.cc(i, j, k);
                        if (inv[o] == 0.0f) continue;
                        const size_t q = d.cp(i, j, k);
                        float s = rU[d.cu(i, j, k)] * p[q - 1] + rU[d.cu(i + 1, j, k)] * p[q + 1] +
                                  rV[d.cv(i, j, k)] * p[q - sy] + rV[d.cv(i, j + 1, k)] * p[q + sy] +
                                  rW[d.cw(i, j, k)] * p[q - sz] + rW[d.cw(i, j, k + 1)] * p[q + sz];
                        const float pg = (s - div[o]) * inv[o];
                        p[q] = omw * p[q] + omega * pg;
                    }
    auto delta = [](float r, bool ok, float u, float pr, float pl) {
        if (!(r > 0.0f) || !ok) return 0.0f;
        const float g = pr - pl;
        return (u - g) - u;
    };
    for (int k = 0; k < Dc; k++)
        for (int j = 0; j < Hc; j++)
            for (int i = 0; i <= Wc; i++) {
                const size_t f = d.cu(i, j, k);
                const bool ok = (i > 0 && i < Wc) || (i == 0 && (open_mask & ORC_OPEN_XLO)) ||
                                (i == Wc && (open_mask & ORC_OPEN_XHI));
                du[f] = delta(rU[f], ok, cu[f], p[d.cp(i, j, k)], p[d.cp(i - 1, j, k)]);
            }
    for (int k = 0; k < Dc; k++)
        for (int j = 0; j <= Hc; j++)
            for (int i = 0; i < Wc; i++) {
                const size_t f = d.cv(i, j, k);
                const bool ok = (j > 0 && j < Hc) || (j == 0 && (open_mask & ORC_OPEN_YLO)) ||
                                (j == Hc && (open_mask & ORC_OPEN_YHI));
                dv[f] = delta(rV[f], ok, cv[f], p[d.cp(i, j, k)], p[d.cp(i, j - 1, k)]);
            }
    for (int k = 0; k <= Dc; k++)
        for (int j = 0; j < Hc; j++)
            for (int i = 0; i < Wc; i++) {
                const size_t f = d.cw(i, j, k);
                const bool ok = (k > 0 && k < Dc) || (k == 0 && (open_mask & ORC_OPEN_ZLO)) ||
                                (k == Dc && (open_mask & ORC_OPEN_ZHI));
                dw[f] = delta(rW[f], ok, cw[f], p[d.cp(i, j, k)], p[d.cp(i, j, k - 1)]);
            }
}

// Coarse correction added to the fine faces that lie on coarse faces (ZQ_MapCoarse2Fine3D,
// ZQ_SmokeSimulationUtils3D.cpp:3172-3296): an interior fine face gets it when both fine cells
// beside it are fluid, a domain face when the inside cell is.  Two choices about the reference's
// own slips: its x = W face tests the cell at x = W, one past the row (:3209; read here: the inside
// cell x = W-1, as the y and z faces do); its y = H faces take the delta of coarse v row Hc-1
// instead of Hc (:3251), which is kept.
void orc_coarse_to_fine(int W, int H, int D, int s, const float* du, const float* dv, const float* dw,
                        const uint8_t* occ, float* u, float* v, float* w) {
    const CD d = make(W, H, D, s);
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int ic = 0; ic <= d.Wc; ic++) {
                const int x = ic * s;
                bool ok;
                if (ic == 0) ok = !occ[d.c(0, j, k)];
                else if (ic == d.Wc) ok = !occ[d.c(W - 1, j, k)];
                else ok = !occ[d.c(x, j, k)] && !occ[d.c(x - 1, j, k)];
                if (ok) u[d.fu(x, j, k)] += du[d.cu(ic, j / s, k / s)];
            }
    for (int k = 0; k < D; k++)
        for (int jc = 0; jc <= d.Hc; jc++)
            for (int i = 0; i < W; i++) {
                const int y = jc * s;
                bool ok;
                int row = jc;
                if (jc == 0) ok = !occ[d.c(i, 0, k)];
                else if (jc == d.Hc) {
                    ok = !occ[d.c(i, H - 1, k)];
                    row = d.Hc - 1;
                } else ok = !occ[d.c(i, y, k)] && !occ[d.c(i, y - 1, k)];
                if (ok) v[d.fv(i, y, k)] += dv[d.cv(i / s, row, k / s)];
            }
    for (int kc = 0; kc <= d.Dc; kc++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                const int z = kc * s;
                bool ok;
                if (kc == 0) ok = !occ[d.c(i, j, 0)];
                else if (kc == d.Dc) ok = !occ[d.c(i, j, D - 1)];
                else ok = !occ[d.c(i, j, z)] && !occ[d.c(i, j, z - 1)];
                if (ok) w[d.fw(i, j, z)] += dw[d.cw(i / s, j / s, kc)];
            }
}

// Sub-projection of the regular blocks (ZQ_SubProjection3D, ZQ_SmokeSimulationUtils3D.cpp:3318-3460):
// a block's faces in the reference's order -- u (sk, sj, si = 0..s), v (sk, sj = 0..s, si), w
// (sk = 0..s, sj, si) -- form x (F = 3 s^2 (s+1) values); its interior faces, in the order u
// (sk, sj, si = 1..s-1), v (sk, sj = 1..s-1, si), w (sk = 1..s-1, sj, si), become y = M x
// (R = 3 s^2 (s-1) rows), each a dot product summed from column 0 up, one rounding per operation.
// The reference hands this product to ZQ_CUDA_MatrixMul in the un-vendored ZQlibCuda (:13-15), so
// its summation order is unpinned.  Blocks with a solid cell are left alone.
void orc_sub_project_regular(int W, int H, int D, int s, const float* M, const uint8_t* regular, float* u, float* v,
                             float* w) {
    const CD d = make(W, H, D, s);
    const int F = 3 * s * s * (s + 1), R = 3 * s * s * (s - 1);
    std::vector<float> x(F), y(R);
    for (int bk = 0; bk < d.Dc; bk++)
        for (int bj = 0; bj < d.Hc; bj++)
            for (int bi = 0; bi < d.Wc; bi++) {
                if (!regular[d.cc(bi, bj, bk)]) continue;
                const int x0 = bi * s, y0 = bj * s, z0 = bk * s;
                int c = 0;
                for (int sk = 0; sk < s; sk++)
                    for (int sj = 0; sj < s; sj++)
                        for (int si = 0; si <= s; si++) x[c++] = u[d.fu(x0 + si, y0 + sj, z0 + sk)];
                for (int sk = 0; sk < s; sk++)
                    for (int sj = 0; sj <= s; sj++)
                        for (int si = 0; si < s; si++) x[c++] = v[d.fv(x0 + si, y0 + sj, z0 + sk)];
                for (int sk = 0; sk <= s; sk++)
                    for (int sj = 0; sj < s; sj++)
                        for (int si = 0; si < s; si++) x[c++] = w[d.fw(x0 + si, y0 + sj, z0 + sk)];
                for (int r = 0; r < R; r++) {
                    float a = 0.0f;
                    for (int cc = 0; cc < F; cc++) a = a + M[(size_t)r * F + cc] * x[cc];
                    y[r] = a;
                }
                int r = 0;
                for (int sk = 0; sk < s; sk++)
                    for (int sj = 0; sj < s; sj++)
                        for (int si = 1; si < s; si++) u[d.fu(x0 + si, y0 + sj, z0 + sk)] = y[r++];
                for (int sk = 0; sk < s; sk++)
                    for (int sj = 1; sj < s; sj++)
                        for (int si = 0; si < s; si++) v[d.fv(x0 + si, y0 + sj, z0 + sk)] = y[r++];
                for (int sk = 1; sk < s; sk++)
                    for (int sj = 0; sj < s; sj++)
                        for (int si = 0; si < s; si++) w[d.fw(x0 + si, y0 + sj, z0 + sk)] = y[r++];
            }
}

}  // extern "C"
