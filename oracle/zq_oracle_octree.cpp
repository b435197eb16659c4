// zq_oracle_octree.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's octree projection with the red-black solver
// (OPEN_OCTREE_POISSON_SOR / CLOSED_OCTREE_POISSON_SOR, ZQ_SmokeOctreeGrid3D.cpp): the distance
// field to the solid cells, the four leaf levels (fine cells and 2^3, 4^3, 8^3 blocks), the
// per-leaf neighbour lists in the order the reference builds them, the red-black Gauss-Seidel
// sweeps over the levels and the face correction.  The sub-projection of the leaf blocks reuses
// orc_sub_project_regular (zq_oracle_coarse.cpp) from zq_oracle.py.
// Only tests/ may load this; the product library never links or calls it.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "zq_oracle.h"

namespace {

struct OG {  // fine grid W, H, D (multiples of 8); level L has (W >> L, H >> L, D >> L) cells
    int W, H, D;
    int w(int L) const { return W >> L; }
    int h(int L) const { return H >> L; }
    int d(int L) const { return D >> L; }
    size_t at(int L, int x, int y, int z) const { return ((size_t)z * h(L) + y) * w(L) + x; }
    bool in(int L, int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < w(L) && y < h(L) && z < d(L); }
};

struct Leaves {
    const uint8_t* l[4];
    bool at(const OG& g, int L, int x, int y, int z) const { return l[L][g.at(L, x, y, z)] != 0; }
};

struct Nb {
    int lvl, x, y, z;
};

// Face weight between a level-L cell and a level-M neighbour: the area of the smaller cell's face
// over the distance between centres, 4^min(L,M) / (len_L + len_M) with len = 0.5 * 2^L (the weight
// tables of CPU_PoissonSolver3D_OpenOctreePoisson.cpp:408,452,496,540).
float weight(int L, int M) {
    const int m = L < M ? L : M;
    const float area = (float)(1 << (2 * m));
    const float lenL = 0.5f * (float)(1 << L), lenM = 0.5f * (float)(1 << M);
    return area / (lenL + lenM);
}

// The neighbours of a level-L cell (x, y, z) across the face on axis `ax` (0 x, 1 y, 2 z) and side
// `sd` (-1, +1), appended in the order of _selectInfos_for_CUDAOpenPoissonRedBlack3
// (ZQ_SmokeOctreeGrid3D.cpp:6330-9305; level 0 :6368-6941, 1 :6943-7626, 2 :7628-8405, 3 :8407-9305):
// a domain face gives one level-3 entry with out-of-range coordinates (an open boundary: it enters
// the diagonal only); otherwise the same-level cell across the face if it is a leaf, else the first
// coarser leaf whose block the face bounds, else (L > 0) the leaves behind the face found by
// descending one level at a time over the 2 x 2 sub-squares, slower axis outer (x faces: z then y;
// y faces: z then x; z faces: y then x), where a non-leaf square is descended into and a fine cell
// that is no leaf (a solid) gives nothing.
void face_neighbours(const OG& g, const Leaves& lf, int L, int x, int y, int z, int ax, int sd, std::vector<Nb>& out) {
    int c[3] = {x, y, z};
    const int dimL[3] = {g.w(L), g.h(L), g.d(L)};
    const int sh = 3 - L;
    if ((sd < 0 && c[ax] == 0) || (sd > 0 && c[ax] == dimL[ax] - 1)) {
        int b[3] = {c[0] >> sh, c[1] >> sh, c[2] >> sh};
        b[ax] = sd < 0 ? (c[ax] >> sh) - 1 : (c[ax] + 1) >> sh;
        out.push_back(Nb{3, b[0], b[1], b[2]});
        return;
    }
    int n[3] = {c[0], c[1], c[2]};
    n[ax] += sd;
    if (lf.at(g, L, n[0], n[1], n[2])) {
        out.push_back(Nb{L, n[0], n[1], n[2]});
        return;
    }
    for (int M = L + 1; M <= 3; M++) {
        const int k = M - L;
        const int face = sd < 0 ? c[ax] : c[ax] + 1;  // the face's coordinate at level L
        if (face % (1 << k) != 0) continue;
        if (lf.at(g, M, n[0] >> k, n[1] >> k, n[2] >> k)) {
            out.push_back(Nb{M, n[0] >> k, n[1] >> k, n[2] >> k});
            return;
        }
    }
    if (L == 0) return;
    // smaller leaves: the layer of level m adjacent to the face is at 2 n + 1 (sd < 0) or 2 n (sd > 0)
    // of the level above; the two face axes in loop order (outer, inner)
    const int outer = ax == 2 ? 1 : 2, inner = ax == 0 ? 1 : 0;
    struct Rec {
        const OG& g;
        const Leaves& lf;
        int ax, sd, outer, inner;
        std::vector<Nb>& out;
        void go(int m, int q[3]) {
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) {
                    int r[3];
                    r[ax] = sd < 0 ? 2 * q[ax] + 1 : 2 * q[ax];
                    r[outer] = 2 * q[outer] + a;
                    r[inner] = 2 * q[inner] + b;
                    if (lf.at(g, m, r[0], r[1], r[2])) out.push_back(Nb{m, r[0], r[1], r[2]});
                    else if (m > 0) go(m - 1, r);
                }
        }
    } rec{g, lf, ax, sd, outer, inner, out};
    rec.go(L - 1, n);
}

void cell_neighbours(const OG& g, const Leaves& lf, int L, int x, int y, int z, std::vector<Nb>& out) {
    for (int ax = 0; ax < 3; ax++)
        for (int sd = -1; sd <= 1; sd += 2) face_neighbours(g, lf, L, x, y, z, ax, sd, out);
}

struct List {  // one (level, colour) list: cells and their neighbour ranges
    std::vector<int> cells;  // x, y, z per cell
    std::vector<int> begin;  // neighbour range [begin[i], begin[i + 1])
    std::vector<Nb> nb;
};

}  // namespace

extern "C" {

// Distance of every cell to the nearest solid cell by breadth-first filling from the solid cells
// over the six face neighbours, the seeds queued in x-fastest scan order (_buildOctreeInfo and
// _seedFilling, ZQ_SmokeOctreeGrid3D.cpp:5328-5375, 5034-5167); with no solid cell every distance
// is W + H + D.
void orc_oct_distance(int W, int H, int D, const uint8_t* occ, int32_t* dist) {
    const size_t N = (size_t)W * H * D;
    bool seed = false;
    for (size_t i = 0; i < N; i++) {
        dist[i] = occ[i] ? 0 : -1;
        seed |= occ[i] != 0;
    }
    if (!seed) {
        for (size_t i = 0; i < N; i++) dist[i] = W + H + D;
        return;
    }
    std::vector<int> q;
    q.reserve(N);
    for (size_t i = 0; i < N; i++)
        if (dist[i] == 0) q.push_back((int)i);
    for (size_t h = 0; h < q.size(); h++) {
        const int c = q[h], i = c % W, j = (c / W) % H, k = c / (W * H);
        const int nd = dist[c] + 1;
        const int cand[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        for (auto& n : cand) {
            if (n[0] < 0 || n[1] < 0 || n[2] < 0 || n[0] >= W || n[1] >= H || n[2] >= D) continue;
            const int o = (n[2] * H + n[1]) * W + n[0];
            if (dist[o] < 0) {
                dist[o] = nd;
                q.push_back(o);
            }
        }
    }
}

// Leaf flags of the four levels (_buildOctreeInfo, ZQ_SmokeOctreeGrid3D.cpp:5377-5534): an 8^3 block
// is a leaf when every distance in it exceeds t3; a 4^3 block outside level-3 leaves when every
// distance exceeds t2; a 2^3 block outside coarser leaves when every distance exceeds t1; a fine
// cell when it lies in no coarser leaf and is not solid.
void orc_oct_leaves(int W, int H, int D, const uint8_t* occ, const int32_t* dist, int t1, int t2, int t3, uint8_t* l0,
                    uint8_t* l1, uint8_t* l2, uint8_t* l3) {
    const OG g{W, H, D};
    const int th[4] = {0, t1, t2, t3};
    uint8_t* l[4] = {l0, l1, l2, l3};
    for (int L = 3; L >= 1; L--) {
        const int s = 1 << L;
        for (int z = 0; z < g.d(L); z++)
            for (int y = 0; y < g.h(L); y++)
                for (int x = 0; x < g.w(L); x++) {
                    bool inside = false;
                    for (int M = L + 1; M <= 3; M++)
                        inside |= l[M][g.at(M, x >> (M - L), y >> (M - L), z >> (M - L))] != 0;
                    bool flag = !inside;
                    for (int k = 0; k < s && flag; k++)
                        for (int j = 0; j < s && flag; j++)
                            for (int i = 0; i < s && flag; i++)
                                if (dist[g.at(0, x * s + i, y * s + j, z * s + k)] <= th[L]) flag = false;
                    l[L][g.at(L, x, y, z)] = flag ? 1 : 0;
                }
    }
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const bool coarse = l1[g.at(1, x >> 1, y >> 1, z >> 1)] || l2[g.at(2, x >> 2, y >> 2, z >> 2)] ||
                                    l3[g.at(3, x >> 3, y >> 3, z >> 3)];
                l0[g.at(0, x, y, z)] = (!coarse && !occ[g.at(0, x, y, z)]) ? 1 : 0;
            }
}

// Neighbour entries (level, x, y, z) of the level-L cell (x, y, z) in the reference's order; out
// holds up to 6 * 4^L entries of 4 ints; returns the count.
int orc_oct_neighbors(int W, int H, int D, const uint8_t* l0, const uint8_t* l1, const uint8_t* l2, const uint8_t* l3,
                      int L, int x, int y, int z, int32_t* out) {
    const OG g{W, H, D};
    const Leaves lf{{l0, l1, l2, l3}};
    std::vector<Nb> nb;
    cell_neighbours(g, lf, L, x, y, z, nb);
    for (size_t i = 0; i < nb.size(); i++) {
        out[4 * i] = nb[i].lvl;
        out[4 * i + 1] = nb[i].x;
        out[4 * i + 2] = nb[i].y;
        out[4 * i + 3] = nb[i].z;
    }
    return (int)nb.size();
}

// Divergence of the fine cells, u(i+1) - u(i) + v(j+1) - v(j) + w(k+1) - w(k) left to right, and of
// each coarser level as the sum of its eight children in x-fastest order
// (CPU_SolveOpenOctreePoissonRedBlack3_MAC and Calculate_Divergence_Octree_from_previous_level_Kernel,
// CPU_PoissonSolver3D_OpenOctreePoisson.cpp:596-613, 375-400).
void orc_oct_divergence(int W, int H, int D, const float* u, const float* v, const float* w, float* d0, float* d1,
                        float* d2, float* d3) {
    const OG g{W, H, D};
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                const size_t cu = ((size_t)k * H + j) * (W + 1) + i, cv = ((size_t)k * (H + 1) + j) * W + i;
                const size_t cw = ((size_t)k * H + j) * W + i;
                d0[g.at(0, i, j, k)] = u[cu + 1] - u[cu] + v[cv + W] - v[cv] + w[cw + (size_t)W * H] - w[cw];
            }
    float* d[4] = {d0, d1, d2, d3};
    for (int L = 1; L <= 3; L++)
        for (int z = 0; z < g.d(L); z++)
            for (int y = 0; y < g.h(L); y++)
                for (int x = 0; x < g.w(L); x++) {
                    const float* p = d[L - 1];
                    const int X = 2 * x, Y = 2 * y, Z = 2 * z, M = L - 1;
                    d[L][g.at(L, x, y, z)] = p[g.at(M, X, Y, Z)] + p[g.at(M, X + 1, Y, Z)] + p[g.at(M, X, Y + 1, Z)] +
                                             p[g.at(M, X + 1, Y + 1, Z)] + p[g.at(M, X, Y, Z + 1)] +
                                             p[g.at(M, X + 1, Y, Z + 1)] + p[g.at(M, X, Y + 1, Z + 1)] +
                                             p[g.at(M, X + 1, Y + 1, Z + 1)];
                }
}

// Closed domains: the mean divergence over the non-solid fine cells, summed in float in scan order
// and divided by their count (_solveGlobalClosedPoissonSOR, ZQ_SmokeOctreeGrid3D.cpp:2935-2964).
float orc_oct_closed_mean(int W, int H, int D, const uint8_t* occ, const float* u, const float* v, const float* w) {
    float s = 0.0f;
    int count = 0;
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                if (occ[((size_t)k * H + j) * W + i]) continue;
                const size_t cu = ((size_t)k * H + j) * (W + 1) + i, cv = ((size_t)k * (H + 1) + j) * W + i;
                const size_t cw = ((size_t)k * H + j) * W + i;
                count++;
                s += u[cu + 1] - u[cu] + v[cv + W] - v[cv] + w[cw + (size_t)W * H] - w[cw];
            }
    return count ? s / (float)count : 0.0f;
}

// `iters` passes of red-black Gauss-Seidel over the levels in the order level 0 red, level 0 black,
// level 1 red, ..., level 3 black, from zero pressures (CPU_SolveOpenOctreePoissonRedBlack3_MAC,
// CPU_PoissonSolver3D_OpenOctreePoisson.cpp:615-647): each cell sums, over its neighbour list in
// order, coeff += weight and sigma += weight * p (an out-of-range entry, the open boundary, adds
// its weight to coeff only), then p = (sigma - rhs) / coeff (:404-425).  rhs is the level's
// divergence; with `closed`, div - mean * 8^L (the mean divergence per fine cell times the cell's
// volume).  The red cells have even x + y + z at their level.  Lists as _selectInfos_for_CUDA...3
// builds them, including its level-3 test of the block at z = 0 of each column
// (ZQ_SmokeOctreeGrid3D.cpp:8416): a level-3 block is solved when the block (x, y, 0) is a leaf.
void orc_oct_solve(int W, int H, int D, const uint8_t* l0, const uint8_t* l1, const uint8_t* l2, const uint8_t* l3,
                   const float* d0, const float* d1, const float* d2, const float* d3, int iters, int closed,
                   float mean, float* p0, float* p1, float* p2, float* p3) {
    const OG g{W, H, D};
    const Leaves lf{{l0, l1, l2, l3}};
    const float* dv[4] = {d0, d1, d2, d3};
    float* p[4] = {p0, p1, p2, p3};
    List lists[4][2];
    for (int L = 0; L <= 3; L++) {
        std::memset(p[L], 0, sizeof(float) * g.w(L) * g.h(L) * g.d(L));
        for (int z = 0; z < g.d(L); z++)
            for (int y = 0; y < g.h(L); y++)
                for (int x = 0; x < g.w(L); x++) {
                    const bool take = L == 3 ? lf.at(g, 3, x, y, 0) : lf.at(g, L, x, y, z);
                    if (!take) continue;
                    List& li = lists[L][(x + y + z) % 2];
                    li.cells.insert(li.cells.end(), {x, y, z});
                    li.begin.push_back((int)li.nb.size());
                    cell_neighbours(g, lf, L, x, y, z, li.nb);
                }
        for (auto& li : lists[L]) li.begin.push_back((int)li.nb.size());
    }
    float wt[4][4];
    for (int L = 0; L < 4; L++)
        for (int M = 0; M < 4; M++) wt[L][M] = weight(L, M);
    for (int it = 0; it < iters; it++)
        for (int L = 0; L <= 3; L++) {
            const float vol = (float)(1 << (3 * L));
            for (int col = 0; col < 2; col++) {
                const List& li = lists[L][col];
                const size_t n = li.cells.size() / 3;
                for (size_t c = 0; c < n; c++) {
                    const int x = li.cells[3 * c], y = li.cells[3 * c + 1], z = li.cells[3 * c + 2];
                    float coeff = 0.0f, sigma = 0.0f;
                    for (int e = li.begin[c]; e < li.begin[c + 1]; e++) {
                        const Nb& b = li.nb[e];
                        const float wgt = wt[L][b.lvl];
                        coeff += wgt;
                        if (g.in(b.lvl, b.x, b.y, b.z)) sigma += wgt * p[b.lvl][g.at(b.lvl, b.x, b.y, b.z)];
                    }
                    const float dd = dv[L][g.at(L, x, y, z)];
                    sigma -= closed ? dd - mean * vol : dd;
                    p[L][g.at(L, x, y, z)] = sigma / coeff;
                }
            }
        }
}

// Face correction (Adjust_MAC_{u,v,w}_OpenOctreePoisson_Kernel, CPU_PoissonSolver3D_OpenOctreePoisson.cpp:6-373):
// a face between the leaves on its two sides -- the fine cell beside it if that is a leaf, else the
// coarser leaf whose block the face bounds -- takes face -= (p2 - p1) / (len1 + len2); a domain face
// beside a leaf takes face -= (p - 0) / (4 + len) at the low side and (0 - p) / (4 + len) at the high
// side; faces inside a leaf or beside a solid cell are left alone.  (The reference's CPU file leaves
// the whole u and v loops at the first such face, :96,115; the kernels it restates skip just that
// face, which is what is done here.)
void orc_oct_adjust(int W, int H, int D, const uint8_t* l0, const uint8_t* l1, const uint8_t* l2, const uint8_t* l3,
                    const float* p0, const float* p1, const float* p2, const float* p3, float* u, float* v, float* w) {
    const OG g{W, H, D};
    const Leaves lf{{l0, l1, l2, l3}};
    const float* p[4] = {p0, p1, p2, p3};
    // the leaf holding fine cell c beside a face at coordinate f (along the face's axis): the cell
    // itself at level 0, else a coarser leaf whose block the face bounds
    auto side = [&](int f, const int c[3], float& len, float& pv) {
        for (int M = 0; M <= 3; M++) {
            if (M > 0 && f % (1 << M) != 0) continue;
            if (lf.at(g, M, c[0] >> M, c[1] >> M, c[2] >> M)) {
                len = 0.5f * (float)(1 << M);
                pv = p[M][g.at(M, c[0] >> M, c[1] >> M, c[2] >> M)];
                return true;
            }
        }
        return false;
    };
    const int dims[3] = {W, H, D};
    for (int ax = 0; ax < 3; ax++) {
        const int ex[3] = {W + (ax == 0), H + (ax == 1), D + (ax == 2)};
        float* f = ax == 0 ? u : ax == 1 ? v : w;
        for (int k = 0; k < ex[2]; k++)
            for (int j = 0; j < ex[1]; j++)
                for (int i = 0; i < ex[0]; i++) {
                    const int fc[3] = {i, j, k};
                    const int x = fc[ax];  // the face's coordinate along its axis
                    float* face = &f[((size_t)k * ex[1] + j) * ex[0] + i];
                    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
                    lo[ax] -= 1;
                    float len1, len2, pa, pb;
                    if (x == 0) {
                        if (!side(0, hi, len2, pb)) continue;
                        *face -= (pb - 0.0f) / (4.0f + len2);
                    } else if (x == dims[ax]) {
                        if (!side(x, lo, len1, pa)) continue;
                        *face -= (0.0f - pa) / (4.0f + len1);
                    } else {
                        if (!side(x, lo, len1, pa)) continue;
                        if (!side(x, hi, len2, pb)) continue;
                        *face -= (pb - pa) / (len1 + len2);
                    }
                }
    }
}

// Region labels of _find_special_label (ZQ_SmokeOctreeGrid3D.cpp:5169-5320); lab holds -1 on the
// masked cells and 0 elsewhere.  A pass seeds the first unlabelled cell of the lowest plane that has
// one and, in every plane above it, the first unlabelled cell of row 0 only (the plane loop is not
// left after the first seed, the row loop is); the seeds take the pass's label and the fill starts
// from the last one.  The fill queues the z neighbours without labelling them and labels the y and x
// neighbours as it queues them (order z+1, z-1, y+1, y-1, x+1, x-1), so a cell reached only along z
// stays unlabelled and seeds a later pass.  Two such cells above each other queue each other without
// end; the reference's queue (W*H*D entries) then overflows: returns -1 there.
static int reference_region_labels(int W, int H, int D, std::vector<int32_t>& lab) {
    const size_t N = (size_t)W * H * D, plane = (size_t)W * H;
    int regions = 0;
    std::vector<size_t> queue;
    for (;;) {
        size_t seed = N;
        bool seeded = false;
        for (int z = 0; z < D; z++) {
            const int rows = seeded ? 1 : H;
            bool here = false;
            for (int y = 0; y < rows && !here; y++)
                for (int x = 0; x < W; x++) {
                    const size_t c = (size_t)z * plane + (size_t)y * W + x;
                    if (lab[c] == -1) {
                        lab[c] = regions + 1;
                        seed = c;
                        here = true;
                        break;
                    }
                }
            seeded = seeded || here;
        }
        if (!seeded) return regions;
        const int r = ++regions;
        queue.assign(1, seed);
        for (size_t h = 0; h < queue.size(); h++) {
            const size_t c = queue[h];
            const int x = (int)(c % W), y = (int)((c / W) % H), z = (int)(c / plane);
            auto visit = [&](size_t n, bool label) {
                if (lab[n] != -1) return;
                if (label) lab[n] = r;
                queue.push_back(n);
            };
            if (z + 1 < D) visit(c + plane, false);
            if (z >= 1) visit(c - plane, false);
            if (y + 1 < H) visit(c + W, true);
            if (y >= 1) visit(c - W, true);
            if (x + 1 < W) visit(c + 1, true);
            if (x >= 1) visit(c - 1, true);
            if (queue.size() > N) return -1;
        }
    }
}

int orc_oct_special_regions(int W, int H, int D, const uint8_t* l0, const uint8_t* l1, int iters, float* u, float* v,
                            float* w, int32_t* labels) {
    const size_t N = (size_t)W * H * D;
    const OG g{W, H, D};
    auto cell = [&](int x, int y, int z) { return ((size_t)z * H + y) * W + x; };
    std::vector<int32_t> lab(N, 0);
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++)
                if (l0[g.at(0, x, y, z)] || l1[g.at(1, x >> 1, y >> 1, z >> 1)]) lab[cell(x, y, z)] = -1;
    const int regions = reference_region_labels(W, H, D, lab);
    if (regions < 0) return -1;
    const size_t Wu = (size_t)W + 1, Hv = (size_t)H + 1;
    for (int r = 1; r <= regions; r++) {
        int x0 = W, y0 = H, z0 = D, x1 = -1, y1 = -1, z1 = -1, count = 0;
        float mean = 0.0f;
        for (int z = 0; z < D; z++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    if (lab[cell(x, y, z)] != r) continue;
                    count++;
                    mean += u[((size_t)z * H + y) * Wu + x + 1] - u[((size_t)z * H + y) * Wu + x] +
                            v[((size_t)z * Hv + y + 1) * W + x] - v[((size_t)z * Hv + y) * W + x] +
                            w[((size_t)(z + 1) * H + y) * W + x] - w[((size_t)z * H + y) * W + x];
                    x0 = std::min(x0, x), y0 = std::min(y0, y), z0 = std::min(z0, z);
                    x1 = std::max(x1, x), y1 = std::max(y1, y), z1 = std::max(z1, z);
                }
        mean /= count;
        const int sx = x1 - x0 + 1, sy = y1 - y0 + 1, sz = z1 - z0 + 1;
        std::vector<uint8_t> bo((size_t)sx * sy * sz);
        std::vector<float> bu((size_t)(sx + 1) * sy * sz), bv((size_t)sx * (sy + 1) * sz), bw((size_t)sx * sy * (sz + 1));
        for (int k = 0; k < sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++) bo[((size_t)k * sy + j) * sx + i] = lab[cell(x0 + i, y0 + j, z0 + k)] != r;
        for (int k = 0; k < sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i <= sx; i++)
                    bu[((size_t)k * sy + j) * (sx + 1) + i] = u[((size_t)(z0 + k) * H + y0 + j) * Wu + x0 + i];
        for (int k = 0; k < sz; k++)
            for (int j = 0; j <= sy; j++)
                for (int i = 0; i < sx; i++)
                    bv[((size_t)k * (sy + 1) + j) * sx + i] = v[((size_t)(z0 + k) * Hv + y0 + j) * W + x0 + i];
        for (int k = 0; k <= sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++)
                    bw[((size_t)k * sy + j) * sx + i] = w[((size_t)(z0 + k) * H + y0 + j) * W + x0 + i];
        orc_project(sx, sy, sz, bu.data(), bv.data(), bw.data(), bo.data(), 0, mean, 1.0f, std::min(iters, count),
                    nullptr);
        for (int k = 0; k < sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++)  // the reference's copy-back stops short of the high x face
                    u[((size_t)(z0 + k) * H + y0 + j) * Wu + x0 + i] = bu[((size_t)k * sy + j) * (sx + 1) + i];
        for (int k = 0; k < sz; k++)
            for (int j = 0; j <= sy; j++)
                for (int i = 0; i < sx; i++)
                    v[((size_t)(z0 + k) * Hv + y0 + j) * W + x0 + i] = bv[((size_t)k * (sy + 1) + j) * sx + i];
        for (int k = 0; k <= sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++)
                    w[((size_t)(z0 + k) * H + y0 + j) * W + x0 + i] = bw[((size_t)k * sy + j) * sx + i];
    }
    if (labels) std::memcpy(labels, lab.data(), N * sizeof(int32_t));
    return regions;
}

}  // extern "C"
