"""The oracle's octree projection (OPEN_OCTREE_POISSON_SOR / CLOSED_OCTREE_POISSON_SOR), pinned by
checks that do not restate it:

- distance field: scipy's taxicab distance transform;
- leaf flags: a numpy block-minimum closed form, and every fine cell covered exactly once;
- neighbour lists: brute-force face adjacency through a per-cell owner map (the set of neighbours,
  their shared face counts, the domain faces) and their order (the first fine cell of each
  neighbour on the face, ranked by interleaving the face's two coordinates, slower axis high);
- the converged red-black solve: a direct sparse solve (float64) of the system built from the
  brute-force adjacency with weights (shared fine faces) / (len_L + len_M);
- the converged projection: zero divergence in every fine cell of level-0, 2 and 3 leaves, and (w
  written back per row) of level-1 leaves; the reference's two slips shown on their own.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import zq_oracle as O
from util import faces


def scene(W, H, D, seed, spheres=1, sprinkle=0.0, shell=False, centres=None):
    rng = np.random.default_rng(seed)
    occ = (rng.random((D, H, W)) < sprinkle).astype(np.uint8)
    zz, yy, xx = np.mgrid[:D, :H, :W]
    for n in range(spheres if centres is None else len(centres)):
        c = rng.random(3) * [W, H, D] if centres is None else centres[n]
        r = 0.12 * min(W, H, D)
        occ[((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2) < r * r] = 1
    if shell:
        occ[0] = occ[-1] = 1
        occ[:, 0] = occ[:, -1] = 1
        occ[:, :, 0] = occ[:, :, -1] = 1
    return occ


def owner_levels(leaves, occ):
    """Per fine cell: the level of the leaf holding it (0..3), 4 for a solid cell."""
    D, H, W = occ.shape
    lev = np.full((D, H, W), 4, np.int8)
    for L in (3, 2, 1):
        up = np.repeat(np.repeat(np.repeat(leaves[L], 1 << L, 0), 1 << L, 1), 1 << L, 2)
        lev[(up == 1) & (lev == 4)] = L
    lev[leaves[0] == 1] = 0
    return lev


def test_distance_is_the_taxicab_distance_to_the_nearest_solid():
    occ = scene(24, 16, 16, 1, spheres=2, sprinkle=0.01)
    want = ndi.distance_transform_cdt(occ == 0, metric="taxicab")
    np.testing.assert_array_equal(O.octree_distance(occ), want)
    empty = np.zeros((8, 16, 24), np.uint8)
    assert (O.octree_distance(empty) == 24 + 16 + 8).all()


@pytest.mark.parametrize("thresh", [(2, 4, 16), (1, 2, 4), (0, 1, 2)])
def test_leaves_closed_form_and_exact_cover(thresh):
    occ = scene(64, 48, 32, 2, spheres=2, sprinkle=0.002)
    D, H, W = occ.shape
    dist = ndi.distance_transform_cdt(occ == 0, metric="taxicab")
    l0, l1, l2, l3 = O.octree_leaves(occ, thresh)

    def bmin(s):
        return dist.reshape(D // s, s, H // s, s, W // s, s).min(axis=(1, 3, 5))

    def up(a, f):
        return np.repeat(np.repeat(np.repeat(a, f, 0), f, 1), f, 2)

    w3 = bmin(8) > thresh[2]
    w2 = (up(w3, 2) == 0) & (bmin(4) > thresh[1])
    w1 = (up(w3, 4) == 0) & (up(w2, 2) == 0) & (bmin(2) > thresh[0])
    covered = up(w3, 8) | up(w2, 4) | up(w1, 2)
    w0 = ~covered & (occ == 0)
    for got, want, n in ((l3, w3, 3), (l2, w2, 2), (l1, w1, 1), (l0, w0, 0)):
        np.testing.assert_array_equal(got, want.astype(np.uint8), err_msg="level %d" % n)
    cover = up(l3, 8).astype(int) + up(l2, 4) + up(l1, 2) + l0 + occ
    assert (cover == 1).all()
    assert l3.any() or thresh[2] >= 16


def brute_neighbours(lev, L, x, y, z):
    """Per face (x-, x+, y-, y+, z-, z+): the owners behind the face in order of first appearance when
    the face's fine cells are ranked by interleaved coordinates (slower axis high), with shared face
    counts; 'B' for a domain face."""
    D, H, W = lev.shape
    s = 1 << L
    X0, Y0, Z0 = x * s, y * s, z * s
    out = []
    for ax in range(3):
        for sd in (-1, 1):
            c0 = (X0, Y0, Z0)[ax]
            f = c0 - 1 if sd < 0 else c0 + s
            if f < 0 or f >= (W, H, D)[ax]:
                out.append(("B",))
                continue
            slow, fast = (2, 1) if ax == 0 else (2, 0) if ax == 1 else (1, 0)
            cells = []
            for a in range(s):
                for b in range(s):
                    p = [X0, Y0, Z0]
                    p[ax] = f
                    p[slow] += a
                    p[fast] += b
                    key = 0
                    for bit in range(L):
                        key |= ((a >> bit) & 1) << (2 * bit + 1) | ((b >> bit) & 1) << (2 * bit)
                    cells.append((key, p))
            cells.sort(key=lambda t: t[0])
            seen = {}
            for _, p in cells:
                m = int(lev[p[2], p[1], p[0]])
                if m == 4:
                    continue
                own = (m, p[0] >> m, p[1] >> m, p[2] >> m)
                seen[own] = seen.get(own, 0) + 1
            out.extend(seen.items())
    return out


@pytest.mark.parametrize("thresh", [(1, 2, 4), (2, 4, 16)])
def test_neighbour_lists_match_brute_force_adjacency(thresh):
    occ = scene(48, 32, 40, 3, spheres=2, sprinkle=0.003)
    leaves = O.octree_leaves(occ, thresh)
    lev = owner_levels(leaves, occ)
    rng = np.random.default_rng(0)
    checked = 0
    for L in range(4):
        cells = np.argwhere(leaves[L])
        if len(cells) > 150:
            cells = cells[rng.choice(len(cells), 150, replace=False)]
        for z, y, x in cells:
            got = O.octree_neighbors(leaves, L, int(x), int(y), int(z))
            want = brute_neighbours(lev, L, int(x), int(y), int(z))
            assert len(got) == len(want), (L, x, y, z)
            for g, wnt in zip(got, want):
                if wnt[0] == "B":
                    # an open-boundary entry: level 3, coordinates outside the level-3 grid
                    assert g[0] == 3 and not (0 <= g[1] < 48 // 8 and 0 <= g[2] < 32 // 8 and 0 <= g[3] < 40 // 8)
                else:
                    (m, bx, by, bz), cnt = wnt
                    assert tuple(g) == (m, bx, by, bz), (L, x, y, z)
                    assert cnt == 4 ** min(L, m)  # one entry covers the smaller cell's whole face
            checked += 1
    assert checked > 300


def system(leaves, lev, solved):
    """Sparse A and the index of every solved leaf: A p = -rhs, A = diag(coeff) - weights, built from
    brute-force adjacency (leaves not solved are held at p = 0)."""
    index = {}
    for L in range(4):
        for z, y, x in np.argwhere(solved[L]):
            index[(L, int(x), int(y), int(z))] = len(index)
    rows, cols, vals = [], [], []
    for (L, x, y, z), r in index.items():
        diag = 0.0
        for e in brute_neighbours(lev, L, x, y, z):
            if e[0] == "B":
                diag += 4.0 ** L / (0.5 * 2 ** L + 4.0)
                continue
            (m, bx, by, bz), cnt = e
            wgt = cnt / (0.5 * 2 ** L + 0.5 * 2 ** m)
            diag += wgt
            c = index.get((m, bx, by, bz))
            if c is not None:
                rows.append(r)
                cols.append(c)
                vals.append(-wgt)
        rows.append(r)
        cols.append(r)
        vals.append(diag)
    n = len(index)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n)), index


def solved_sets(leaves):
    """The cells the solve updates: every leaf of levels 0-2; level-3 blocks whose column's block at
    z = 0 is a leaf (the reference's level-3 list test, ZQ_SmokeOctreeGrid3D.cpp:8416), leaves only."""
    l3 = leaves[3] & np.broadcast_to(leaves[3][:1], leaves[3].shape)
    return [leaves[0], leaves[1], leaves[2], l3]


@pytest.mark.parametrize("thresh", [(1, 2, 4), (2, 4, 16)])
def test_converged_solve_matches_a_direct_sparse_solve(thresh):
    occ = scene(32, 32, 40, 5, centres=[(14.0, 17.0, 27.0)])
    occ[30, 5, 25] = 1
    rng = np.random.default_rng(7)
    u, v, w = faces(rng, 32, 32, 40)
    leaves = O.octree_leaves(occ, thresh)
    lev = owner_levels(leaves, occ)
    divs = O.octree_divergence(u, v, w)
    ps = O.octree_solve(leaves, divs, 4000)
    solved = solved_sets(leaves)
    A, index = system(leaves, lev, solved)
    rhs = np.zeros(len(index))
    for (L, x, y, z), r in index.items():
        rhs[r] = -float(divs[L][z, y, x])
    p = spla.spsolve(A.tocsc(), rhs)
    got = np.array([ps[L][z, y, x] for (L, x, y, z) in index])
    scale = np.abs(p).max()
    np.testing.assert_allclose(got, p, rtol=0, atol=2e-4 * scale)
    assert any(L == 3 for L, *_ in index) or thresh[2] >= 16


def fine_div(u, v, w):
    return (u[:, :, 1:] - u[:, :, :-1]) + (v[:, 1:, :] - v[:, :-1, :]) + (w[1:] - w[:-1])


def test_converged_projection_is_divergence_free_per_leaf_and_per_cell():
    W, H, D = 48, 40, 32
    occ = scene(W, H, D, 9, centres=[(14.0, 12.0, 22.0), (34.0, 28.0, 25.0)])
    occ[20, 30, 9] = 1
    rng = np.random.default_rng(11)
    u, v, w = faces(rng, W, H, D)
    thresh = (1, 2, 4)
    leaves = O.octree_leaves(occ, thresh)
    solved = solved_sets(leaves)
    assert leaves[3].any() and (solved[3] == leaves[3]).all(), "scene must keep every level-3 leaf solved"
    divs = O.octree_divergence(u, v, w)
    ps = O.octree_solve(leaves, divs, 6000)
    O.octree_adjust(leaves, ps, u, v, w)
    scale = max(np.abs(a).max() for a in (u, v, w))
    # after the correction every leaf's net outflow vanishes
    after = O.octree_divergence(u, v, w)
    for L in range(4):
        assert np.abs(after[L][leaves[L] == 1]).max() < 2e-3 * scale, "level %d" % L
    # the block operator then leaves every fluid fine cell divergence-free (rows written one by one)
    ru, rv, rw = u.copy(), v.copy(), w.copy()
    for L in (1, 2, 3):
        O.sub_project_regular(ru, rv, rw, 1 << L, O.sub_matrix(1 << L), leaves[L])
    d = fine_div(ru, rv, rw)
    assert np.abs(d[occ == 0]).max() < 2e-3 * scale
    # with the reference's level-1 w write-back the level-1 interiors no longer balance
    O.octree_sub_project(u, v, w, leaves)
    up1 = np.repeat(np.repeat(np.repeat(leaves[1], 2, 0), 2, 1), 2, 2) == 1
    d2 = fine_div(u, v, w)
    assert np.abs(d2[(occ == 0) & ~up1]).max() < 2e-3 * scale
    for bk, bj, bi in np.argwhere(leaves[1])[:20]:
        z, y, x = 2 * bk + 1, 2 * bj, 2 * bi
        assert (w[z, y:y + 2, x:x + 2] == w[z, y, x]).all()


def test_level3_leaves_outside_the_z0_test_keep_zero_pressure():
    """A level-3 leaf whose column's block at z = 0 is no leaf is never updated (reference slip kept)."""
    W, H, D = 32, 32, 32
    occ = np.zeros((D, H, W), np.uint8)
    occ[2, 4, 4] = 1  # spoils the z = 0 level-3 block of column (0, 0) only
    leaves = O.octree_leaves(occ, (1, 2, 4))
    assert leaves[3][0, 0, 0] == 0 and leaves[3][2, 0, 0] == 1 and leaves[3][0, 1, 1] == 1
    rng = np.random.default_rng(3)
    u, v, w = faces(rng, W, H, D)
    ps = O.octree_solve(leaves, O.octree_divergence(u, v, w), 50)
    assert ps[3][2, 0, 0] == 0.0 and ps[3][3, 0, 0] == 0.0
    assert ps[3][2, 1, 1] != 0.0


def test_closed_shell_leaves_every_level3_leaf_unsolved_and_balances_the_rest():
    """With a wall shell no z = 0 block is a leaf, so no level-3 leaf is solved; the other leaves
    converge to net outflow mean * volume (the closed right-hand side)."""
    W, H, D = 48, 48, 40
    occ = scene(W, H, D, 4, spheres=1, shell=True)
    rng = np.random.default_rng(5)
    u, v, w = faces(rng, W, H, D)
    thresh = (1, 2, 4)
    leaves = O.octree_leaves(occ, thresh)
    assert leaves[3].any()
    mean = O.lib().orc_oct_closed_mean(W, H, D, occ, u, v, w)
    ps = O.octree_solve(leaves, O.octree_divergence(u, v, w), 5000, closed=True, mean=mean)
    assert (ps[3] == 0).all()
    O.octree_adjust(leaves, ps, u, v, w)
    after = O.octree_divergence(u, v, w)
    scale = max(np.abs(a).max() for a in (u, v, w))
    for L in range(3):
        sel = leaves[L] == 1
        np.testing.assert_allclose(after[L][sel], mean * 8.0 ** L, rtol=0, atol=2e-3 * scale, err_msg="level %d" % L)


def test_closed_mean_is_the_float_mean_over_fluid_cells():
    W, H, D = 16, 16, 8
    occ = scene(W, H, D, 6, spheres=1, shell=True)
    rng = np.random.default_rng(2)
    u, v, w = faces(rng, W, H, D)
    got = O.lib().orc_oct_closed_mean(W, H, D, occ, u, v, w)
    d = fine_div(u.astype(np.float64), v.astype(np.float64), w.astype(np.float64))
    assert abs(got - d[occ == 0].mean()) < 1e-5


def region_labels(cells, W=8, H=8, D=8):
    """The oracle's special-region labels of a mask given as (x, y, z) cells (as level-0 leaves)."""
    l0 = np.zeros((D, H, W), np.uint8)
    for x, y, z in cells:
        l0[z, y, x] = 1
    leaves = [l0, np.zeros((D // 2, H // 2, W // 2), np.uint8)]
    n, lab = O.octree_special_regions(*O.zeros_faces(W, H, D), leaves, 0)
    return n, {c: int(lab[c[2], c[1], c[0]]) for c in cells}


def test_special_region_labels_in_one_plane_are_scan_order_components():
    """With every masked cell in one plane the reference's labelling (ZQ_SmokeOctreeGrid3D.cpp:
    5169-5320) is a 4-connected flood fill numbered by each component's first cell in scan order,
    which is how scipy numbers its components."""
    rng = np.random.default_rng(8)
    W, H, D = 24, 16, 8
    plane = rng.random((H, W)) < 0.45
    want, n = ndi.label(plane)  # 4-connectivity in 2D
    l0 = np.zeros((D, H, W), np.uint8)
    l0[5] = plane
    leaves = [l0, np.zeros((D // 2, H // 2, W // 2), np.uint8)]
    got, lab = O.octree_special_regions(*O.zeros_faces(W, H, D), leaves, 0)
    assert got == n and n > 5
    np.testing.assert_array_equal(lab[5], want)
    assert not lab[np.arange(D) != 5].any()


def test_special_region_labels_keep_the_reference_fill():
    """Hand-worked cases of _find_special_label (ZQ_SmokeOctreeGrid3D.cpp:5169-5320).
    - A cell reached only along z is queued but not labelled, so it seeds a region of its own: cells
      (0,0,0), (1,0,0), (1,0,1) give (0,0,0) and the second plane's row-0 seed (1,0,1) label 1, and
      (1,0,0) label 2, where a flood fill would give one region.
    - After the first seed a pass also seeds the first cell of row 0 of every plane above, under the
      same label: (2,3,0), (4,0,2), (5,0,2) are one region, not two; (4,2,2), off row 0, is not
      seeded with (2,3,0).
    - Two cells above each other reached only along z queue each other without end (the reference's
      queue then overflows): a column of three cells off row 0 raises; a column of two gives two
      regions."""
    n, lab = region_labels([(0, 0, 0), (1, 0, 0), (1, 0, 1)])
    assert n == 2 and lab == {(0, 0, 0): 1, (1, 0, 1): 1, (1, 0, 0): 2}
    n, lab = region_labels([(2, 3, 0), (4, 0, 2), (5, 0, 2)])
    assert n == 1 and set(lab.values()) == {1}
    n, lab = region_labels([(2, 3, 0), (4, 2, 2)])  # above the first seed only row 0 is searched
    assert n == 2 and lab == {(2, 3, 0): 1, (4, 2, 2): 2}
    n, lab = region_labels([(1, 1, 0), (1, 1, 1)])
    assert n == 2 and lab == {(1, 1, 0): 1, (1, 1, 1): 2}
    with pytest.raises(RuntimeError):
        region_labels([(1, 1, 0), (1, 1, 1), (1, 1, 2)])
    # in row 0 every plane's cell is a seed of the first pass, so the same column ends
    n, lab = region_labels([(1, 0, 0), (1, 0, 1), (1, 0, 2)])
    assert n == 1 and set(lab.values()) == {1}


def test_special_regions_are_projected_alone():
    """Octree guiding's special regions (ZQ_SmokeOctreeGrid3D.cpp:546-720), each projected alone by the
    closed solver with the region's mean divergence: converged, every region cell's divergence is that
    mean, and only faces between two cells of one region change.  In this scene every region cell has a
    neighbour in its plane and no region reaches row 0, so the reference's labels are the 6-connected
    components in scan order (scipy's).  With zero iterations nothing changes."""
    W, H, D = 32, 24, 16
    occ = scene(W, H, D, 12, centres=[(7.0, 8.0, 8.0), (24.0, 15.0, 8.0)])
    leaves = O.octree_leaves(occ, (1, 2, 4))
    lev = owner_levels(leaves, occ)
    want, n = ndi.label((lev == 0) | (lev == 1))
    assert n >= 2
    rng = np.random.default_rng(4)
    u, v, w = faces(rng, W, H, D)
    u0, v0, w0 = u.copy(), v.copy(), w.copy()
    n0, _ = O.octree_special_regions(u, v, w, leaves, 0)
    assert n0 == n
    for a, b in ((u, u0), (v, v0), (w, w0)):
        np.testing.assert_array_equal(a, b)
    got, labels = O.octree_special_regions(u, v, w, leaves, 100000)
    assert got == n
    np.testing.assert_array_equal(labels, want)
    d0 = fine_div(u0.astype(np.float64), v0.astype(np.float64), w0.astype(np.float64))
    d1 = fine_div(u.astype(np.float64), v.astype(np.float64), w.astype(np.float64))
    for r in range(1, n + 1):
        m = want == r
        np.testing.assert_allclose(d1[m], d0[m].mean(), rtol=0, atol=2e-3, err_msg="region %d" % r)
    # a face may change only between two cells of one region
    pad = np.pad(want, 1, constant_values=-1)
    same_u = (pad[1:-1, 1:-1, :-1] == pad[1:-1, 1:-1, 1:]) & (pad[1:-1, 1:-1, :-1] > 0)
    same_v = (pad[1:-1, :-1, 1:-1] == pad[1:-1, 1:, 1:-1]) & (pad[1:-1, :-1, 1:-1] > 0)
    same_w = (pad[:-1, 1:-1, 1:-1] == pad[1:, 1:-1, 1:-1]) & (pad[:-1, 1:-1, 1:-1] > 0)
    for a, b, s, name in ((u, u0, same_u, "u"), (v, v0, same_v, "v"), (w, w0, same_w, "w")):
        assert not np.any((a != b) & ~s), name
        assert np.any((a != b) & s), name
