"""Pins the oracle's coarse + sub-grid projection (COARSE_OPEN_POISSON_SOR, the reference's default
globalGridSolver, ZQ_SmokeSimulationUtils3D.cpp:3688-3795).

What fixes each piece:
- sub-block operator: the reference's own mat.txt (s = 2, tests/golden/sub_matrix_s2.txt) and, for
  larger s, what a projection must do (divergence-free inputs pass unchanged, the output's cell
  divergences are uniform, applying it twice changes nothing);
- face fractions, coarse occupancy, block flags, averages: independent numpy restatements
  (reshape-and-sum, exact arithmetic);
- coarse solver: with unit fractions (s = 1) it must equal the full-grid SOR bit for bit, and at
  convergence it must solve the reference's face-fraction matrix (ZQ_InitMatOpenPoisson3D) built
  densely here;
- coarse -> fine: the fine flux added through each coarse face must be fraction * s^2 * delta
  (ties ZQ_MapCoarse2Fine3D to ZQ_InitCoarseGrid3D);
- blocks with solids: each fluid component ends with uniform divergence equal to its mean,
  faces touching solids do not move;
- the whole pipeline: with a converged coarse solve the fine divergence vanishes.
"""
import os

import numpy as np
import pytest

import zq_oracle as O
from util import faces

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def div_of(u, v, w):
    return (u[:, :, 1:] - u[:, :, :-1]) + (v[:, 1:] - v[:, :-1]) + (w[1:] - w[:-1])


def block_div(x, s):
    """Cell divergences of one block from its faces in the reference's flat order."""
    nu = s * s * (s + 1)
    u = x[:nu].reshape(s, s, s + 1).astype(np.float64)
    v = x[nu:2 * nu].reshape(s, s + 1, s).astype(np.float64)
    w = x[2 * nu:].reshape(s + 1, s, s).astype(np.float64)
    return div_of(u, v, w)


def with_interior(x, y, s):
    z = x.copy()
    z[O.interior_rows(s)] = y
    return z


# ---------------------------------------------------------------- sub-block operator
def test_sub_matrix_matches_reference_fixture():
    ref = np.loadtxt(os.path.join(GOLDEN, "sub_matrix_s2.txt"))
    M = O.sub_matrix(2)
    assert M.shape == ref.shape == (12, 36)
    # the fixture is printed with 5 decimals
    np.testing.assert_allclose(M, ref, rtol=0, atol=5.1e-6)


@pytest.mark.parametrize("s", [2, 3, 4])
def test_sub_matrix_is_the_block_projection(s):
    rng = np.random.default_rng(7 + s)
    M = O.sub_matrix(s).astype(np.float64)
    R, F = M.shape
    assert (R, F) == (3 * s * s * (s - 1), 3 * s * s * (s + 1))
    rows = O.interior_rows(s)
    for _ in range(4):
        x = rng.standard_normal(F)
        y = M @ x
        d = block_div(with_interior(x, y, s), s)
        # uniform divergence equal to the block mean (least-squares residual of the pinned system)
        np.testing.assert_allclose(d, block_div(x, s).mean(), atol=2e-5)
        # idempotent
        np.testing.assert_allclose(M @ with_interior(x, y, s), y, atol=2e-5)
    # a divergence-free field (discrete curl of a random potential) passes unchanged
    a = rng.standard_normal((3, s + 1, s + 1, s + 1))
    U, V, Wf, _ = O.block_face_index(s)
    x = np.zeros(F)
    for k in range(s):
        for j in range(s):
            for i in range(s + 1):
                x[U(i, j, k)] = (a[2, k, j + 1, i] - a[2, k, j, i]) - (a[1, k + 1, j, i] - a[1, k, j, i])
    for k in range(s):
        for j in range(s + 1):
            for i in range(s):
                x[V(i, j, k)] = (a[0, k + 1, j, i] - a[0, k, j, i]) - (a[2, k, j, i + 1] - a[2, k, j, i])
    for k in range(s + 1):
        for j in range(s):
            for i in range(s):
                x[Wf(i, j, k)] = (a[1, k, j, i + 1] - a[1, k, j, i]) - (a[0, k, j + 1, i] - a[0, k, j, i])
    assert np.abs(block_div(x, s)).max() < 1e-12
    np.testing.assert_allclose(M @ x, x[rows], atol=2e-5)


# ---------------------------------------------------------------- set-up and averages
def ratios_numpy(occ, s):
    D, H, W = occ.shape
    Dc, Hc, Wc = D // s, H // s, W // s
    o = occ.astype(bool)
    sf = np.float32(s * s)

    def frac(solid2d, a, b):  # (a*s, b*s) boolean plane -> (a, b) fractions
        n = solid2d.reshape(a, s, b, s).sum(axis=(1, 3)).astype(np.float32)
        return np.float32(1) - n / sf

    rU = np.zeros((Dc, Hc, Wc + 1), np.float32)
    for i in range(Wc + 1):
        x = i * s
        sol = o[:, :, 0] if i == 0 else o[:, :, W - 1] if i == Wc else (o[:, :, x] | o[:, :, x - 1])
        rU[:, :, i] = frac(sol, Dc, Hc)
    rV = np.zeros((Dc, Hc + 1, Wc), np.float32)
    for j in range(Hc + 1):
        y = j * s
        sol = o[:, 0] if j == 0 else o[:, H - 1] if j == Hc else (o[:, y] | o[:, y - 1])
        rV[:, j] = frac(sol, Dc, Wc)
    rW = np.zeros((Dc + 1, Hc, Wc), np.float32)
    for k in range(Dc + 1):
        z = k * s
        sol = o[0] if k == 0 else o[D - 1] if k == Dc else (o[z] | o[z - 1])
        rW[k] = frac(sol, Hc, Wc)
    cocc = ((rU[:, :, :-1] == 0) & (rU[:, :, 1:] == 0) & (rV[:, :-1] == 0) & (rV[:, 1:] == 0) &
            (rW[:-1] == 0) & (rW[1:] == 0)).astype(np.uint8)
    regular = (~o.reshape(Dc, s, Hc, s, Wc, s).any(axis=(1, 3, 5))).astype(np.uint8)
    return rU, rV, rW, cocc, regular


@pytest.mark.parametrize("s,frac", [(2, 0.3), (3, 0.5), (4, 0.85), (4, 0.0)])
def test_coarse_setup_matches_numpy(s, frac):
    rng = np.random.default_rng(s * 100 + int(frac * 10))
    occ = (rng.random((3 * s, 2 * s, 4 * s)) < frac).astype(np.uint8)
    occ[:s, :s, :s] = 1  # one fully solid block: a solid coarse cell when its faces close
    got = O.coarse_setup(occ, s)
    want = ratios_numpy(occ, s)
    for g, w_ in zip(got, want):
        assert g.dtype == w_.dtype and np.array_equal(g, w_)
    # a fully solid block whose neighbours are all solid is a solid coarse cell
    occ2 = np.zeros_like(occ)
    occ2[:2 * s, :2 * s, :2 * s] = 1
    assert O.coarse_setup(occ2, s)[3][0, 0, 0] == 1


def test_fine_to_coarse_exact_sums(rng):
    s = 4
    W, H, D = 12, 8, 16
    # multiples of 2^-8 below 2^10: every partial sum is exact in float32
    u, v, w = [np.round(a * 64).astype(np.float32) / np.float32(256) for a in faces(rng, W, H, D, 8.0)]
    cu, cv, cw = O.fine_to_coarse(u, v, w, s)
    sf = np.float32(s * s)
    eu = u[:, :, ::s].reshape(D // s, s, H // s, s, W // s + 1).sum(axis=(1, 3), dtype=np.float64)
    ev = v[:, ::s, :].reshape(D // s, s, H // s + 1, W // s, s).sum(axis=(1, 4), dtype=np.float64)
    ew = w[::s].reshape(D // s + 1, H // s, s, W // s, s).sum(axis=(2, 4), dtype=np.float64)
    assert np.array_equal(cu, (eu.astype(np.float32) / sf))
    assert np.array_equal(cv, (ev.astype(np.float32) / sf))
    assert np.array_equal(cw, (ew.astype(np.float32) / sf))


# ---------------------------------------------------------------- coarse solver
@pytest.mark.parametrize("open_mask,omega", [(O.ALL_FACES, 1.0), (O.ALL_FACES, 1.7), (O.OPEN_YHI, 1.0),
                                             (O.OPEN_XLO | O.OPEN_ZHI, 1.5)])
def test_coarse_solver_with_unit_fractions_is_the_fine_sor(open_mask, omega):
    rng = np.random.default_rng(31 + open_mask)
    W, H, D = 11, 9, 8
    occ = (rng.random((D, H, W)) < 0.15).astype(np.uint8)
    u, v, w = faces(rng, W, H, D)
    rU, rV, rW, cocc, _ = O.coarse_setup(occ, 1)
    p, du, dv, dw = O.coarse_project(u, v, w, cocc, rU, rV, rW, 40, omega, open_mask)
    u1, v1, w1 = u.copy(), v.copy(), w.copy()
    pf = O.project(u1, v1, w1, occ, open_mask, 0.0, omega, 40, want_p=True)
    assert np.array_equal(p, pf)
    assert np.array_equal(du, u1 - u) and np.array_equal(dv, v1 - v) and np.array_equal(dw, w1 - w)


def face_ratio_matrix(cocc, rU, rV, rW):
    """ZQ_InitMatOpenPoisson3D (ZQ_SmokeSimulationUtils3D.cpp:476-592) as a dense matrix over fluid cells."""
    Dc, Hc, Wc = cocc.shape
    fluid = [(i, j, k) for k in range(Dc) for j in range(Hc) for i in range(Wc) if not cocc[k, j, i]]
    idx = {c: n for n, c in enumerate(fluid)}
    A = np.zeros((len(fluid), len(fluid)))
    for (i, j, k), r in idx.items():
        nbs = []
        if k > 0: nbs.append(((i, j, k - 1), rW[k, j, i]))
        if k < Dc - 1: nbs.append(((i, j, k + 1), rW[k + 1, j, i]))
        if j > 0: nbs.append(((i, j - 1, k), rV[k, j, i]))
        if j < Hc - 1: nbs.append(((i, j + 1, k), rV[k, j + 1, i]))
        if i > 0: nbs.append(((i - 1, j, k), rU[k, j, i]))
        if i < Wc - 1: nbs.append(((i + 1, j, k), rU[k, j, i + 1]))
        count = 6.0 - len(nbs)
        for c, rr in nbs:
            if c in idx:
                count += rr
                A[r, idx[c]] = rr
        A[r, r] = -count
    return A, fluid


def test_coarse_solver_converges_to_the_reference_system():
    rng = np.random.default_rng(5)
    s, W, H, D = 2, 12, 10, 8
    occ = np.zeros((D, H, W), np.uint8)
    occ[2:6, 3:7, 4:7] = 1              # an obstacle cutting through several blocks
    occ[0:2, 0:2, 0:2] = 1              # one fully solid block in a corner
    u, v, w = faces(rng, W, H, D)
    rU, rV, rW, cocc, _ = O.coarse_setup(occ, s)
    cu, cv, cw = O.fine_to_coarse(u, v, w, s)
    p, du, dv, dw = O.coarse_project(cu, cv, cw, cocc, rU, rV, rW, 3000, 1.6)
    A, fluid = face_ratio_matrix(cocc, rU, rV, rW)
    cdiv = div_of(cu.astype(np.float64), cv.astype(np.float64), cw.astype(np.float64))
    b = np.array([cdiv[k, j, i] for (i, j, k) in fluid])
    x = np.linalg.solve(A, b)
    got = np.array([p[k + 1, j + 1, i + 1] for (i, j, k) in fluid])
    np.testing.assert_allclose(got, x, atol=1e-4 * np.abs(x).max())
    # fraction-weighted coarse flux after the delta balances: the system is the one the delta solves
    nd = div_of(cu + rU * du, cv + rV * dv, cw + rW * dw)
    assert np.abs(nd[cocc == 0]).max() < 1e-4 * np.abs(cdiv).max()


# ---------------------------------------------------------------- coarse -> fine
def test_coarse_to_fine_adds_fraction_times_delta():
    rng = np.random.default_rng(11)
    s, W, H, D = 3, 12, 9, 6
    occ = (rng.random((D, H, W)) < 0.3).astype(np.uint8)
    u, v, w = faces(rng, W, H, D)
    rU, rV, rW, cocc, _ = O.coarse_setup(occ, s)
    du, dv, dw = [np.round(a * 8).astype(np.float32) / np.float32(8) for a in faces(rng, W // s, H // s, D // s)]
    # integer-valued faces so the sums below are exact
    u, v, w = [np.round(a * 16).astype(np.float32) for a in (u, v, w)]
    u1, v1, w1 = u.copy(), v.copy(), w.copy()
    O.coarse_to_fine(du, dv, dw, occ, u1, v1, w1, s)
    sf = s * s
    Dc, Hc, Wc = D // s, H // s, W // s
    eu, ev, ew = [(b.astype(np.float64) - a) for a, b in ((u, u1), (v, v1), (w, w1))]
    gu = eu[:, :, ::s].reshape(Dc, s, Hc, s, Wc + 1).sum(axis=(1, 3))
    gv = ev[:, ::s, :].reshape(Dc, s, Hc + 1, Wc, s).sum(axis=(1, 4))
    gw = ew[::s].reshape(Dc + 1, Hc, s, Wc, s).sum(axis=(2, 4))
    np.testing.assert_allclose(gu, rU.astype(np.float64) * sf * du, atol=1e-9)
    np.testing.assert_allclose(gw, rW.astype(np.float64) * sf * dw, atol=1e-9)
    # v: every face but the top row, which takes the delta of row Hc-1 (ZQ_SmokeSimulationUtils3D.cpp:3251)
    np.testing.assert_allclose(gv[:, :-1], rV[:, :-1].astype(np.float64) * sf * dv[:, :-1], atol=1e-9)
    np.testing.assert_allclose(gv[:, -1], rV[:, -1].astype(np.float64) * sf * dv[:, -2], atol=1e-9)
    # only faces on coarse planes move
    mu = np.ones(W + 1, bool)
    mu[::s] = False
    assert np.array_equal(u1[:, :, mu], u[:, :, mu])
    mv = np.ones(H + 1, bool)
    mv[::s] = False
    assert np.array_equal(v1[:, mv], v[:, mv])


# ---------------------------------------------------------------- blocks
def test_irregular_block_components_have_uniform_divergence():
    rng = np.random.default_rng(3)
    s = 4
    _, _, _, F = O.block_face_index(s)
    occ_b = np.zeros((s, s, s), np.uint8)
    occ_b[:, 2, :] = 1          # a wall splitting the block into two fluid components
    occ_b[3, 3, 3] = 1
    x = rng.standard_normal(F).astype(np.float32)
    y = O.sub_solve_block(x, occ_b, s)
    z = with_interior(x, y, s)
    d = block_div(z, s)
    d0 = block_div(x, s)
    fl = occ_b == 0
    for comp in (np.s_[:, :2, :], np.s_[:, 3:, :]):
        m = fl[comp]
        np.testing.assert_allclose(d[comp][m], d0[comp][m].mean(), atol=1e-5)
    # faces beside a solid cell keep their values
    U, V, Wf, _ = O.block_face_index(s)
    for k in range(s):
        for i in range(s):
            assert z[V(i, 2, k)] == x[V(i, 2, k)] and z[V(i, 3, k)] == x[V(i, 3, k)]
    # a block with a single fluid cell is returned unchanged
    one = np.ones((s, s, s), np.uint8)
    one[1, 1, 1] = 0
    assert np.array_equal(O.sub_solve_block(x, one, s), x[O.interior_rows(s)])


def test_irregular_block_without_solids_agrees_with_operator():
    rng = np.random.default_rng(9)
    s = 3
    x = rng.standard_normal(O.block_face_index(s)[3]).astype(np.float32)
    y = O.sub_solve_block(x, np.zeros((s, s, s), np.uint8), s)
    np.testing.assert_allclose(y, O.sub_matrix(s).astype(np.float64) @ x, atol=2e-5)


@pytest.mark.parametrize("s", [2, 4])
def test_regular_blocks_end_with_uniform_divergence(s):
    rng = np.random.default_rng(s)
    W, H, D = 4 * s, 2 * s, 3 * s
    u, v, w = faces(rng, W, H, D)
    occ = np.zeros((D, H, W), np.uint8)
    regular = np.ones((D // s, H // s, W // s), np.uint8)
    u1, v1, w1 = u.copy(), v.copy(), w.copy()
    O.sub_project_regular(u1, v1, w1, s, O.sub_matrix(s), regular)
    d0 = div_of(u.astype(np.float64), v.astype(np.float64), w.astype(np.float64))
    d1 = div_of(u1.astype(np.float64), v1.astype(np.float64), w1.astype(np.float64))
    bm = d0.reshape(D // s, s, H // s, s, W // s, s).mean(axis=(1, 3, 5))
    want = np.repeat(np.repeat(np.repeat(bm, s, 0), s, 1), s, 2)
    np.testing.assert_allclose(d1, want, atol=5e-5)
    # block boundary faces are untouched
    assert np.array_equal(u1[:, :, ::s], u[:, :, ::s])
    # the sequential float32 product equals the float64 product to rounding
    x = O._block_faces(u, v, w, s, 1, 0, 2)
    y = O._block_faces(u1, v1, w1, s, 1, 0, 2)[O.interior_rows(s)]
    np.testing.assert_allclose(y, O.sub_matrix(s).astype(np.float64) @ x.astype(np.float64), atol=1e-5)


def test_whole_coarse_projection_removes_divergence():
    rng = np.random.default_rng(21)
    s, W, H, D = 4, 16, 12, 12
    occ = np.zeros((D, H, W), np.uint8)
    zz, yy, xx = np.mgrid[:D, :H, :W]
    occ[((xx - 7.3) ** 2 + (yy - 5.6) ** 2 + (zz - 6.1) ** 2) < 2.6 ** 2] = 1   # a sphere across blocks
    u, v, w = faces(rng, W, H, D)
    # faces of solid cells still (a resting obstacle), so solid cells add no divergence of their own
    so = occ.astype(bool)
    u[:, :, 1:][so] = 0
    u[:, :, :-1][so] = 0
    v[:, 1:][so] = 0
    v[:, :-1][so] = 0
    w[1:][so] = 0
    w[:-1][so] = 0
    u0, v0, w0 = u.copy(), v.copy(), w.copy()
    O.coarse_projection(u, v, w, occ, s, iters=4000, omega=1.7)
    d = div_of(u.astype(np.float64), v.astype(np.float64), w.astype(np.float64))
    fl = occ == 0
    # every block below the top row ends divergence-free
    assert np.abs(d[:, :H - s][fl[:, :H - s]]).max() < 2e-4
    # the top row: its y = H faces took the delta of coarse v row Hc-1 instead of Hc
    # (ZQ_SmokeSimulationUtils3D.cpp:3251), which leaves each regular block there with a mean
    # divergence of (dv[Hc-1] - dv[Hc]) * s^2 / s^3
    rU, rV, rW, cocc, regular = O.coarse_setup(occ, s)
    cu, cv, cw = O.fine_to_coarse(u0, v0, w0, s)
    _, du, dv, dw = O.coarse_project(cu, cv, cw, cocc, rU, rV, rW, 4000, 1.7)
    Hc = H // s
    for bk in range(D // s):
        for bi in range(W // s):
            if not regular[bk, Hc - 1, bi]:
                continue
            blk = d[bk * s:(bk + 1) * s, H - s:, bi * s:(bi + 1) * s]
            want = (np.float64(dv[bk, Hc - 1, bi]) - dv[bk, Hc, bi]) / s
            np.testing.assert_allclose(blk, want, atol=2e-4)


def test_guiding_adds_the_blend_to_the_coarse_delta():
    """With no coarse iterations the delta is zero, so at s = 1 (coarse = fine, no sub-projection) the
    fluid faces become u + c (G - u) exactly and every other face stays."""
    rng = np.random.default_rng(17)
    W, H, D = 10, 8, 6
    occ = (rng.random((D, H, W)) < 0.2).astype(np.uint8)
    u, v, w = faces(rng, W, H, D)
    G = faces(rng, W, H, D)
    c = np.float32(0.3)
    u1, v1, w1 = u.copy(), v.copy(), w.copy()
    O.coarse_projection(u1, v1, w1, occ, 1, 0, guide=G, guide_coeff=0.3)
    fl = occ == 0
    mu = np.zeros(u.shape, bool)
    mu[:, :, 1:-1] = fl[:, :, 1:] & fl[:, :, :-1]
    mu[:, :, 0] = fl[:, :, 0]
    mu[:, :, -1] = fl[:, :, -1]
    want = np.where(mu, u + (np.float32(0) + c * (G[0] - u)), u)
    assert np.array_equal(u1, want)


def flux_kkt(cu, cv, cw, cocc, rU, rV, rW, vol=None):
    """The KKT systems of ZQ_InitMatOpenFlux3D (ZQ_SmokeSimulationUtils3D.cpp:749-936; vol None) and
    ZQ_InitMatClosedFlux3D (:937-1150; with the volume-fraction multiplier row), right-hand side of
    ZQ_SolveFlux3D (:3065-3079), solved densely; returns the face unknowns as du, dv, dw."""
    Dc, Hc, Wc = cocc.shape
    closed = vol is not None
    fu, fv, fw, lam = {}, {}, {}, {}
    n = 0
    for k in range(Dc):
        for j in range(Hc):
            for i in range(Wc + 1):
                if rU[k, j, i] != 0 and (not closed or 0 < i < Wc):
                    fu[i, j, k] = n; n += 1
    for k in range(Dc):
        for j in range(Hc + 1):
            for i in range(Wc):
                if rV[k, j, i] != 0 and (not closed or 0 < j < Hc):
                    fv[i, j, k] = n; n += 1
    for k in range(Dc + 1):
        for j in range(Hc):
            for i in range(Wc):
                if rW[k, j, i] != 0 and (not closed or 0 < k < Dc):
                    fw[i, j, k] = n; n += 1
    for k in range(Dc):
        for j in range(Hc):
            for i in range(Wc):
                if not cocc[k, j, i]:
                    lam[i, j, k] = n; n += 1
    kap = n if closed else None
    n += closed
    A = np.zeros((n, n))
    b = np.zeros(n)
    for (i, j, k), r in fu.items():
        c = rU[k, j, i]
        A[r, r] = 2 * c
        if i > 0: A[r, lam[i - 1, j, k]] = -c
        if i < Wc: A[r, lam[i, j, k]] = c
    for (i, j, k), r in fv.items():
        c = rV[k, j, i]
        A[r, r] = 2 * c
        if j > 0: A[r, lam[i, j - 1, k]] = -c
        if j < Hc: A[r, lam[i, j, k]] = c
    for (i, j, k), r in fw.items():
        c = rW[k, j, i]
        A[r, r] = 2 * c
        if k > 0: A[r, lam[i, j, k - 1]] = -c
        if k < Dc: A[r, lam[i, j, k]] = c
    for (i, j, k), r in lam.items():
        for key, tab, rat, sgn in (((i + 1, j, k), fu, rU[k, j, i + 1], -1), ((i, j, k), fu, rU[k, j, i], 1),
                                   ((i, j + 1, k), fv, rV[k, j + 1, i], -1), ((i, j, k), fv, rV[k, j, i], 1),
                                   ((i, j, k + 1), fw, rW[k + 1, j, i], -1), ((i, j, k), fw, rW[k, j, i], 1)):
            if rat > 0 and key in tab:
                A[r, tab[key]] = sgn * rat
        if closed and vol[k, j, i] > 0:
            A[r, kap] = vol[k, j, i]
            A[kap, r] = vol[k, j, i]
        b[r] = (np.float64(cu[k, j, i + 1]) - cu[k, j, i] + cv[k, j + 1, i] - cv[k, j, i] + cw[k + 1, j, i] - cw[k, j, i])
    x = np.linalg.lstsq(A, b, rcond=None)[0]
    du, dv, dw = (np.zeros(a.shape) for a in (cu, cv, cw))
    for (i, j, k), r in fu.items(): du[k, j, i] = x[r]
    for (i, j, k), r in fv.items(): dv[k, j, i] = x[r]
    for (i, j, k), r in fw.items(): dw[k, j, i] = x[r]
    return du, dv, dw


@pytest.mark.parametrize("closed", [False, True])
def test_flux_form_converges_to_the_reference_kkt_system(closed):
    """The *_FLUX_SOR solvers' CPU systems: face unknowns with a multiplier per fluid cell.  Eliminating
    the faces (du = (l_l - l_r)/2) leaves the face-fraction Poisson problem with the flux form's diagonal
    (open domain faces by their own fraction) and, closed, the divergence's mean per unit volume
    removed; the converged flux-form solve must reproduce the KKT face unknowns."""
    rng = np.random.default_rng(23 + closed)
    s, W, H, D = 2, 12, 10, 8
    occ = np.zeros((D, H, W), np.uint8)
    occ[2:6, 3:7, 4:7] = 1
    occ[0, :, :5] = 1                      # solids on a domain face: fractions below 1 there
    occ[3, 0, 8] = 1
    u, v, w = faces(rng, W, H, D)
    rU, rV, rW, cocc, _ = O.coarse_setup(occ, s)
    cu, cv, cw = O.fine_to_coarse(u, v, w, s)
    vol = O.coarse_volume(occ, s) if closed else None
    mask = 0 if closed else O.ALL_FACES
    _, du, dv, dw = O.coarse_project(cu, cv, cw, cocc, rU, rV, rW, 4000, 1.6, mask, vol, flux=True)
    ku, kv, kw = flux_kkt(cu, cv, cw, cocc, rU, rV, rW, vol)
    scale = max(np.abs(ku).max(), np.abs(kv).max(), np.abs(kw).max())
    for g, r in ((du, ku), (dv, kv), (dw, kw)):
        np.testing.assert_allclose(g, r, rtol=0, atol=2e-4 * scale)
    if not closed:
        # the Poisson form (open faces count 1) differs where the domain faces are partly solid
        _, pu, _, _ = O.coarse_project(cu, cv, cw, cocc, rU, rV, rW, 4000, 1.6, mask)
        assert np.abs(pu - ku).max() > 1e-3 * scale
