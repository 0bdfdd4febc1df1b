"""Three resolution levels on the multires oracle (pf_oracle_mr.c).

The reference's multires layout has two levels (multires_layout3d.h:25-26) and its own
projection iteration stalls on three (a CG on the unsymmetric rows A, whose row a
carries the 1/V_a of its cell).  What the finite-volume rows of
multires_pressure3d.cpp:783-812 fix at any number of levels is checked on the oracle's
rows for a level 2-1-0 layout: V A is symmetric (a conductance Laplacian), its row
sums vanish on rows with no AIR neighbour (constants span the null space of the
Neumann part) and are positive otherwise, and a projection with the exact solution of
the rows leaves a discrete divergence at round-off.  The export of the rows is itself
checked against the oracle's own projection on a two-level layout.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import pyoracle as O
from _helpers import bubble_lattice, refine_box, rel_err, three_level_bubble_layout


def _phase():
    ph = O.Phase(1.0, 0.01, 1.0, 1.0)
    ph.rho_tilde_0 = O.calibrate_rho_tilde0(ph, 1.0)
    return ph


def _grid(n, leaves, dt):
    ph = _phase()
    g = O.MRGrid(n, n, n, 1.0, faithful=False)
    g.set_leaves(leaves)
    x, y, z, u, v, w, t = bubble_lattice(n)
    ps = O.Particles(x, y, z, u, v, w, type=t)
    g.mark_cells(ps)
    g.p2g(ps, ph, 1.0)
    g.apply_gravity(dt, -9.81)
    g.apply_walls()
    return g, ph


def _system(g, ph, dt):
    cells, (ptr, col, val), rhs, pin = g.system(ph, dt)
    A = sp.csr_matrix((val, col, ptr), shape=(len(cells), len(cells)))
    vol = (2.0 ** g.leaves()[cells // 64, 0]) ** 3
    return cells, A, rhs, pin, vol


def exact_solution(A, rhs, vol):
    """x with A x = rhs to FP64 round-off: CG on the symmetric positive definite V A"""
    S = (sp.diags(vol) @ A).tocsr()
    b = vol * rhs
    x, info = sla.cg(S, b, rtol=1e-15, atol=0.0, maxiter=20000, M=sp.diags(1.0 / S.diagonal()))
    assert info == 0
    return x


def test_refine_restatement_matches_oracle_layouts():
    # the numpy refineFineCellBox against the oracle's two-level layouts
    n = 24
    box = [4, 3, 2, 13, 9, 8]
    coarse = O.mr_layout_fine_box(n, n, n, None, level=1)
    assert np.array_equal(refine_box(n, n, n, coarse, box), O.mr_layout_fine_box(n, n, n, box, level=1))
    lv = three_level_bubble_layout(64)
    assert set(np.unique(lv[:, 0]).tolist()) == {0, 1, 2}


def test_row_export_reproduces_the_two_level_projection():
    n, dt = 16, 0.02
    leaves = O.mr_layout_bubble_band(n, n, n)
    g, ph = _grid(n, leaves, dt)
    h = g.clone()
    st = g.project(ph, dt, O.mr_solve_cfg(max_iterations=20000, abs_tol=1e-13, rel_tol=0.0))
    assert st.converged == 1
    cells, A, rhs, pin, vol = _system(h, ph, dt)
    assert len(cells) == st.active_cells and pin == st.pinned_cell
    x = sla.spsolve(A.tocsc(), rhs)
    h.apply_pressure(ph, dt, x)
    assert rel_err(h.dense(O.P), g.dense(O.P)) < 1e-9
    for f in range(3):
        assert rel_err(h.dense(f), g.dense(f)) < 1e-9, f


def test_three_level_rows_are_a_conductance_laplacian():
    n, dt = 64, 0.02
    g, ph = _grid(n, three_level_bubble_layout(n), dt)
    cells, A, rhs, pin, vol = _system(g, ph, dt)
    lvl = g.leaves()[cells // 64, 0]
    assert set(np.unique(lvl).tolist()) == {0, 1, 2} and pin == -1
    S = (sp.diags(vol) @ A).tocsr()
    assert abs(S - S.T).max() <= 1e-14 * abs(S).max()
    rs = np.asarray(S.sum(axis=1)).ravel()
    d = S.diagonal()
    assert (rs >= -1e-12 * d).all()
    # AIR lies above the level-0 band around the free surface: rows of level 1 and 2
    # have only FLUID and SOLID neighbours
    assert (np.abs(rs[lvl > 0]) <= 1e-12 * d[lvl > 0]).all()
    assert (rs[lvl == 0] > 1e-6 * d[lvl == 0]).sum() > 0


def test_three_level_exact_projection_is_divergence_free():
    n, dt = 64, 0.02
    g, ph = _grid(n, three_level_bubble_layout(n), dt)
    cells, A, rhs, pin, vol = _system(g, ph, dt)
    x = exact_solution(A, rhs, vol)
    assert np.abs(A @ x - rhs).max() < 1e-10 * np.abs(rhs).max()
    g.apply_pressure(ph, dt, x)
    _, _, after, _ = g.system(ph, dt)
    assert np.abs(after).max() < 1e-9 * np.abs(rhs).max()


def test_block_grid_galerkin_operator_is_seven_point():
    # the multires cycle's first coarse space: leaf cells of level <= 2 aggregated by
    # the 4^3 block that holds them (P piecewise constant, R = P^T).  R (V A) P couples
    # face-adjacent blocks only, is symmetric, keeps the constants in the null space of
    # its Neumann part, and its face weights are the conductance sums over the
    # block planes that the GPU cycle assembles per fine face
    n, dt = 64, 0.02
    g, ph = _grid(n, three_level_bubble_layout(n), dt)
    cells, A, rhs, pin, vol = _system(g, ph, dt)
    lv = g.leaves()[cells // 64]
    L, loc = lv[:, 0], cells % 64
    s = 1 << L
    fine = np.stack([lv[:, 1] * 4 * s + (loc & 3) * s, lv[:, 2] * 4 * s + ((loc >> 2) & 3) * s,
                     lv[:, 3] * 4 * s + (loc >> 4) * s], 1)
    blk = fine // 4
    nb = n // 4
    bid = blk[:, 0] + nb * (blk[:, 1] + nb * blk[:, 2])
    ub, col = np.unique(bid, return_inverse=True)
    P = sp.csr_matrix((np.ones(len(cells)), (np.arange(len(cells)), col)), shape=(len(cells), len(ub)))
    S = (sp.diags(vol) @ A).tocsr()
    Sc = (P.T @ S @ P).tocoo()
    coords = np.stack([ub % nb, (ub // nb) % nb, ub // (nb * nb)], 1)
    off = Sc.row != Sc.col
    dist = np.abs(coords[Sc.row[off]] - coords[Sc.col[off]]).sum(1)
    assert (dist == 1).all() and (Sc.data[off] <= 0).all()
    Scc = Sc.tocsr()
    assert abs(Scc - Scc.T).max() <= 1e-14 * abs(Scc).max()
    rs = np.asarray(Scc.sum(axis=1)).ravel()
    assert np.allclose(rs, P.T @ np.asarray(S.sum(axis=1)).ravel(), rtol=0, atol=1e-12 * abs(Scc).max())
    assert (rs >= -1e-12 * Scc.diagonal()).all()
