"""The reference's mesh refinement and DOF numbering in the host setup (CPU only).

Golden vectors: tests/golden/refinery_golden.json holds the input meshes and the refined meshes
that the reference's refinery tests compare against (geometry/test_aux/standard_quad.cpp: the
unit square in four vertex/edge orientations; standard_hexa.cpp: the unit cube in three;
tetris_quad.cpp / tetris_hexa.cpp: 4-cell meshes with shared, differently oriented entities),
written by tests/golden/extract_refinery.py.  Every array (coordinates, vertices-at-edge/quad/
hexa, edges-at-quad/hexa, quads-at-hexa) must match exactly.
Then the numbering of the benchmark hierarchies built on it: nesting of the vertex sets, the
Lagrange-2 numbering (vertices, edges, cells), the reference's grid-transfer golden matrix
(in test_oracle_golden.py), and the solver KAT of poisson_dirichlet in FEAT's numbering.
"""
import json
import os

import numpy as np
import pytest

from feat3_amd import problems as pb
from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "refinery_golden.json")
CASES = json.load(open(GOLDEN))["cases"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_refinement_matches_reference_golden(case):
    dim = case["dim"]
    fine = pb.refine_mesh(dim, case["coarse"])
    for key, want in case["fine"].items():
        got = fine[key]
        assert got.size == len(want), key
        if key == "vtx":
            assert np.array_equal(got, np.array(want, np.float64)), key
        else:
            assert np.array_equal(got, np.array(want, np.uint32)), key


def test_refinement_rejects_inconsistent_mesh():
    c = dict(CASES[0]["coarse"])
    c["e_q"] = [0, 1, 2, 2]   # local edge 3 named as edge 2: no fine edge joins those vertices
    from feat3_amd.capi import Feat3HipError
    with pytest.raises(Feat3HipError):
        pb.refine_mesh(2, c)


def test_feat_order_level1_square():
    # corners (0,0) (1,0) (0,1) (1,1), midpoints of edges (0,1) (2,3) (0,2) (1,3), centre
    assert pb.feat_order(2, 1, 1).tolist() == [0, 2, 6, 8, 1, 7, 3, 5, 4]
    # the same as the golden refined unit square's coordinates
    vtx = np.array(CASES[0]["fine"]["vtx"]).reshape(-1, 2)
    assert np.array_equal(pb.lattice_order(vtx), pb.feat_order(2, 1, 1))


@pytest.mark.parametrize("dim,deg,lvl", [(2, 1, 5), (3, 1, 3), (2, 2, 4)])
def test_feat_order_is_nested_permutation(dim, deg, lvl):
    o = pb.feat_order(dim, deg, lvl)
    m = (deg << lvl) + 1
    assert np.array_equal(np.sort(o), np.arange(m ** dim))
    # the vertices of the coarser level come first, at the same positions (lattice index x2)
    oc = pb.feat_order(dim, deg, lvl - 1)
    mc = (deg << (lvl - 1)) + 1
    cc = np.stack(np.unravel_index(oc, (mc,) * dim)[::-1], 1)   # (x, y[, z]) of each coarse DOF
    fine_lex = np.ravel_multi_index(tuple((2 * cc)[:, ::-1].T), (m,) * dim)
    assert np.array_equal(o[: oc.size], fine_lex)


@pytest.mark.parametrize("dim,lvl", [(2, 3), (3, 2)])
def test_lagrange2_numbering_is_next_vertex_numbering(dim, lvl):
    """DofMappingUniform of Lagrange-2: vertices, then edges, then cells -- the vertex numbering
    of the next refinement (whose vertices are the old ones, edge midpoints, quad centres).  In 3D
    Lagrange-2 also has face DOFs; the setup is 2D for Q2."""
    if dim == 2:
        assert np.array_equal(pb.feat_order(2, 2, lvl), pb.feat_order(2, 1, lvl + 1))
    else:
        n = (2 ** (lvl + 1) + 1) ** 3
        assert pb.feat_order(3, 1, lvl + 1).size == n


@pytest.mark.parametrize("dim,lvl", [(2, 3), (3, 2)])
def test_feat_order_follows_refined_meshes(dim, lvl):
    """feat3_setup_feat_order equals refining UnitCubeFactory's mesh step by step through the
    one-step refinement the golden tests pin, then ranking the vertex coordinates."""
    c = CASES[0]["coarse"] if dim == 2 else CASES[4]["coarse"]
    # orientation 0 of standard_quad / standard_hexa is UnitCubeFactory's mesh
    mesh = {k: np.array(v) for k, v in c.items()}
    for _ in range(lvl):
        mesh = pb.refine_mesh(dim, mesh)
    vtx = mesh["vtx"].reshape(-1, dim)
    assert np.array_equal(pb.lattice_order(vtx), pb.feat_order(dim, 1, lvl))


def test_lattice_order_rejects_repeated_nodes():
    from feat3_amd.capi import Feat3HipError
    with pytest.raises(Feat3HipError):
        pb.lattice_order(np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 0.0]]))


@pytest.mark.parametrize("dim,deg,lf", [(2, 1, 3), (3, 1, 2), (2, 2, 2)])
def test_feat_hierarchy_is_symmetric_permutation(dim, deg, lf):
    """Every level of the FEAT-numbered hierarchy is the lexicographic one under the same DOF
    permutation on rows and columns (values moved, bit for bit), columns ascending, Dirichlet
    lists ascending."""
    H = pb.hierarchy(dim, deg, lf, 1)
    G = pb.hierarchy(dim, deg, lf, 1, numbering="feat")
    for i, (L, M) in enumerate(zip(H.levels, G.levels)):
        o = G.orders[i]
        Ad, Md = L.A.to_dense(), M.A.to_dense()
        assert np.array_equal(Md, Ad[np.ix_(o, o)])
        for r in range(M.A.rows):
            c = M.A.col_ind[M.A.row_ptr[r]:M.A.row_ptr[r + 1]]
            assert np.all(np.diff(c.astype(np.int64)) > 0)
        assert np.array_equal(M.bnd, np.sort(pb.inverse_order(o)[L.bnd]))
        assert np.all(np.diff(M.bnd.astype(np.int64)) > 0)
        if L.P is not None:
            oc = G.orders[i + 1]
            assert np.array_equal(M.P.to_dense(), L.P.to_dense()[np.ix_(o, oc)])
            assert np.array_equal(M.R.to_dense(), M.P.to_dense().T)


def test_poisson_dirichlet_kat_in_feat_numbering():
    """applications/CMakeLists.txt:141-150 + poisson_dirichlet.cpp:266-276: PCG-MG (Richardson
    Jacobi 0.7, V(4,4)) on Q1 Poisson levels 5..0 with the ExpBubble load converges in 4 +- 1
    iterations -- now on the hierarchy in the reference's own numbering."""
    H = pb.hierarchy(2, 1, 5, 0, numbering="feat")
    b = pb.filtered_rhs(H, func=2)
    S = po.System.from_hierarchy(H)
    prm = po.params(omega_jac=0.7, coarse=po.COARSE_SMOOTHER)
    x = np.zeros(H.n)
    st, it, _ = S.solve(x, b, prm)
    assert st == 2 and abs(it - 4) <= 1
    # the same solution as the lexicographic hierarchy's, permuted
    Hl = pb.hierarchy(2, 1, 5, 0)
    xl = np.zeros(Hl.n)
    po.System.from_hierarchy(Hl).solve(xl, pb.filtered_rhs(Hl, func=2), prm)
    assert np.max(np.abs(x - xl[H.orders[0]])) <= 1e-12 * np.max(np.abs(xl))
