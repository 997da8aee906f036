"""Mesh bundle derivation and the icosahedral generator (CPU).

The derivation restates the reference's scripts/prepare_mesh.py:117-300; re-deriving the
shipped pi bundle from its primitive arrays pins it (integers exact, floats to 1e-15
relative, the MFCT up/down triangle scan exact)."""

import numpy as np
import pytest

from fesom_jax_amd.mesh import (MESH_ARRAY_FIELDS, check_cw_orientation, level_masks,
                                load_mesh, save_mesh, wet_counts)
from fesom_jax_amd.meshgen import derive_bundle, icosa_counts, icosahedral_mesh, icosphere


def test_pi_bundle_rederived(pi_mesh):
    m = pi_mesh
    arr, meta = derive_bundle(m.coord_nod2D, m.geo_coord_nod2D, m.coast_flag, m.elem_nodes,
                              m.edges, m.edge_tri, m.nlevels, m.nlevels_nod2D, m.zbar, m.depth)
    for k in MESH_ARRAY_FIELDS:
        a, r = np.asarray(arr[k]), np.asarray(getattr(m, k))
        assert a.shape == r.shape, k
        if r.dtype.kind in "iu":
            assert np.array_equal(a, r), k
        else:
            scale = np.max(np.abs(r)) or 1.0
            assert np.max(np.abs(a - r)) <= 1e-15 * scale, k
    for k in ("nod2D", "elem2D", "edge2D", "nl", "edge2D_in"):
        assert meta[k] == getattr(m, k)


def test_pi_conventions(pi_mesh):
    """el1 lies left of n1->n2, boundary edges come last, triangles are clockwise,
    up/dn triangles contain n1/n2 (the conventions the kernels rely on)."""
    m = pi_mesh
    check_cw_orientation(m.coord_nod2D, m.elem_nodes)
    et = m.edge_tri
    assert np.all(et[:m.edge2D_in, 1] >= 0) and np.all(et[m.edge2D_in:, 1] < 0)
    ud = m.edge_up_dn_tri
    for i in np.nonzero(ud[:, 0] >= 0)[0][:500]:
        assert m.edges[i, 0] in m.elem_nodes[ud[i, 0]]
    for i in np.nonzero(ud[:, 1] >= 0)[0][:500]:
        assert m.edges[i, 1] in m.elem_nodes[ud[i, 1]]


def test_level_masks():
    lm, im = level_masks(np.array([1, 2]), np.array([4, 3]), 5)
    assert lm.tolist() == [[True, True, True, False, False], [False, True, False, False, False]]
    assert im.tolist() == [[True, True, True, True, False], [False, True, True, False, False]]


def test_orientation_gate_rejects_ccw(pi_mesh):
    en = pi_mesh.elem_nodes.copy()
    en[0, 1], en[0, 2] = en[0, 2], en[0, 1]
    with pytest.raises(ValueError):
        check_cw_orientation(pi_mesh.coord_nod2D, en)


def test_tiny_and_empty_meshes():
    """The generator's smallest outputs (one triangle with three coastal edges, two
    triangles sharing one edge) keep the conventions; an all-land draw is an error."""
    m = icosahedral_mesh(1, 4, seed=2, land_fraction=0.85, min_levels=2, polar_cap_deg=60.0)
    assert (m.nod2D, m.elem2D, m.edge2D, m.edge2D_in) == (3, 1, 3, 0)
    check_cw_orientation(m.coord_nod2D, m.elem_nodes)
    m = icosahedral_mesh(1, 4, seed=1, land_fraction=0.7, min_levels=2, polar_cap_deg=60.0)
    assert (m.nod2D, m.elem2D, m.edge2D, m.edge2D_in) == (4, 2, 5, 1)
    assert np.all(m.edge_tri[:1, 1] >= 0) and np.all(m.edge_tri[1:, 1] < 0)
    with pytest.raises(ValueError, match="no ocean"):
        icosahedral_mesh(1, 4, seed=0, land_fraction=0.9, min_levels=2, polar_cap_deg=60.0)
    # more degree-8 nodes than a 42-node sphere can hold without overlapping flips
    with pytest.raises(ValueError, match="degree 8"):
        icosahedral_mesh(1, 4, seed=0, degree8=100)


def test_icosphere_counts():
    for lev in range(4):
        v, f = icosphere(lev)
        n, e, _ = icosa_counts(lev)
        assert v.shape[0] == n and f.shape[0] == e
        assert np.allclose(np.linalg.norm(v, axis=1), 1.0)


@pytest.mark.parametrize("level,nl,deg8", [(3, 48, 0), (4, 70, 0), (3, 70, 12)])
def test_icosahedral_mesh_conventions(level, nl, deg8, tmp_path):
    m = icosahedral_mesh(level, nl, seed=5, degree8=deg8)
    n0, _, _ = icosa_counts(level)
    assert 0.6 < m.nod2D / n0 < 0.8                 # ~30% land carved out
    deg = np.bincount(m.edges.reshape(-1), minlength=m.nod2D)
    assert deg.max() == (8 if deg8 else 6)
    key = np.sort(m.edges, axis=1) @ np.array([m.nod2D, 1])
    assert np.unique(key).size == m.edge2D           # no edge twice
    # every triangle side is an edge
    en = m.elem_nodes
    sides = np.sort(np.concatenate([en[:, [0, 1]], en[:, [1, 2]], en[:, [2, 0]]]), axis=1)
    assert np.isin(sides @ np.array([m.nod2D, 1]), key).all()
    assert np.all(m.elem_area > 0)
    check_cw_orientation(m.coord_nod2D, m.elem_nodes)
    assert m.nlevels.min() >= 10 and m.nlevels.max() <= nl
    et = m.edge_tri
    assert np.all(et[:m.edge2D_in, 1] >= 0) and np.all(et[m.edge2D_in:, 1] < 0)
    # el1 to the left of n1 -> n2 (rotated plane, cyclic-trimmed)
    c = m.coord_nod2D
    tr = lambda x: np.where(x > np.pi, x - 2 * np.pi, np.where(x < -np.pi, x + 2 * np.pi, x))
    dx = tr(c[m.edges[:, 1], 0] - c[m.edges[:, 0], 0]); dy = c[m.edges[:, 1], 1] - c[m.edges[:, 0], 1]
    ex = tr(m.elem_center_x[et[:, 0]] - c[m.edges[:, 0], 0]); ey = m.elem_center_y[et[:, 0]] - c[m.edges[:, 0], 1]
    assert np.all(dx * ey - dy * ex > 0)
    # every node's level count is the max over its triangles
    nmax = np.zeros(m.nod2D, int)
    np.maximum.at(nmax, m.elem_nodes.reshape(-1), np.repeat(m.nlevels, 3))
    assert np.array_equal(nmax, m.nlevels_nod2D)
    # round trip through the bundle format
    save_mesh(m, tmp_path / "b")
    m2 = load_mesh(tmp_path / "b")
    for k in MESH_ARRAY_FIELDS:
        assert np.array_equal(getattr(m, k), getattr(m2, k)), k
    w = wet_counts(m)
    assert w["sum_Ln"] == int(m.node_layer_mask.sum())
    assert w["sum_Le"] == int(m.elem_layer_mask.sum())


def test_compute_w_matches_oracle_and_closes(pi_mesh, pi_case):
    from fesom_jax_amd.inputs import compute_w
    from oracle import fct_numpy as O
    w = compute_w(pi_mesh, pi_case["uv"], pi_case["helem"])
    assert np.array_equal(w, O.compute_w(pi_mesh, pi_case["uv"], pi_case["helem"]))
    m = pi_mesh
    # bottom interface carries no flux; the antisymmetric edge divergence sums to zero
    idx = np.arange(m.nod2D)
    assert np.all(w[idx, m.nlevels_nod2D - 1] == 0.0)
    total = np.sum(w[:, 0] * m.area[:, 0])
    assert abs(total) <= 1e-9 * np.sum(np.abs(w[:, 0] * m.area[:, 0]))
